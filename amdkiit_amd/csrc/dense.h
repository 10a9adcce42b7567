// Density-sphere (dense) transforms, exchange-correlation and the pointwise
// pieces of the local Kohn-Sham potential, all in FP64.
//
// Real fields a(r), b(r) travel packed as a + i b through one complex FFT;
// the sphere side holds their half-sphere coefficients A(G), B(G) as the two
// "bands" of a 2 x ldc complex array, so rows_inv/cols_inv of fft_passes.h
// serve the G->r direction unchanged.
#pragma once
#include "wave.h"

namespace amdk {

// z-FFT exp(-i) of full columns of a(r) + i b(r), keep planes |m3| <= R.
template <int TX>
__global__ void __launch_bounds__(256) k_z_fwd_dense(PassGeom g, const double* __restrict__ a,
                                                     const double* __restrict__ b, double2* __restrict__ W2,
                                                     FftPlan1D p3, const double2* __restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  double2* tw = reinterpret_cast<double2*>(smem_raw);
  double2* b0 = tw + g.n3;
  double2* b1 = b0 + g.n3 * TX;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const int x0 = blockIdx.x * TX, y = blockIdx.y;
  load_tw(tw, twg, n3);
  for (int idx = threadIdx.x; idx < n3 * TX; idx += blockDim.x) {
    const int z = idx / TX, tx = idx % TX, x = x0 + tx;
    double2 v = make_double2(0, 0);
    if (x < n1) {
      const size_t r = ((size_t)z * n2 + y) * n1 + x;
      v.x = a[r];
      v.y = b ? b[r] : 0.0;
    }
    b0[idx] = v;
  }
  __syncthreads();
  double2* res = lds_fft<double2, -1>(b0, b1, tw, p3, TX, 1, TX, true);
  for (int idx = threadIdx.x; idx < g.nplane * TX; idx += blockDim.x) {
    const int p = idx / TX, tx = idx % TX, x = x0 + tx;
    if (x < n1) W2[((size_t)p * n2 + y) * g.p1 + x] = res[wrap(p - g.R, n3) * TX + tx];
  }
}

// z-FFT exp(+i) of the planes |m3| <= R, full column out:
// oa = Re, ob = Im (either may be null); ov = scale * Re in type T if non-null.
template <typename T, int TX>
__global__ void __launch_bounds__(256) k_z_inv_dense(PassGeom g, const double2* __restrict__ W2,
                                                     double* __restrict__ oa, double* __restrict__ ob,
                                                     T* __restrict__ ov, double scale, FftPlan1D p3,
                                                     const double2* __restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  double2* tw = reinterpret_cast<double2*>(smem_raw);
  double2* b0 = tw + g.n3;
  double2* b1 = b0 + g.n3 * TX;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const int x0 = blockIdx.x * TX, y = blockIdx.y;
  load_tw(tw, twg, n3);
  for (int i = threadIdx.x; i < n3 * TX; i += blockDim.x) b0[i] = make_double2(0, 0);
  __syncthreads();
  for (int idx = threadIdx.x; idx < g.nplane * TX; idx += blockDim.x) {
    const int p = idx / TX, tx = idx % TX, x = x0 + tx;
    if (x < n1) b0[wrap(p - g.R, n3) * TX + tx] = W2[((size_t)p * n2 + y) * g.p1 + x];
  }
  __syncthreads();
  double2* res = lds_fft<double2, +1>(b0, b1, tw, p3, TX, 1, TX, true);
  for (int idx = threadIdx.x; idx < n3 * TX; idx += blockDim.x) {
    const int z = idx / TX, tx = idx % TX, x = x0 + tx;
    if (x < n1) {
      const size_t r = ((size_t)z * n2 + y) * n1 + x;
      if (oa) oa[r] = res[idx].x;
      if (ob) ob[r] = res[idx].y;
      if (ov) ov[r] = (T)(scale * res[idx].x);
    }
  }
}

// x-FFT exp(-i) of the rows and unpack two real fields:
// A = (F(G) + conj F(-G)) / 2, B = (F(G) - conj F(-G)) / 2i, times `scale`.
template <int RPB>
__global__ void __launch_bounds__(256) k_rows_fwd_unpack(PassGeom g, const double2* __restrict__ W1,
                                                         double2* __restrict__ out, int ldc, double scale,
                                                         FftPlan1D p1, const double2* __restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  double2* tw = reinterpret_cast<double2*>(smem_raw);
  double2* b0 = tw + g.n1;
  double2* b1 = b0 + 2 * RPB * g.n1;
  const int n1 = g.n1;
  const int r0 = blockIdx.x * RPB;
  const int nr = min(RPB, g.nrow - r0);
  load_tw(tw, twg, n1);
  for (int lr = 0; lr < nr; ++lr) {
    const RowInfo ri = g.rows[r0 + lr];
    const double2* id = W1 + (size_t)ri.slot_d * g.p1;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) b0[(2 * lr) * n1 + i] = id[i];
    if (ri.slot_m >= 0) {
      const double2* im = W1 + (size_t)ri.slot_m * g.p1;
      for (int i = threadIdx.x; i < n1; i += blockDim.x) b0[(2 * lr + 1) * n1 + i] = im[i];
    } else {
      for (int i = threadIdx.x; i < n1; i += blockDim.x) b0[(2 * lr + 1) * n1 + i] = make_double2(0, 0);
    }
  }
  __syncthreads();
  double2* res = lds_fft<double2, -1>(b0, b1, tw, p1, 2 * nr, n1, 1, false);
  for (int lr = 0; lr < nr; ++lr) {
    const RowInfo ri = g.rows[r0 + lr];
    const double2* rd = res + (2 * lr) * n1;
    const double2* rm = (ri.slot_m >= 0) ? rd + n1 : rd;
    for (int e = threadIdx.x; e < ri.cnt; e += blockDim.x) {
      const int m1 = ri.m1_0 + e;
      const double2 P = rd[wrap(m1, n1)], M = rm[wrap(-m1, n1)];
      const double2 fa = make_double2(P.x + M.x, P.y + M.y), fb = make_double2(P.x - M.x, P.y - M.y);
      out[ri.start + e] = make_double2(0.5 * scale * fa.x, 0.5 * scale * fb.y);
      out[(size_t)ldc + ri.start + e] = make_double2(0.5 * scale * fa.y, -0.5 * scale * fb.x);
    }
  }
}

__device__ __forceinline__ void gvec_of(const int* mill, int g, const Recip& rc, double& gx, double& gy, double& gz) {
  const double m1 = mill[3 * g], m2 = mill[3 * g + 1], m3 = mill[3 * g + 2];
  gx = m1 * rc.g[0] + m2 * rc.g[3] + m3 * rc.g[6];
  gy = m1 * rc.g[1] + m2 * rc.g[4] + m3 * rc.g[7];
  gz = m1 * rc.g[2] + m2 * rc.g[5] + m3 * rc.g[8];
}

// Hartree + local: vg = 4 pi/G^2 (rho + rho_I) + V_ps (G != 0), vg(0) = 0
// (potential.f90:4-46) and the partial sums of E_H and E_loc
// (total_energy.f90:64-131, pseudopotential.f90:22-34).  With f0 non-null
// (GGA) also the fields for grad rho: pair f0 = (i Gx rho, i Gy rho), pair
// f1 = (i Gz rho, 0), zero on the padding up to ldc.
__global__ void __launch_bounds__(256) k_hartree(const int* __restrict__ mill, int ng, Recip rc,
                                                 const double2* __restrict__ rhog, const double2* __restrict__ sr,
                                                 const double2* __restrict__ sv, double2* __restrict__ vg,
                                                 double* __restrict__ part, int ldc, double2* __restrict__ f0,
                                                 double2* __restrict__ f1) {
  __shared__ double sh[4];
  double eh = 0, el = 0;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    double gx, gy, gz;
    gvec_of(mill, g, rc, gx, gy, gz);
    const double g2 = gx * gx + gy * gy + gz * gz;
    const double2 r = rhog[g], ri = sr[g], v = sv[g];
    if (f0) {
      f0[g] = make_double2(-gx * r.y, gx * r.x);
      f0[ldc + g] = make_double2(-gy * r.y, gy * r.x);
      f1[g] = make_double2(-gz * r.y, gz * r.x);
      f1[ldc + g] = make_double2(0, 0);
    }
    if (g2 > 1e-12) {
      const double f = 4.0 * M_PI / g2;
      const double2 rt = make_double2(r.x + ri.x, r.y + ri.y);
      eh += f * (rt.x * rt.x + rt.y * rt.y);
      el += v.x * r.x + v.y * r.y;
      vg[g] = make_double2(f * rt.x + v.x, f * rt.y + v.y);
    } else {
      el += 0.5 * (v.x * r.x + v.y * r.y);
      vg[g] = make_double2(0, 0);
    }
  }
  if (f0)
    for (int g = ng + blockIdx.x * blockDim.x + threadIdx.x; g < ldc; g += gridDim.x * blockDim.x)
      f0[g] = f0[ldc + g] = f1[g] = f1[ldc + g] = make_double2(0, 0);
  const double a = block_sum(eh, sh);
  const double b = block_sum(el, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = b;
  }
}


// vg += A1 - i Gx B1 - i Gy A2 - i Gz B2  (xc.f90:264-296); A2/B2 may be null (LDA)
__global__ void k_vassemble(const int* __restrict__ mill, int ng, int ldc, Recip rc,
                            const double2* __restrict__ p1, const double2* __restrict__ p2,
                            double2* __restrict__ vg) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    double gx, gy, gz;
    gvec_of(mill, g, rc, gx, gy, gz);
    const double2 A1 = p1[g], B1 = p1[ldc + g];
    double2 v = vg[g];
    v.x += A1.x + gx * B1.y;
    v.y += A1.y - gx * B1.x;
    if (p2) {
      const double2 A2 = p2[g], B2 = p2[ldc + g];
      v.x += gy * A2.y + gz * B2.y;
      v.y += -gy * A2.x - gz * B2.x;
    }
    vg[g] = v;
  }
}

// ---------------------------------------------------------------------------
// exchange-correlation energy density f = rho eps and df/drho, df/dsigma
enum { XC_TETER93 = 0, XC_PZ = 1, XC_PBE = 2 };

__device__ __forceinline__ void xc_teter93(double rho, double& f, double& vr) {
  const double a0 = 0.4581652932831429, a1 = 2.217058676663745, a2 = 0.7405551735357053, a3 = 0.01968227878617998;
  const double b1 = 1.0, b2 = 4.504130959426697, b3 = 1.110667363742916, b4 = 0.02359291751427506;
  const double rs = cbrt(3.0 / (4.0 * M_PI * rho));
  const double n = a0 + rs * (a1 + rs * (a2 + rs * a3));
  const double d = rs * (b1 + rs * (b2 + rs * (b3 + rs * b4)));
  const double dn = a1 + rs * (2 * a2 + rs * 3 * a3);
  const double dd = b1 + rs * (2 * b2 + rs * (3 * b3 + rs * 4 * b4));
  const double eps = -n / d;
  const double deps = -(dn * d - n * dd) / (d * d);
  f = rho * eps;
  vr = eps - rs / 3.0 * deps;
}

__device__ __forceinline__ void xc_pz81(double rho, double& f, double& vr) {
  const double cx = -0.75 * cbrt(3.0 / M_PI);
  const double r13 = cbrt(rho);
  const double ex = cx * r13;
  double ec, vc;
  const double rs = cbrt(3.0 / (4.0 * M_PI * rho));
  if (rs >= 1.0) {
    const double g = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sq = sqrt(rs);
    const double den = 1.0 + b1 * sq + b2 * rs;
    ec = g / den;
    vc = ec * (1.0 + 7.0 / 6.0 * b1 * sq + 4.0 / 3.0 * b2 * rs) / den;
  } else {
    const double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    const double lr = log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    vc = A * lr + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lr + (2.0 * D - C) / 3.0 * rs;
  }
  f = rho * (ex + ec);
  vr = 4.0 / 3.0 * ex + vc;
}

__device__ __forceinline__ void xc_pbe(double rho, double sigma, double& f, double& vr, double& vs) {
  // Constants in closed form (evaluated once, not per point):
  //   CX = -(3/4) (3/pi)^(1/3), CS = 1 / (4 (9 pi^4)^(1/3)) (s^2 = CS sigma rho^-8/3),
  //   CT = pi / (16 (3 pi^2)^(1/3)) (t^2 = CT sigma rho^-7/3), RS0 = (3/(4 pi))^(1/3).
  // One cube root, one square root, two logarithms and one exponential per
  // point; the powers of rho and the divisions share their reciprocals.
  constexpr double CX = -0.7385587663820223, CS = 0.026121172985233605, CT = 0.0634682060977037;
  constexpr double RS0 = 0.6203504908994001;
  // exchange
  const double kappa = 0.804, mu = 0.2195149727645171;
  const double irho = 1.0 / rho;
  const double r13 = cbrt(rho);
  const double ir13 = 1.0 / r13;
  const double ex = CX * rho * r13;  // rho * eps_x^LDA
  const double rm73 = irho * irho * ir13;
  const double rm83 = rm73 * ir13;
  const double s2 = CS * sigma * rm83;
  const double iden = 1.0 / (1.0 + mu * s2 / kappa);
  const double Fx = 1.0 + kappa - kappa * iden;
  const double dF = mu * iden * iden;  // dFx/ds2
  const double fx = ex * Fx;
  const double vrx = (4.0 / 3.0) * ex * irho * Fx - ex * dF * (8.0 / 3.0) * s2 * irho;
  const double vsx = ex * dF * CS * rm83;
  // correlation: PW92 (PW_MOD constants) + H
  const double A = 0.0310907, a1 = 0.21370, B1 = 7.5957, B2 = 3.5876, B3 = 1.6382, B4 = 0.49294;
  const double beta = 0.06672455060314922, gamma = (1.0 - M_LN2) / (M_PI * M_PI);
  const double rs = RS0 * ir13;
  const double srs = sqrt(rs);
  const double Q = 2.0 * A * (B1 * srs + B2 * rs + B3 * rs * srs + B4 * rs * rs);
  const double dQ = 2.0 * A * (0.5 * B1 / srs + B2 + 1.5 * B3 * srs + 2.0 * B4 * rs);
  const double iQ = 1.0 / Q;
  const double lq = log(1.0 + iQ);
  const double ec = -2.0 * A * (1.0 + a1 * rs) * lq;
  const double dec = -2.0 * A * a1 * lq + 2.0 * A * (1.0 + a1 * rs) * dQ * iQ / (Q + 1.0);
  const double t2 = CT * sigma * rm73;
  const double ee = exp(-ec * (1.0 / gamma));
  const double Ac = (beta / gamma) / (ee - 1.0);
  const double dAde = Ac * Ac * ee * (1.0 / beta);
  const double y = Ac * t2;
  const double N = 1.0 + y;
  const double iDn = 1.0 / (1.0 + y + y * y);
  const double R = t2 * N * iDn;
  const double iarg = 1.0 / (1.0 + (beta / gamma) * R);
  const double H = -gamma * log(iarg);
  const double dRdt2 = N * iDn - y * y * (2.0 + y) * iDn * iDn;
  const double dRdA = -t2 * t2 * y * (2.0 + y) * iDn * iDn;
  const double dHdt2 = beta * dRdt2 * iarg;
  const double dHdA = beta * dRdA * iarg;
  const double drs = -rs * irho * (1.0 / 3.0);
  const double fc = rho * (ec + H);
  const double vrc = ec + H + rho * (dec * drs + dHdA * dAde * dec * drs) + dHdt2 * (-7.0 / 3.0) * t2;
  const double vsc = rho * dHdt2 * CT * rm73;
  f = fx + fc;
  vr = vrx + vrc;
  vs = vsx + vsc;
}

// XC on the grid with the reference's low-density smoothing (xc.f90:100-133).
// In: rho, grad components (GGA only).  Out: v1 -> va; v2*grad -> gx,gy,gz in
// place; partial E_xc sums (times Omega/N later) and grid sums of rho.
__global__ void __launch_bounds__(256) k_xc(int kind, size_t nr, const double* __restrict__ rho,
                                            double* __restrict__ va, double* __restrict__ gx,
                                            double* __restrict__ gy, double* __restrict__ gz,
                                            double* __restrict__ part, double* __restrict__ part_rho) {
  __shared__ double sh[4];
  const double gc = 1.0e-8;
  double exc = 0, rsum = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nr; i += (size_t)gridDim.x * blockDim.x) {
    rsum += rho[i];
    const double r = fmax(rho[i], 0.0);
    double v1 = 0, v2 = 0;
    if (r > 0.1 * gc) {
      double smoo = 1.0, dsmoo = 0.0;
      if (r <= 4.0 * gc) {
        const double texp = exp(3.0 * (1.0 - r / gc));
        smoo = 1.0 / (1.0 + texp);
        dsmoo = 3.0 / gc * texp * smoo * smoo;
      }
      double f = 0, vr = 0, vs = 0;
      if (kind == XC_TETER93) xc_teter93(r, f, vr);
      else if (kind == XC_PZ) xc_pz81(r, f, vr);
      else {
        const double s = gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i];
        xc_pbe(r, fmax(s, 1e-60), f, vr, vs);
      }
      exc += smoo * f;
      v1 = dsmoo * f + smoo * vr;
      v2 = smoo * 2.0 * vs;
    }
    va[i] = v1;
    if (kind == XC_PBE) {
      gx[i] *= v2;
      gy[i] *= v2;
      gz[i] *= v2;
    }
  }
  const double e = block_sum(exc, sh);
  const double rs = block_sum(rsum, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = e;
    part_rho[blockIdx.x] = rs;
  }
}

// rho = sum of G partial grids (fixed order) and the block sums of rho; G is
// a compile-time count so that a thread's G loads are all in flight at once
template <int G>
__global__ void __launch_bounds__(256) k_rho_combine(const double* __restrict__ part, size_t stride, size_t n,
                                                     double* __restrict__ rho, double* __restrict__ sums) {
  __shared__ double sh[4];
  double s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double v = part[i];
#pragma unroll
    for (int g = 1; g < G; ++g) v += part[g * stride + i];
    rho[i] = v;
    s += v;
  }
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) sums[blockIdx.x] = r;
}

// sum of a real grid (for the electron count)
__global__ void __launch_bounds__(256) k_grid_sum(const double* __restrict__ x, size_t n, double* __restrict__ part) {
  __shared__ double sh[4];
  double s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += x[i];
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}


// E_kin and the density norm of the bands in one pass over C0:
// part4 = sum_b f_b sum_G |G|^2 |c|^2 over the half sphere (total_energy.f90:38-61),
// part10 = sum_b f_b DOTP(c_b, c_b).
template <typename V>
__global__ void __launch_bounds__(256) k_c0_sums(const V* __restrict__ C, int ldc, int npw, int nb,
                                                 const double* __restrict__ g2pw, const double* __restrict__ occ,
                                                 double* __restrict__ part4, double* __restrict__ part10) {
  // work item e = (band group, G): G fastest, so a wave reads consecutive
  // coefficients of one band; C0_BG band groups keep every thread busy
  constexpr int C0_BG = 8;
  __shared__ double sh[4];
  double s = 0, q = 0;
  const int nbg = (nb + C0_BG - 1) / C0_BG;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)npw * C0_BG;
       e += (size_t)gridDim.x * blockDim.x) {
    const int g = (int)(e % npw), b0 = (int)(e / npw) * nbg, b1 = min(nb, b0 + nbg);
    const double g2 = g2pw[g];
#pragma unroll 4
    for (int b = b0; b < b1; ++b) {
      const V c = C[(size_t)b * ldc + g];
      const double a2 = (double)c.x * c.x + (double)c.y * c.y;
      s += occ[b] * g2 * a2;
      q += occ[b] * (g == 0 ? (double)c.x * c.x : 2.0 * a2);
    }
  }
  const double r4 = block_sum(s, sh);
  const double r10 = block_sum(q, sh);
  if (threadIdx.x == 0) {
    part4[blockIdx.x] = r4;
    part10[blockIdx.x] = r10;
  }
}

// Local/Hartree force POTFOR (nuclear_grad.f90:4-50):
// F_a = 2 Omega sum_{G != 0} Re[i G e^{-iGR_a} (rho_s 4pi/G^2 conj(rho_tot) + V_s conj(rho_e))]
// Block (x: atom group grp[x] of one species, y: every gridDim.y-th
// quadruple of sphere rows, one row per wave): the group's x phase rows in
// LDS, each row's y-z factor per atom in registers, three components per
// atom accumulated per lane and reduced in a fixed order into
// part[(a * gridDim.y + y) * 3 + k] (without the 2 Omega).
__global__ void __launch_bounds__(256) k_potfor(const RowInfo* __restrict__ rows, int nrow, Recip rc,
                                                const int4* __restrict__ grp, const int* __restrict__ order,
                                                const double* __restrict__ vps, const double* __restrict__ rhos,
                                                int ldg, const double2* __restrict__ rhog,
                                                const double2* __restrict__ sr, const double2* __restrict__ tab,
                                                int W, int hmax, double* __restrict__ part) {
  __shared__ double2 ex[PHASE_AG][PHASE_WMAX];
  __shared__ double2 yzs[4][PHASE_AG];  // per wave: this row's y-z factors
  __shared__ double red[4][3 * PHASE_AG];
  const int4 gi = grp[blockIdx.x];
  const int s = gi.z;
  stage_x_phases(tab, W, order, gi, ex);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double f[PHASE_AG][3];
#pragma unroll
  for (int j = 0; j < PHASE_AG; ++j) f[j][0] = f[j][1] = f[j][2] = 0.0;
  for (int r = 4 * blockIdx.y + w; r < nrow; r += 4 * gridDim.y) {
    const RowInfo ri = rows[r];
    if (lane < PHASE_AG) {
      const double2* ty = tab + (size_t)order[gi.x + min(lane, gi.y - 1)] * 3 * W + W + hmax;
      yzs[w][lane] = zmul(ty[ri.m2], ty[W + ri.m3]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double gy0 = ri.m2 * rc.g[3] + ri.m3 * rc.g[6], gy1 = ri.m2 * rc.g[4] + ri.m3 * rc.g[7],
                 gy2 = ri.m2 * rc.g[5] + ri.m3 * rc.g[8];
    for (int e = lane; e < ri.cnt; e += 64) {
      const int g = ri.start + e, m1 = ri.m1_0 + e;
      if (g == 0) continue;  // G = 0
      const double gx = m1 * rc.g[0] + gy0, gy = m1 * rc.g[1] + gy1, gz = m1 * rc.g[2] + gy2;
      const double g2 = gx * gx + gy * gy + gz * gz;
      const double2 re = rhog[g], rt = sr[g];
      const double2 rtc = make_double2(re.x + rt.x, -(re.y + rt.y));  // conj(rho_tot)
      const double fh = 4.0 * M_PI / g2;
      const double rs_ = rhos[(size_t)s * ldg + g], vs_ = vps[(size_t)s * ldg + g];
      const double2 t = make_double2(rs_ * fh * rtc.x + vs_ * re.x, rs_ * fh * rtc.y - vs_ * re.y);
#pragma unroll
      for (int j = 0; j < PHASE_AG; ++j) {
        const double2 et = zmul(zmul(ex[j][m1 + hmax], yzs[w][j]), t);
        // Re(i G et) = -G Im(et)
        f[j][0] += -gx * et.y;
        f[j][1] += -gy * et.y;
        f[j][2] += -gz * et.y;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PHASE_AG; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double v = f[j][k];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[w][3 * j + k] = v;
    }
  __syncthreads();
  if (threadIdx.x < 3 * gi.y) {
    const int j = threadIdx.x / 3, k = threadIdx.x - 3 * j;
    const double v = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
    part[((size_t)order[gi.x + j] * gridDim.y + blockIdx.y) * 3 + k] = v;
  }
}

// F_a,k = scale * sum_y part[a][y][k] (POTFOR, ordered) - 2 sum_{p of a} sum_b
// f_b dfnl_k(p, b) Dnl(p, b) (NLFOR, nuclear_grad.f90:51-89; UPF_NLFOR with
// D = hlsg).  dfnl_k and Dnl are column-major (P, nb), dfnl_k at dfnl + k P nb.
__global__ void __launch_bounds__(256) k_force_final(const double* __restrict__ part, int ns, double scale, int P, int nb,
                                                     const int* __restrict__ pstart, const double* __restrict__ occ,
                                                     const double* __restrict__ dfnl, const double* __restrict__ dnl,
                                                     double* __restrict__ force) {
  __shared__ double sh[4];
  const int a = blockIdx.x;
  const int p0 = P > 0 ? pstart[a] : 0, np = P > 0 ? pstart[a + 1] - p0 : 0;
  for (int k = 0; k < 3; ++k) {
    double s = 0;
    for (int e = threadIdx.x; e < np * nb; e += blockDim.x) {
      const int p = p0 + e % np, b = e / np;
      const size_t o = (size_t)b * P + p;
      s += occ[b] * dfnl[(size_t)k * P * nb + o] * dnl[o];
    }
    const double nlf = block_sum(s, sh);
    if (threadIdx.x == 0) {
      double loc = 0;
      for (int y = 0; y < ns; ++y) loc += part[((size_t)a * ns + y) * 3 + k];
      force[3 * a + k] = scale * loc - 2.0 * nlf;
    }
  }
}

// Minimum-image rule of the ion-ion term (apply_pbc.f90:1-60 with m = 1):
// mode 1 simple cubic (each component folded into (-A/2, A/2]); mode 2 FCC
// with conventional constant A: rotated by 45 degrees about z into the
// tetragonal frame of the lattice, folded into its box and moved by the
// body-centre vector when outside |x| + |y| + sqrt2 |z| < 2 m_a; mode 3 the
// nearest image in fractional coordinates (s -= NINT(s)), as the reference
// does for the other lattices.  a: lattice vectors as columns (a[3 j + i] =
// (a_j)_i), ainv: rows of the inverse.
struct PbcInfo {
  int mode;
  double A;
  double a[9];
  double ainv[9];
};

__device__ __forceinline__ void min_image(const PbcInfo& p, double d[3]) {
  if (p.mode == 1) {
    for (int k = 0; k < 3; ++k) {
      const double x = fmod(d[k], p.A);
      d[k] = x - p.A * trunc(x / (0.5 * p.A));
    }
  } else if (p.mode == 2) {
    const double s2 = sqrt(2.0), ma = p.A * s2 / 4.0;
    double x = (d[0] + d[1]) / s2, y = (-d[0] + d[1]) / s2, z = d[2];
    x = fmod(x, 2.0 * ma);
    y = fmod(y, 2.0 * ma);
    z = fmod(z, 2.0 * ma * s2);
    x -= 2.0 * ma * trunc(x / ma);
    y -= 2.0 * ma * trunc(y / ma);
    z -= 2.0 * ma * s2 * trunc(z / (s2 * ma));
    if (!(fabs(x) + fabs(y) + s2 * fabs(z) < 2.0 * ma)) {
      x -= copysign(ma, x);
      y -= copysign(ma, y);
      z -= copysign(s2 * ma, z);
    }
    d[0] = (x - y) / s2;
    d[1] = (x + y) / s2;
    d[2] = z;
  } else if (p.mode == 3) {
    double s[3];
    for (int j = 0; j < 3; ++j) {
      const double v = p.ainv[3 * j] * d[0] + p.ainv[3 * j + 1] * d[1] + p.ainv[3 * j + 2] * d[2];
      s[j] = v - copysign(floor(fabs(v) + 0.5), v);  // NINT: halves away from zero
    }
    for (int i = 0; i < 3; ++i) d[i] = p.a[i] * s[0] + p.a[3 + i] * s[1] + p.a[6 + i] * s[2];
  }
}

// Short-range ion-ion energy and forces with the minimum image above
// (e_sr, total_energy.f90:133-239, nlp = 0).  One wave per atom i: lane l
// takes j = l, l + 64, ..., and the lanes' sums are combined by a fixed
// butterfly; energy counted for j > i.
__global__ void __launch_bounds__(64) k_esr(const double* __restrict__ tau, int nat, const int* __restrict__ sp,
                                            const double* __restrict__ zv, PbcInfo pbc, double rckj,
                                            double* __restrict__ force, double* __restrict__ epart) {
  const int i = blockIdx.x;
  double e = 0, fx = 0, fy = 0, fz = 0;
  for (int j = threadIdx.x; j < nat; j += 64) {
    if (j == i) continue;
    double d[3] = {tau[3 * i] - tau[3 * j], tau[3 * i + 1] - tau[3 * j + 1], tau[3 * i + 2] - tau[3 * j + 2]};
    min_image(pbc, d);
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r = sqrt(r2);
    const double arg = r / rckj;
    if (arg > 20.0) continue;
    const double z2 = zv[sp[i]] * zv[sp[j]];
    const double add = z2 * erfc(arg) / r;
    if (j > i) e += add;
    const double pre = 2.0 * z2 / sqrt(M_PI) * exp(-arg * arg) / rckj;
    const double rep = (add + pre) / r2;
    fx += rep * d[0];
    fy += rep * d[1];
    fz += rep * d[2];
  }
  for (int o = 32; o > 0; o >>= 1) {
    e += __shfl_xor(e, o, 64);
    fx += __shfl_xor(fx, o, 64);
    fy += __shfl_xor(fy, o, 64);
    fz += __shfl_xor(fz, o, 64);
  }
  if (threadIdx.x == 0) {
    force[3 * i] = fx;
    force[3 * i + 1] = fy;
    force[3 * i + 2] = fz;
    epart[i] = e;
  }
}

}  // namespace amdk
