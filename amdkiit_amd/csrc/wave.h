// Wavefunction-side kernels: structure factors and projector tables for a
// geometry, the Gamma-point clean-up, deterministic FP64 reductions, the
// preconditioned descent of PCG, and the real-space/G-space pointwise pieces
// that do not belong to an FFT pass.
#pragma once
#include "fft_core.h"

namespace amdk {

// ---------------------------------------------------------------------------
// block reductions (sh holds one double per wave of the block)
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) r += sh[i];
  __syncthreads();
  return r;  // valid in thread 0
}
__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = -1.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) r = fmax(r, sh[i]);
  __syncthreads();
  return r;
}

constexpr int RED_BLOCKS = 1024;

// Final pass over the slots set in `slots` (bit q), block q for slot q:
// out[q] = sum, or max when bit q of `maxes` is set, of the RED_BLOCKS
// partials of slot q.  Launched with 16 blocks.
__global__ void k_red_slots(const double* __restrict__ part, unsigned slots, unsigned maxes, double* __restrict__ out) {
  __shared__ double sh[4];
  const int q = blockIdx.x;
  if (!((slots >> q) & 1u)) return;  // uniform over the block
  const bool is_max = (maxes >> q) & 1u;
  double v = is_max ? -1.0 : 0.0;
  for (int i = threadIdx.x; i < RED_BLOCKS; i += blockDim.x) {
    const double x = part[(size_t)q * RED_BLOCKS + i];
    v = is_max ? fmax(v, x) : v + x;
  }
  const double r = is_max ? block_max(v, sh) : block_sum(v, sh);
  if (threadIdx.x == 0) out[q] = r;
}

// Band reductions walk (band, 256-wide chunk of G) pairs block by block, so
// the band index comes from one scalar division per chunk.
// sum_b w_b * DOTP(X_b, X_b) (full-sphere norm from the half sphere, G=0 once)
// and max |X_b(G)| over b < nmax: part[0][blk] = sum, part[1][blk] = max.
template <typename V>
__global__ void __launch_bounds__(256) k_norms(const V* __restrict__ X, int ldx, int npw, int nb,
                                               const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[4];
  double s = 0, mx2 = 0;
  const int nch = (npw + 255) / 256, tot = nb * nch;
  for (int c = blockIdx.x; c < tot; c += gridDim.x) {
    const int b = c / nch, g = (c - b * nch) * 256 + threadIdx.x;
    if (g < npw) {
      const V x = X[(size_t)b * ldx + g];
      const double a2 = (double)x.x * x.x + (double)x.y * x.y;
      s += w[b] * (g == 0 ? (double)x.x * x.x : 2.0 * a2);
      mx2 = fmax(mx2, a2);
    }
  }
  const double rs = block_sum(s, sh);
  const double rm = block_max(sqrt(mx2), sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rs;
    part[RED_BLOCKS + blockIdx.x] = rm;
  }
}

// The gradient's three band reductions in one pass over C2 (bands b < nocc,
// weights w_b): part5 = sum_b w_b DOTP(C2_b, C2_b) (CNORM), part6 = max |C2|
// (GEMAX), part7 = sum_b w_b DOTP(P_b, P_b) for P = f_b^-1 K(G)^-1 C2 (PCG).
template <typename V>
__global__ void __launch_bounds__(256) k_grad_norms(const V* __restrict__ X, int ldx, int npw, int nocc,
                                                    const double* __restrict__ vpp, const double* __restrict__ finv,
                                                    const double* __restrict__ w, double* __restrict__ part5,
                                                    double* __restrict__ part6, double* __restrict__ part7) {
  __shared__ double sh[4];
  double s = 0, p = 0, mx2 = 0;
  const int nch = (npw + 255) / 256, tot = nocc * nch;
  for (int c = blockIdx.x; c < tot; c += gridDim.x) {
    const int b = c / nch, g = (c - b * nch) * 256 + threadIdx.x;
    if (g < npw) {
      const V x = X[(size_t)b * ldx + g];
      const double a2 = (double)x.x * x.x + (double)x.y * x.y;
      const double f = finv[b] * vpp[g];
      s += w[b] * (g == 0 ? (double)x.x * x.x : 2.0 * a2);
      p += w[b] * f * f * (g == 0 ? (double)x.x * x.x : 2.0 * a2);
      mx2 = fmax(mx2, a2);
    }
  }
  const double rs = block_sum(s, sh);
  const double rp = block_sum(p, sh);
  const double rm = block_max(sqrt(mx2), sh);
  if (threadIdx.x == 0) {
    part5[blockIdx.x] = rs;
    part6[blockIdx.x] = rm;
    part7[blockIdx.x] = rp;
  }
}

// ---------------------------------------------------------------------------
// structure factors (exp_igrxfactor): per-atom axis phases e^{-i m s_k},
// s_k = g_k . R_a, for m in [-h, h], h = n_k/2.
__global__ void k_phase_tables(const double* __restrict__ tau, int nat, const double* __restrict__ gm,
                               int n1, int n2, int n3, double2* __restrict__ tab) {
  // tab layout: [a][axis][m + hmax], hmax = max(n)/2, width W = 2*hmax+1
  const int hmax = max(n1, max(n2, n3)) / 2, W = 2 * hmax + 1;
  const int tot = nat * 3 * W;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
    const int m = i % W - hmax;
    const int ax = (i / W) % 3;
    const int a = i / (3 * W);
    const double s = gm[3 * ax] * tau[3 * a] + gm[3 * ax + 1] * tau[3 * a + 1] + gm[3 * ax + 2] * tau[3 * a + 2];
    double sn, cs;
    sincos(-(double)m * s, &sn, &cs);
    tab[i] = make_double2(cs, sn);
  }
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 eigr_of(const double2* tab, int a, int W, int hmax, int m1, int m2, int m3) {
  const double2* t = tab + (size_t)a * 3 * W;
  return zmul(zmul(t[m1 + hmax], t[W + m2 + hmax]), t[2 * W + m3 + hmax]);
}

// The structure-factor and local-force kernels walk the atoms in groups of
// at most PHASE_AG atoms of one species (grp[q] = {first index into order[],
// count, species}; order[] lists the atoms species by species), and stage the
// group's x-axis phase rows e^{-i m1 s_1} (W = 2 hmax + 1 entries, grids up
// to 192 per axis) in LDS, rows past the count as zeros; the loops over a
// group run all PHASE_AG slots without branches, slots past the count taking
// the group's last atom against a zero row.
constexpr int PHASE_AG = 8, PHASE_WMAX = 193;

__device__ __forceinline__ void stage_x_phases(const double2* __restrict__ tab, int W, const int* __restrict__ order,
                                               int4 gi, double2 (*ex)[PHASE_WMAX]) {
  for (int e = threadIdx.x; e < PHASE_AG * W; e += blockDim.x) {
    const int j = e / W, m = e - j * W;
    ex[j][m] = j < gi.y ? tab[(size_t)order[gi.x + j] * 3 * W + m] : make_double2(0.0, 0.0);
  }
}

// eigrxvps(G) = sum_s V_s(G) sum_{a of s} e^{-iGR_a}, eigrxrhos(G) likewise
// with rho_s (exp_igr.f90:4-140).  One G per thread; each atom group's sum
// is scaled by V_s(G) and rho_s(G) of its species.  The y and z factors
// come from the table (lanes of one plane-wave row read one entry).
__global__ void __launch_bounds__(256) k_strfac(const int* __restrict__ mill, int ng, const int4* __restrict__ grp,
                                                int ngrp, const int* __restrict__ order,
                                                const double* __restrict__ vps, const double* __restrict__ rhos,
                                                int ldg, const double2* __restrict__ tab, int W, int hmax,
                                                double2* __restrict__ sv, double2* __restrict__ sr) {
  __shared__ double2 ex[PHASE_AG][PHASE_WMAX];
  const int g = blockIdx.x * 256 + threadIdx.x;
  const bool ok = g < ng;
  const int gc = ok ? g : ng - 1;
  const int m1 = mill[3 * gc] + hmax, m2 = mill[3 * gc + 1] + hmax, m3 = mill[3 * gc + 2] + hmax;
  double2 av = make_double2(0.0, 0.0), ar = make_double2(0.0, 0.0);
  for (int q = 0; q < ngrp; ++q) {
    const int4 gi = grp[q];
    __syncthreads();  // the previous group's rows are no longer read
    stage_x_phases(tab, W, order, gi, ex);
    __syncthreads();
    double2 se = make_double2(0.0, 0.0);
    for (int j = 0; j < gi.y; ++j) {
      const double2* ty = tab + (size_t)order[gi.x + j] * 3 * W + W;
      const double2 e = zmul(zmul(ex[j][m1], ty[m2]), ty[W + m3]);
      se.x += e.x;
      se.y += e.y;
    }
    const double v = vps[(size_t)gi.z * ldg + gc], r = rhos[(size_t)gi.z * ldg + gc];
    av.x += v * se.x;
    av.y += v * se.y;
    ar.x += r * se.x;
    ar.y += r * se.y;
  }
  if (ok) {
    sv[g] = av;
    sr[g] = ar;
  }
}

// beta_p(G) = (-i)^l twnl_{s,k}(G) e^{-iGR_a} for projector p = (a, s, k, l)
template <typename V>
__global__ void k_beta(const int* __restrict__ mill, int npw, int ldc, const int4* __restrict__ plist, int np,
                       const double* __restrict__ twnl, int maxproj, const double2* __restrict__ tab, int W,
                       int hmax, V* __restrict__ beta) {
  const int p = blockIdx.y;
  const int4 pi = plist[p];  // a, s, k, l
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ldc; g += gridDim.x * blockDim.x) {
    V o;
    if (g < npw) {
      const double2 e = eigr_of(tab, pi.x, W, hmax, mill[3 * g], mill[3 * g + 1], mill[3 * g + 2]);
      const double t = twnl[((size_t)pi.y * maxproj + pi.z) * ldc + g];
      double2 b = make_double2(t * e.x, t * e.y);
      switch (pi.w & 3) {  // multiply by (-i)^l
        case 1: b = make_double2(b.y, -b.x); break;
        case 2: b = make_double2(-b.x, -b.y); break;
        case 3: b = make_double2(-b.y, b.x); break;
        default: break;
      }
      o.x = b.x;
      o.y = b.y;
    } else {
      o.x = 0;
      o.y = 0;
    }
    beta[(size_t)p * ldc + g] = o;
  }
}

// derivative projectors: out_p(G) = -i G_dir beta_p(G)
template <typename V>
__global__ void k_dbeta(const V* __restrict__ beta, const int* __restrict__ mill, int npw, int ldc, int np,
                        Recip rc, int dir, V* __restrict__ out) {
  const int p = blockIdx.y;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ldc; g += gridDim.x * blockDim.x) {
    V o;
    o.x = 0;
    o.y = 0;
    if (g < npw) {
      const double gk = mill[3 * g] * rc.g[dir] + mill[3 * g + 1] * rc.g[3 + dir] + mill[3 * g + 2] * rc.g[6 + dir];
      const V b = beta[(size_t)p * ldc + g];
      o.x = (double)b.y * gk;
      o.y = -(double)b.x * gk;
    }
    out[(size_t)p * ldc + g] = o;
  }
}


// ---------------------------------------------------------------------------
// PCG (cg.f90:63-145): P = f^-1 K(G)^-1 C2 (its norm from k_grad_norms, P
// itself formed inside k_pcg_update), d = P + gamma d_prev, C0 += step d.
// k_pcg_scalar: st[0] = ||P||^2 of this step (the sum of its RED_BLOCKS
// partials), st[2] = gamma = forget * st[0] / st[1] (0 at istep 1), then
// st[1] = st[0].  istep is 1-based within one optimisation.
__device__ __forceinline__ void pcg_scalar_set(double p2, double* st, int istep) {
  st[0] = p2;
  double gamma = 0.0;
  if (istep > 1) {
    const double forget = istep <= 10 ? 0.60 : (istep <= 20 ? 0.91 : 0.95);
    gamma = forget * st[0] / st[1];
  }
  st[2] = gamma;
  st[1] = st[0];
}
__global__ void k_pcg_scalar(const double* __restrict__ part, double* st, int istep) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < RED_BLOCKS; i += blockDim.x) v += part[i];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) pcg_scalar_set(v, st, istep);
}

// d = P + gamma d ; C0 += step d  (d stored in H; P = f^-1 K^-1 C2 formed on
// the fly, zero in the padding).  Grid: x over plane waves, y over bands.
template <typename V>
__global__ void k_pcg_update(V* __restrict__ C0, V* __restrict__ H, const V* __restrict__ C2, int ldc, int npw,
                             const double* __restrict__ vpp, const double* __restrict__ finv,
                             const double* __restrict__ st, double step) {
  const double gamma = st[2];
  const int b = blockIdx.y;
  const double fb = finv[b];
  const size_t off = (size_t)b * ldc;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ldc; g += gridDim.x * blockDim.x) {
    const size_t i = off + g;
    double px = 0.0, py = 0.0;
    if (g < npw) {
      const double f = fb * vpp[g];
      px = f * (double)C2[i].x;
      py = f * (double)C2[i].y;
    }
    V d;
    d.x = px + gamma * (double)H[i].x;
    d.y = py + gamma * (double)H[i].y;
    H[i] = d;
    V c = C0[i];
    c.x = (double)c.x + step * (double)d.x;
    c.y = (double)c.y + step * (double)d.y;
    C0[i] = c;
  }
}

}  // namespace amdk
