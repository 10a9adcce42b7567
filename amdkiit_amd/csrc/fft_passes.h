// The five passes of the sphere-sparse, band-pair-batched 3D FFT and the
// kernels that fuse the Kohn-Sham local operator and the density into them.
//
//  rows_inv   sphere half-rows (m2,m3) of two real bands -> complex rows
//             psi1 + i psi2, x-FFT (G->r)                    -> W1 [pair][slot][x]
//  cols_inv   rows of plane m3 -> y-FFT (G->r)               -> W2 [pair][plane][y][x]
//  z_vpsi     z-FFT (G->r), * V(r)/N, z-FFT (r->G), truncate -> W2 (in place)
//  z_rho      z-FFT (G->r), rho += f1 Re^2 + f2 Im^2 over all pairs -> rho(r)
//  cols_fwd   y-FFT (r->G), keep the rows inside the sphere  -> W1
//  rows_fwd   x-FFT (r->G), extract the sphere, undo the two-band packing,
//             add the kinetic term: C2 = -(f/2)(|G|^2 c + 2 (V psi)(G))
//
// Only the planes |m3| <= R are ever materialised, so the z pass reads and
// writes (2R+1)/n3 of a full grid per band pair and the full grid exists only
// inside LDS.
#pragma once
#include "fft_core.h"

namespace amdk {

struct PassGeom {
  int n1, n2, n3;
  int p1;  // row pitch (elements) of the W1/W2 work arrays, >= n1, 128-byte multiple
  int nrow, nslot, nplane, R;
  const RowInfo* rows;
  const int* slot_m2;
  const int* plane_off;
};

struct Recip {  // 2*pi*B, columns g1 g2 g3 (bohr^-1)
  double g[9];
};

__device__ __forceinline__ int wrap(int m, int n) { return m < 0 ? m + n : m; }

template <typename V>
__device__ __forceinline__ void load_tw(V* tw, const V* __restrict__ twg, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
}

// ---------------------------------------------------------------------------
// rows_inv: gather the half-sphere of bands (2*pair, 2*pair+1), pack them as
// psi1 + i psi2 on the direct row and conj(psi1) + i conj(psi2) on the mirror
// row, then x-FFT with exp(+i).
template <typename T, int RPB>
__global__ void __launch_bounds__(256) k_rows_inv(PassGeom g, const double2* __restrict__ C,
                                                  int ldc, int nb, int pair0,
                                                  typename cplx_of<T>::type* __restrict__ W1, FftPlan1D p1,
                                                  const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* b0 = tw + g.n1;
  V* b1 = b0 + 2 * RPB * g.n1;
  const int n1 = g.n1;
  const int pl = blockIdx.y;
  const int pair = pair0 + pl;
  const int ba = 2 * pair, bb = 2 * pair + 1;
  const bool has2 = bb < nb;
  const int r0 = blockIdx.x * RPB;
  const int nr = min(RPB, g.nrow - r0);
  load_tw(tw, twg, n1);
  for (int i = threadIdx.x; i < 2 * RPB * n1; i += blockDim.x) { b0[i].x = 0; b0[i].y = 0; }
  __syncthreads();
  for (int lr = 0; lr < nr; ++lr) {
    const RowInfo ri = g.rows[r0 + lr];
    V* rd = b0 + (2 * lr) * n1;
    V* rm = (ri.slot_m >= 0) ? rd + n1 : rd;
    for (int e = threadIdx.x; e < ri.cnt; e += blockDim.x) {
      const double2 s1 = C[(size_t)ba * ldc + ri.start + e];
      V c1, c2;
      c1.x = (T)s1.x;
      c1.y = (T)s1.y;
      if (has2) {
        const double2 s2 = C[(size_t)bb * ldc + ri.start + e];
        c2.x = (T)s2.x;
        c2.y = (T)s2.y;
      } else {
        c2.x = 0;
        c2.y = 0;
      }
      const int m1 = ri.m1_0 + e;
      V d, m;
      d.x = c1.x - c2.y;
      d.y = c1.y + c2.x;
      m.x = c1.x + c2.y;
      m.y = c2.x - c1.y;
      rd[wrap(m1, n1)] = d;
      if (ri.slot_m >= 0 || m1 != 0) rm[wrap(-m1, n1)] = m;
    }
  }
  __syncthreads();
  V* res = lds_fft<V, +1>(b0, b1, tw, p1, 2 * nr, n1, 1, false);
  for (int lr = 0; lr < nr; ++lr) {
    const RowInfo ri = g.rows[r0 + lr];
    V* od = W1 + ((size_t)pl * g.nslot + ri.slot_d) * g.p1;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) od[i] = res[(2 * lr) * n1 + i];
    if (ri.slot_m >= 0) {
      V* om = W1 + ((size_t)pl * g.nslot + ri.slot_m) * g.p1;
      for (int i = threadIdx.x; i < n1; i += blockDim.x) om[i] = res[(2 * lr + 1) * n1 + i];
    }
  }
}

// ---------------------------------------------------------------------------
// cols_inv: plane m3 = plane - R; rows (slots) of that plane -> y-FFT exp(+i).
template <typename T, int TX>
__global__ void __launch_bounds__(256) k_cols_inv(PassGeom g, const typename cplx_of<T>::type* __restrict__ W1,
                                                  typename cplx_of<T>::type* __restrict__ W2, FftPlan1D p2,
                                                  const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* b0 = tw + g.n2;
  V* b1 = b0 + g.n2 * TX;
  const int n1 = g.n1, n2 = g.n2;
  const int x0 = blockIdx.x * TX;
  const int plane = blockIdx.y;
  const int pl = blockIdx.z;
  load_tw(tw, twg, n2);
  for (int i = threadIdx.x; i < n2 * TX; i += blockDim.x) { b0[i].x = 0; b0[i].y = 0; }
  __syncthreads();
  const int s0 = g.plane_off[plane], s1 = g.plane_off[plane + 1];
  for (int idx = threadIdx.x; idx < (s1 - s0) * TX; idx += blockDim.x) {
    const int s = s0 + idx / TX;
    const int tx = idx % TX;
    const int x = x0 + tx;
    if (x < n1) b0[wrap(g.slot_m2[s], n2) * TX + tx] = W1[((size_t)pl * g.nslot + s) * g.p1 + x];
  }
  __syncthreads();
  V* res = lds_fft<V, +1>(b0, b1, tw, p2, TX, 1, TX, true);
  V* out = W2 + ((size_t)pl * g.nplane + plane) * (size_t)n2 * g.p1;
  for (int idx = threadIdx.x; idx < n2 * TX; idx += blockDim.x) {
    const int y = idx / TX, tx = idx % TX, x = x0 + tx;
    if (x < n1) out[(size_t)y * g.p1 + x] = res[idx];
  }
}

// ---------------------------------------------------------------------------
// cols_fwd: W2 plane -> y-FFT exp(-i) -> keep rows inside the sphere -> W1.
template <typename T, int TX>
__global__ void __launch_bounds__(256) k_cols_fwd(PassGeom g, const typename cplx_of<T>::type* __restrict__ W2,
                                                  typename cplx_of<T>::type* __restrict__ W1, FftPlan1D p2,
                                                  const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* b0 = tw + g.n2;
  V* b1 = b0 + g.n2 * TX;
  const int n1 = g.n1, n2 = g.n2;
  const int x0 = blockIdx.x * TX;
  const int plane = blockIdx.y;
  const int pl = blockIdx.z;
  load_tw(tw, twg, n2);
  const V* in = W2 + ((size_t)pl * g.nplane + plane) * (size_t)n2 * g.p1;
  for (int idx = threadIdx.x; idx < n2 * TX; idx += blockDim.x) {
    const int y = idx / TX, tx = idx % TX, x = x0 + tx;
    V v;
    if (x < n1) v = in[(size_t)y * g.p1 + x]; else { v.x = 0; v.y = 0; }
    b0[idx] = v;
  }
  __syncthreads();
  V* res = lds_fft<V, -1>(b0, b1, tw, p2, TX, 1, TX, true);
  const int s0 = g.plane_off[plane], s1 = g.plane_off[plane + 1];
  for (int idx = threadIdx.x; idx < (s1 - s0) * TX; idx += blockDim.x) {
    const int s = s0 + idx / TX;
    const int tx = idx % TX;
    const int x = x0 + tx;
    if (x < n1) W1[((size_t)pl * g.nslot + s) * g.p1 + x] = res[wrap(g.slot_m2[s], n2) * TX + tx];
  }
}

// ---------------------------------------------------------------------------
// z_vpsi: for PG pairs sharing one V tile: z-FFT exp(+i) of the 2R+1 planes,
// multiply by v(r) (already divided by N), z-FFT exp(-i), keep 2R+1 planes.
template <typename T, int TX>
__global__ void __launch_bounds__(256) k_z_vpsi(PassGeom g, typename cplx_of<T>::type* __restrict__ W2,
                                                const T* __restrict__ vr, int npairs, int PG, FftPlan1D p3,
                                                const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* b0 = tw + g.n3;
  V* b1 = b0 + g.n3 * TX;
  T* vt = reinterpret_cast<T*>(b1 + g.n3 * TX);
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const int x0 = blockIdx.x * TX;
  const int y = blockIdx.y;
  const int pg0 = blockIdx.z * PG;
  const int pg1 = min(npairs, pg0 + PG);
  load_tw(tw, twg, n3);
  for (int idx = threadIdx.x; idx < n3 * TX; idx += blockDim.x) {
    const int z = idx / TX, tx = idx % TX, x = x0 + tx;
    vt[idx] = (x < n1) ? vr[((size_t)z * n2 + y) * n1 + x] : (T)0;
  }
  for (int pl = pg0; pl < pg1; ++pl) {
    for (int i = threadIdx.x; i < n3 * TX; i += blockDim.x) { b0[i].x = 0; b0[i].y = 0; }
    __syncthreads();
    V* col = W2 + (size_t)pl * g.nplane * n2 * g.p1 + (size_t)y * g.p1;
    for (int idx = threadIdx.x; idx < g.nplane * TX; idx += blockDim.x) {
      const int p = idx / TX, tx = idx % TX, x = x0 + tx;
      if (x < n1) b0[wrap(p - g.R, n3) * TX + tx] = col[(size_t)p * n2 * g.p1 + x];
    }
    __syncthreads();
    V* r1 = lds_fft<V, +1>(b0, b1, tw, p3, TX, 1, TX, true);
    for (int i = threadIdx.x; i < n3 * TX; i += blockDim.x) {
      const T v = vt[i];
      r1[i].x *= v;
      r1[i].y *= v;
    }
    __syncthreads();
    V* other = (r1 == b0) ? b1 : b0;
    V* r2 = lds_fft<V, -1>(r1, other, tw, p3, TX, 1, TX, true);
    for (int idx = threadIdx.x; idx < g.nplane * TX; idx += blockDim.x) {
      const int p = idx / TX, tx = idx % TX, x = x0 + tx;
      if (x < n1) col[(size_t)p * n2 * g.p1 + x] = r2[wrap(p - g.R, n3) * TX + tx];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// z_rho: z-FFT exp(+i) of every pair in the batch and accumulate
// rho(r) += w1 * Re(psi)^2 + w2 * Im(psi)^2 in FP64 registers; one block owns
// one (y, x-tile) column set, so the sum over pairs is ordered and atomic-free.
template <typename T, int TX, int MAXE>
__global__ void __launch_bounds__(256) k_z_rho(PassGeom g, const typename cplx_of<T>::type* __restrict__ W2,
                                               double* __restrict__ rho, int npairs,
                                               const double* __restrict__ wpair, int accumulate,
                                               FftPlan1D p3, const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* b0 = tw + g.n3;
  V* b1 = b0 + g.n3 * TX;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const int x0 = blockIdx.x * TX;
  const int y = blockIdx.y;
  double acc[MAXE];
#pragma unroll
  for (int k = 0; k < MAXE; ++k) acc[k] = 0.0;
  load_tw(tw, twg, n3);
  for (int pl = 0; pl < npairs; ++pl) {
    for (int i = threadIdx.x; i < n3 * TX; i += blockDim.x) { b0[i].x = 0; b0[i].y = 0; }
    __syncthreads();
    const V* col = W2 + (size_t)pl * g.nplane * n2 * g.p1 + (size_t)y * g.p1;
    for (int idx = threadIdx.x; idx < g.nplane * TX; idx += blockDim.x) {
      const int p = idx / TX, tx = idx % TX, x = x0 + tx;
      if (x < n1) b0[wrap(p - g.R, n3) * TX + tx] = col[(size_t)p * n2 * g.p1 + x];
    }
    __syncthreads();
    V* r = lds_fft<V, +1>(b0, b1, tw, p3, TX, 1, TX, true);
    const double w1 = wpair[2 * pl], w2 = wpair[2 * pl + 1];
#pragma unroll
    for (int k = 0; k < MAXE; ++k) {
      const int i = threadIdx.x + k * blockDim.x;
      if (i < n3 * TX) {
        const double re = (double)r[i].x, im = (double)r[i].y;
        acc[k] += w1 * re * re + w2 * im * im;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < MAXE; ++k) {
    const int i = threadIdx.x + k * blockDim.x;
    if (i < n3 * TX) {
      const int z = i / TX, tx = i % TX, x = x0 + tx;
      if (x < n1) {
        double* dst = rho + ((size_t)z * n2 + y) * n1 + x;
        *dst = accumulate ? (*dst + acc[k]) : acc[k];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// rows_fwd_hpsi: x-FFT exp(-i) of the direct and mirror rows, then unpack the
// two bands and write C2 = -(f/2) (|G|^2 c + 2 (V psi)(G)) (f = 0 -> factor 1).
template <typename T, int RPB>
__global__ void __launch_bounds__(256) k_rows_fwd_hpsi(PassGeom g, const typename cplx_of<T>::type* __restrict__ W1,
                                                       const double2* __restrict__ C,
                                                       double2* __restrict__ C2, int ldc,
                                                       int nb, int pair0, const double* __restrict__ occ,
                                                       Recip rc, FftPlan1D p1,
                                                       const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* b0 = tw + g.n1;
  V* b1 = b0 + 2 * RPB * g.n1;
  const int n1 = g.n1;
  const int pl = blockIdx.y;
  const int pair = pair0 + pl;
  const int ba = 2 * pair, bb = 2 * pair + 1;
  const bool has2 = bb < nb;
  const int r0 = blockIdx.x * RPB;
  const int nr = min(RPB, g.nrow - r0);
  load_tw(tw, twg, n1);
  for (int lr = 0; lr < nr; ++lr) {
    const RowInfo ri = g.rows[r0 + lr];
    const V* id = W1 + ((size_t)pl * g.nslot + ri.slot_d) * g.p1;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) b0[(2 * lr) * n1 + i] = id[i];
    if (ri.slot_m >= 0) {
      const V* im = W1 + ((size_t)pl * g.nslot + ri.slot_m) * g.p1;
      for (int i = threadIdx.x; i < n1; i += blockDim.x) b0[(2 * lr + 1) * n1 + i] = im[i];
    } else {
      for (int i = threadIdx.x; i < n1; i += blockDim.x) { b0[(2 * lr + 1) * n1 + i].x = 0; b0[(2 * lr + 1) * n1 + i].y = 0; }
    }
  }
  __syncthreads();
  V* res = lds_fft<V, -1>(b0, b1, tw, p1, 2 * nr, n1, 1, false);
  double fa = occ[ba] * 0.5;
  if (fa == 0.0) fa = 1.0;
  double fb = has2 ? occ[bb] * 0.5 : 1.0;
  if (fb == 0.0) fb = 1.0;

  for (int lr = 0; lr < nr; ++lr) {
    const RowInfo ri = g.rows[r0 + lr];
    const V* rd = res + (2 * lr) * n1;
    const V* rm = (ri.slot_m >= 0) ? rd + n1 : rd;
    const double gy0 = ri.m2 * rc.g[3] + ri.m3 * rc.g[6];
    const double gy1 = ri.m2 * rc.g[4] + ri.m3 * rc.g[7];
    const double gy2 = ri.m2 * rc.g[5] + ri.m3 * rc.g[8];
    for (int e = threadIdx.x; e < ri.cnt; e += blockDim.x) {
      const int m1 = ri.m1_0 + e;
      const double gx = m1 * rc.g[0] + gy0, gyv = m1 * rc.g[1] + gy1, gz = m1 * rc.g[2] + gy2;
      const double gg = gx * gx + gyv * gyv + gz * gz;
      const V A = rd[wrap(m1, n1)];
      const V B = rm[wrap(-m1, n1)];
      const V FP = cadd(A, B), FM = csub(A, B);
      const size_t ia = (size_t)ba * ldc + ri.start + e;
      const double2 c1 = C[ia];
      C2[ia] = make_double2(-fa * (gg * c1.x + (double)FP.x), -fa * (gg * c1.y + (double)FM.y));
      if (has2) {
        const size_t ib = (size_t)bb * ldc + ri.start + e;
        const double2 c2 = C[ib];
        C2[ib] = make_double2(-fb * (gg * c2.x + (double)FP.y), -fb * (gg * c2.y - (double)FM.x));
      }
    }
  }
}

}  // namespace amdk
