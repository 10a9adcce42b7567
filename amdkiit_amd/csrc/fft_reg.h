// Register-resident pencil FFT passes for lengths N = A*B (A, B <= 16).
//
// One transform of length N is spread over max(A, B) threads: each thread
// first does a length-B DFT in registers over elements t + A*k2, multiplies
// the twiddle W_N^{t*j2}, one LDS exchange transposes, and each thread then
// does a length-A DFT that yields the outputs t + B*j1.  The same identity
// with A and B swapped continues from exactly those registers, so a forward
// transform after a pointwise product (V(r) psi(r)) needs no reordering:
// inverse FFT -> V -> forward FFT costs two LDS exchanges per column.
//
// Compared with the LDS Stockham passes of fft_passes.h (one LDS round trip
// and one barrier per radix stage) this keeps every butterfly in VGPRs and
// touches LDS twice per transform.  The Stockham passes remain the fallback
// for lengths without an (A, B) split here.
#pragma once
#include "fft_passes.h"
#include "twiddle_const.h"

namespace amdk {

template <int L> struct Split { static constexpr int P = 1, Q = L; };
template <> struct Split<6> { static constexpr int P = 2, Q = 3; };
template <> struct Split<8> { static constexpr int P = 2, Q = 4; };
template <> struct Split<9> { static constexpr int P = 3, Q = 3; };
template <> struct Split<10> { static constexpr int P = 2, Q = 5; };
template <> struct Split<12> { static constexpr int P = 3, Q = 4; };
template <> struct Split<15> { static constexpr int P = 3, Q = 5; };
template <> struct Split<16> { static constexpr int P = 4, Q = 4; };

// In-register DFT of x[0..L): X_k = sum_j x_j exp(DIR 2 pi i jk / L).
template <typename V, int L, int DIR> struct RDft {
  __device__ static __forceinline__ void run(V* x) {
    if constexpr (L == 1) {
      return;
    } else if constexpr (L <= 5) {
      Butterfly<V, L, DIR>::run(x);
    } else {
      using T = decltype(x[0].x);
      constexpr int P = Split<L>::P, Q = Split<L>::Q;
      static_assert(P * Q == L && P > 1, "no split for this DFT length");
      V y[L];
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        V t[Q];
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) t[k2] = x[k1 + P * k2];
        RDft<V, Q, DIR>::run(t);
#pragma unroll
        for (int j2 = 0; j2 < Q; ++j2) {
          const int m = (j2 * k1) % L;
          if (m == 0) {
            y[k1 * Q + j2] = t[j2];
          } else {
            V w;
            w.x = (T)Tw<L>::re[m];
            w.y = (T)(DIR < 0 ? Tw<L>::im[m] : -Tw<L>::im[m]);
            y[k1 * Q + j2] = cmul(t[j2], w);
          }
        }
      }
#pragma unroll
      for (int j2 = 0; j2 < Q; ++j2) {
        V u[P];
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) u[k1] = y[k1 * Q + j2];
        RDft<V, P, DIR>::run(u);
#pragma unroll
        for (int j1 = 0; j1 < P; ++j1) x[j2 + Q * j1] = u[j1];
      }
    }
  }
};

template <int A, int B> struct MaxAB { static constexpr int v = A > B ? A : B; };

// One length-N = A*B transform per "lane group" (column c of a tile).
// In:  thread t < A holds x[t + A*k2], k2 < B, in r[0..B).
// Out: thread t < B holds X[t + B*j1], j1 < A, in r[0..A).
// lds: A*B*stride elements; element (k1, j2) of column c at (k1*B + j2)*stride + c*cstride.
// twN[m] = exp(-2 pi i m / N).  Every thread of the block must call this.
template <typename V, int A, int B, int DIR>
__device__ __forceinline__ void col_fft(V* r, int t, V* lds, int stride, int coff, const V* twN) {
  if (t < A) {
    RDft<V, B, DIR>::run(r);
#pragma unroll
    for (int j2 = 0; j2 < B; ++j2) {
      V v = r[j2];
      if (j2 > 0 && t > 0) {
        V w = twN[t * j2];
        if (DIR > 0) w.y = -w.y;
        v = cmul(v, w);
      }
      lds[(t * B + j2) * stride + coff] = v;
    }
  }
  __syncthreads();
  if (t < B) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) r[k1] = lds[(k1 * B + t) * stride + coff];
    RDft<V, A, DIR>::run(r);
  }
  __syncthreads();
}

__device__ __forceinline__ int signed_freq(int p, int n) { return 2 * p <= n ? p : p - n; }

// Element offset, in a pencil of planes or rows spaced `ps` apart, of the
// frequency signed_freq(t + A k, N) relative to the first stored one (-lo),
// or -1 when the frequency lies outside [lo, hi].  Written so that the
// per-element work is adds and compares on `base` = (t - lo) * ps, hoisted by
// the caller, instead of 64-bit index products.
template <int A, int N>
__device__ __forceinline__ int pencil_off(int t, int k, int lo, int hi, int base, int ps) {
  const int z = t + A * k;
  const bool wrap = 2 * z > N;
  const int f = wrap ? z - N : z;
  return (f >= lo && f <= hi) ? base + (A * k) * ps - (wrap ? N * ps : 0) : -1;
}

// Block order of the z passes: MI355X hands consecutive workgroups
// round-robin to its 8 XCDs; physical block p runs logical block
// (p % 8) * per + p / 8, so each XCD walks one contiguous range of logical
// blocks L = cb * ngrp + grp (launch 8 * per blocks; ids past the last
// logical block return).  Returns false for those.
__device__ __forceinline__ bool z_block_of(int ngrp, int per, int ncb, int& cb, int& grp) {
  const int p = blockIdx.x;
  const int L = (p & 7) * per + (p >> 3);
  if (L >= ncb * ngrp) return false;
  cb = L / ngrp;
  grp = L - cb * ngrp;
  return true;
}

// ---------------------------------------------------------------------------
// z pass of H|psi>: planes |m3| <= R -> z-FFT exp(+i) -> * v(r) -> z-FFT exp(-i)
template <typename T, int A, int B, int TX>
__global__ void __launch_bounds__(TX * (A > B ? A : B))
    k_z_vpsi_r(PassGeom g, typename cplx_of<T>::type* __restrict__ W2, const T* __restrict__ vr, int npairs,
               int PG, const typename cplx_of<T>::type* __restrict__ twg, int ngrp, int per) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* lds = tw + N;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int n1 = g.n1, n2 = g.n2;
  // logical block (column block, pair group), group fastest, walked
  // XCD-contiguously (z_block_of): a V(r) tile's pair groups and the
  // neighbouring column blocks, which share the 128-byte lines at their
  // edges, run on one XCD and share its L2
  int cb, grp;
  if (!z_block_of(ngrp, per, (n1 * n2 + TX - 1) / TX, cb, grp)) return;
  // columns of the n1 x n2 plane in x-fastest order, TX consecutive per block
  const int q = cb * TX + c;
  const bool live = q < n1 * n2;
  const int y = live ? q / n1 : 0, x = q - y * n1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  T vv[A];
#pragma unroll
  for (int j1 = 0; j1 < A; ++j1) {
    const int z = t + B * j1;
    vv[j1] = (live && t < B) ? vr[((size_t)z * n2 + y) * n1 + x] : (T)0;
  }
  __syncthreads();
  const int pg0 = grp * PG, pg1 = min(npairs, pg0 + PG);
  const int ps = n2 * g.p1, base = (t + g.R) * ps;  // plane stride; plane kz sits at kz + R
  const bool act = t < A && live;
  for (int pl = pg0; pl < pg1; ++pl) {
    V* col = W2 + (size_t)pl * g.nplane * ps + (size_t)y * g.p1 + x;
    V r[TM];
#pragma unroll
    for (int k = 0; k < TM; ++k) { r[k].x = 0; r[k].y = 0; }
    if (act) {
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        const int o = pencil_off<A, N>(t, k2, -g.R, g.R, base, ps);
        if (o >= 0) r[k2] = col[o];
      }
    }
    col_fft<V, A, B, +1>(r, t, lds, TX, c, tw);
    if (t < B) {
#pragma unroll
      for (int j1 = 0; j1 < A; ++j1) {
        r[j1].x *= vv[j1];
        r[j1].y *= vv[j1];
      }
    }
    col_fft<V, B, A, -1>(r, t, lds, TX, c, tw);
    if (act) {
#pragma unroll
      for (int j1 = 0; j1 < B; ++j1) {
        const int o = pencil_off<A, N>(t, j1, -g.R, g.R, base, ps);
        if (o >= 0) col[o] = r[j1];
      }
    }
  }
}

// z pass of the density: z-FFT exp(+i) of every pair, rho += w1 Re^2 + w2 Im^2
// in FP64 registers (each thread owns fixed z), ordered over pairs.  The
// pairs are split into gridDim.z consecutive groups; group k accumulates
// into the grid rho + k * part_stride (summed afterwards in a fixed order).
template <typename T, int A, int B, int TX>
__global__ void __launch_bounds__(TX * (A > B ? A : B))
    k_z_rho_r(PassGeom g, const typename cplx_of<T>::type* __restrict__ W2, double* __restrict__ rho, int npairs,
              const double* __restrict__ wpair, int accumulate, const typename cplx_of<T>::type* __restrict__ twg,
              size_t part_stride, int ngrp, int per) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* lds = tw + N;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int n1 = g.n1, n2 = g.n2;
  int cb, grp;  // (column block, pair group), XCD-contiguous as in k_z_vpsi_r
  if (!z_block_of(ngrp, per, (n1 * n2 + TX - 1) / TX, cb, grp)) return;
  const int q = cb * TX + c;  // plane column, x fastest (as in k_z_vpsi_r)
  const bool live = q < n1 * n2;
  const int y = live ? q / n1 : 0, x = q - y * n1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  double acc[A];
#pragma unroll
  for (int j1 = 0; j1 < A; ++j1) acc[j1] = 0.0;
  // FP32 with <= 10 values per thread: the next pair's planes are loaded into
  // registers while this one is transformed (wider forms would lose occupancy)
  constexpr bool PF = sizeof(T) == 4 && TM <= 10;
  V rn[B];
  const int ps = n2 * g.p1, base = (t + g.R) * ps;  // plane stride; plane kz sits at kz + R
  const bool act = t < A && live;
  auto load_pair = [&](int pl, V* dst) {
    const V* col = W2 + (size_t)pl * g.nplane * ps + (size_t)y * g.p1 + x;
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      const int o = act ? pencil_off<A, N>(t, k2, -g.R, g.R, base, ps) : -1;
      if (o >= 0) dst[k2] = col[o];
      else { dst[k2].x = 0; dst[k2].y = 0; }
    }
  };
  const int ppg = (npairs + ngrp - 1) / ngrp;
  const int pa = grp * ppg, pb = min(npairs, pa + ppg);
  rho += grp * part_stride;
  if (PF && pa < pb) load_pair(pa, rn);
  for (int pl = pa; pl < pb; ++pl) {
    V r[TM];
#pragma unroll
    for (int k = 0; k < TM; ++k) { r[k].x = 0; r[k].y = 0; }
    if constexpr (PF) {
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) r[k2] = rn[k2];
      if (pl + 1 < pb) load_pair(pl + 1, rn);
    } else {
      load_pair(pl, r);
    }
    col_fft<V, A, B, +1>(r, t, lds, TX, c, tw);
    const double w1 = wpair[2 * pl], w2 = wpair[2 * pl + 1];
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) {
      const double re = (double)r[j1].x, im = (double)r[j1].y;
      acc[j1] += w1 * re * re + w2 * im * im;
    }
  }
  if (t < B && live) {
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) {
      const int z = t + B * j1;
      double* dst = rho + ((size_t)z * n2 + y) * n1 + x;
      *dst = accumulate ? (*dst + acc[j1]) : acc[j1];
    }
  }
}

// y pass inverse: rows of plane m3 (consecutive m2 from m2_lo) -> y-FFT exp(+i)
template <typename T, int A, int B, int TX>
__global__ void __launch_bounds__(TX * (A > B ? A : B))
    k_cols_inv_r(PassGeom g, const typename cplx_of<T>::type* __restrict__ W1, typename cplx_of<T>::type* __restrict__ W2,
                 const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* lds = tw + N;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int n1 = g.n1, n2 = g.n2;
  const int x = blockIdx.x * TX + c;
  const int plane = blockIdx.y, pl = blockIdx.z;
  const bool live = x < n1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  const int s0 = g.plane_off[plane], s1 = g.plane_off[plane + 1];
  const int lo = s1 > s0 ? g.slot_m2[s0] : 0, hi = lo + (s1 - s0) - 1;
  __syncthreads();
  V r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) { r[k].x = 0; r[k].y = 0; }
  if (t < A && live) {
    const V* rows = W1 + ((size_t)pl * g.nslot + s0) * g.p1 + x;  // row of m2 = lo
    const int base = (t - lo) * g.p1;
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      const int o = pencil_off<A, N>(t, k2, lo, hi, base, g.p1);
      if (o >= 0) r[k2] = rows[o];
    }
  }
  col_fft<V, A, B, +1>(r, t, lds, TX, c, tw);
  if (t < B && live) {
    V* out = W2 + ((size_t)pl * g.nplane + plane) * (size_t)n2 * g.p1 + (size_t)t * g.p1 + x;
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) out[(B * j1) * g.p1] = r[j1];
  }
}

// y pass forward: full plane column -> y-FFT exp(-i) -> rows inside the sphere
template <typename T, int A, int B, int TX>
__global__ void __launch_bounds__(TX * (A > B ? A : B))
    k_cols_fwd_r(PassGeom g, const typename cplx_of<T>::type* __restrict__ W2, typename cplx_of<T>::type* __restrict__ W1,
                 const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* lds = tw + N;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int n1 = g.n1, n2 = g.n2;
  const int x = blockIdx.x * TX + c;
  const int plane = blockIdx.y, pl = blockIdx.z;
  const bool live = x < n1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  const int s0 = g.plane_off[plane], s1 = g.plane_off[plane + 1];
  const int lo = s1 > s0 ? g.slot_m2[s0] : 0, hi = lo + (s1 - s0) - 1;
  __syncthreads();
  V r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) { r[k].x = 0; r[k].y = 0; }
  const V* in = W2 + ((size_t)pl * g.nplane + plane) * (size_t)n2 * g.p1 + (size_t)t * g.p1 + x;
  if (t < A && live) {
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) r[k2] = in[(A * k2) * g.p1];
  }
  col_fft<V, A, B, -1>(r, t, lds, TX, c, tw);
  if (t < B && live) {
    V* rows = W1 + ((size_t)pl * g.nslot + s0) * g.p1 + x;  // row of m2 = lo
    const int base = (t - lo) * g.p1;
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) {
      const int o = pencil_off<B, N>(t, j1, lo, hi, base, g.p1);
      if (o >= 0) rows[o] = r[j1];
    }
  }
}

// x pass inverse: gather half rows of two bands, pack psi1 + i psi2 on the
// direct row and conj(psi1) + i conj(psi2) on the mirror row, x-FFT exp(+i).
// Block: HR half rows -> 2*HR transforms, TM threads each (t fastest).
template <typename T, int A, int B, int HR>
__global__ void __launch_bounds__(2 * HR * (A > B ? A : B))
    k_rows_inv_r(PassGeom g, const double2* __restrict__ C, int ldc, int nb, int pair0,
                 typename cplx_of<T>::type* __restrict__ W1, const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, ST = N + 1;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* lds = tw + N;
  const int t = threadIdx.x % TM, fr = threadIdx.x / TM;  // transform index in block
  const int hr = fr >> 1, mirror = fr & 1;
  const int row = blockIdx.x * HR + hr;
  const int pl = blockIdx.y, pair = pair0 + pl;
  const int ba = 2 * pair, bb = 2 * pair + 1;
  const bool has2 = bb < nb;
  const bool live = row < g.nrow;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  RowInfo ri{};
  if (live) ri = g.rows[row];
  const bool zero_row = live && ri.slot_m < 0;  // (0,0): both halves on the direct row
  __syncthreads();
  V r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) { r[k].x = 0; r[k].y = 0; }
  if (live && t < A && !(zero_row && mirror)) {
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      const int m1 = signed_freq(t + A * k2, N);
      const bool direct_ok = !mirror && m1 >= ri.m1_0 && m1 < ri.m1_0 + ri.cnt;
      const int mm = -m1;
      const bool mirror_ok = (mirror || (zero_row && m1 < 0)) && mm >= ri.m1_0 && mm < ri.m1_0 + ri.cnt;
      if (direct_ok || mirror_ok) {
        const int e = (direct_ok ? m1 : mm) - ri.m1_0;
        const double2 s1 = C[(size_t)ba * ldc + ri.start + e];
        double2 s2 = make_double2(0.0, 0.0);
        if (has2) s2 = C[(size_t)bb * ldc + ri.start + e];
        V v;
        if (direct_ok) {
          v.x = (T)(s1.x - s2.y);
          v.y = (T)(s1.y + s2.x);
        } else {
          v.x = (T)(s1.x + s2.y);
          v.y = (T)(s2.x - s1.y);
        }
        r[k2] = v;
      }
    }
  }
  col_fft<V, A, B, +1>(r, t, lds, 1, fr * ST, tw);
  if (live && t < B && !(zero_row && mirror)) {
    const int slot = mirror ? ri.slot_m : ri.slot_d;
    V* out = W1 + ((size_t)pl * g.nslot + slot) * g.p1;
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) out[t + B * j1] = r[j1];
  }
}

// x pass forward of H|psi>: x-FFT exp(-i) of the direct and mirror rows, then
// unpack the two bands and write C2 = -(f/2)(|G|^2 c + 2 (V psi)(G)) in FP64.
template <typename T, int A, int B, int HR>
__global__ void __launch_bounds__(2 * HR * (A > B ? A : B))
    k_rows_fwd_hpsi_r(PassGeom g, const typename cplx_of<T>::type* __restrict__ W1, const double2* __restrict__ C,
                      double2* __restrict__ C2, int ldc, int nb, int pair0, const double* __restrict__ occ, Recip rc,
                      const typename cplx_of<T>::type* __restrict__ twg) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, ST = N + 1;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  V* lds = tw + N;
  const int t = threadIdx.x % TM, fr = threadIdx.x / TM;
  const int hr = fr >> 1, mirror = fr & 1;
  const int row = blockIdx.x * HR + hr;
  const int pl = blockIdx.y, pair = pair0 + pl;
  const int ba = 2 * pair, bb = 2 * pair + 1;
  const bool has2 = bb < nb;
  const bool live = row < g.nrow;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  RowInfo ri{};
  if (live) ri = g.rows[row];
  const bool zero_row = live && ri.slot_m < 0;
  __syncthreads();
  V r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) { r[k].x = 0; r[k].y = 0; }
  if (live && t < A && !(zero_row && mirror)) {
    const V* in = W1 + ((size_t)pl * g.nslot + (mirror ? ri.slot_m : ri.slot_d)) * g.p1;
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) r[k2] = in[t + A * k2];
  }
  col_fft<V, A, B, -1>(r, t, lds, 1, fr * ST, tw);
  // outputs X[t + B*j1] of every transform -> LDS (same buffer, after the barrier)
  if (t < B) {
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) lds[fr * ST + t + B * j1] = r[j1];
  }
  __syncthreads();
  double fa = occ[ba] * 0.5;
  if (fa == 0.0) fa = 1.0;
  double fb = has2 ? occ[bb] * 0.5 : 1.0;
  if (fb == 0.0) fb = 1.0;
  // unpack: the 2*TM threads of half row hr walk its elements
  if (live) {
    const V* rd = lds + (2 * hr) * ST;
    const V* rm = zero_row ? rd : rd + ST;
    const double gy0 = ri.m2 * rc.g[3] + ri.m3 * rc.g[6];
    const double gy1 = ri.m2 * rc.g[4] + ri.m3 * rc.g[7];
    const double gy2 = ri.m2 * rc.g[5] + ri.m3 * rc.g[8];
    for (int e = t + mirror * TM; e < ri.cnt; e += 2 * TM) {
      const int m1 = ri.m1_0 + e;
      const double gx = m1 * rc.g[0] + gy0, gyv = m1 * rc.g[1] + gy1, gz = m1 * rc.g[2] + gy2;
      const double gg = gx * gx + gyv * gyv + gz * gz;
      const V Av = rd[wrap(m1, N)];
      const V Bv = rm[wrap(-m1, N)];
      const double fpx = (double)Av.x + (double)Bv.x, fpy = (double)Av.y + (double)Bv.y;
      const double fmx = (double)Av.x - (double)Bv.x, fmy = (double)Av.y - (double)Bv.y;
      const size_t ia = (size_t)ba * ldc + ri.start + e;
      const double2 c1 = C[ia];
      C2[ia] = make_double2(-fa * (gg * c1.x + fpx), -fa * (gg * c1.y + fmy));
      if (has2) {
        const size_t ib = (size_t)bb * ldc + ri.start + e;
        const double2 c2 = C[ib];
        C2[ib] = make_double2(-fb * (gg * c2.x + fpy), -fb * (gg * c2.y - fmx));
      }
    }
  }
}

// Up to two field pairs per launch of the dense transforms (blockIdx.z / .y
// selects the pair; work arrays W1/W2 hold one pair after the other).
struct DenseIn {
  const double* a[2];
  const double* b[2];
};
template <typename T>
struct DenseOut {
  double* a[2];
  double* b[2];
  T* v[2];
};


// ---------------------------------------------------------------------------
// Dense (density-sphere) passes in FP64: two real fields travel as a + i b.

// z-FFT exp(-i) of full columns of a(r) + i b(r); keep planes |m3| <= R.
template <int A, int B, int TX>
__global__ void __launch_bounds__(TX * (A > B ? A : B))
    k_z_fwd_dense_r(PassGeom g, DenseIn in, double2* __restrict__ W2, const double2* __restrict__ twg, int nf,
                    int per) {
  constexpr int N = A * B, TM = MaxAB<A, B>::v;
  extern __shared__ unsigned char smem_raw[];
  double2* tw = reinterpret_cast<double2*>(smem_raw);
  double2* lds = tw + N;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int n1 = g.n1, n2 = g.n2;
  int cb, f;  // (column block, field pair f), XCD-contiguous as in k_z_vpsi_r
  if (!z_block_of(nf, per, (n1 * n2 + TX - 1) / TX, cb, f)) return;
  const int q = cb * TX + c;  // plane column, x fastest
  const bool live = q < n1 * n2;
  const int y = live ? q / n1 : 0, x = q - y * n1;
  const double* __restrict__ a = in.a[f];
  const double* __restrict__ b = in.b[f];
  W2 += (size_t)f * g.nplane * g.n2 * g.p1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  double2 r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) r[k] = make_double2(0.0, 0.0);
  if (t < A && live) {
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      const size_t idx = ((size_t)(t + A * k2) * n2 + y) * n1 + x;
      r[k2] = make_double2(a[idx], b ? b[idx] : 0.0);
    }
  }
  col_fft<double2, A, B, -1>(r, t, lds, TX, c, tw);
  if (t < B && live) {
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) {
      const int kz = signed_freq(t + B * j1, N);
      if (kz >= -g.R && kz <= g.R) W2[((size_t)(kz + g.R) * n2 + y) * g.p1 + x] = r[j1];
    }
  }
}

// z-FFT exp(+i) of the planes |m3| <= R to full columns: oa = Re, ob = Im,
// ov = scale * Re in type T (each may be null).
template <typename T, int A, int B, int TX>
__global__ void __launch_bounds__(TX * (A > B ? A : B))
    k_z_inv_dense_r(PassGeom g, const double2* __restrict__ W2, DenseOut<T> out, double scale,
                    const double2* __restrict__ twg, int nf, int per) {
  constexpr int N = A * B, TM = MaxAB<A, B>::v;
  extern __shared__ unsigned char smem_raw[];
  double2* tw = reinterpret_cast<double2*>(smem_raw);
  double2* lds = tw + N;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int n1 = g.n1, n2 = g.n2;
  int cb, f;  // (column block, field pair f), XCD-contiguous as in k_z_vpsi_r
  if (!z_block_of(nf, per, (n1 * n2 + TX - 1) / TX, cb, f)) return;
  const int q = cb * TX + c;  // plane column, x fastest
  const bool live = q < n1 * n2;
  const int y = live ? q / n1 : 0, x = q - y * n1;
  double* __restrict__ oa = out.a[f];
  double* __restrict__ ob = out.b[f];
  T* __restrict__ ov = out.v[f];
  W2 += (size_t)f * g.nplane * g.n2 * g.p1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  double2 r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) r[k] = make_double2(0.0, 0.0);
  if (t < A && live) {
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      const int kz = signed_freq(t + A * k2, N);
      if (kz >= -g.R && kz <= g.R) r[k2] = W2[((size_t)(kz + g.R) * n2 + y) * g.p1 + x];
    }
  }
  col_fft<double2, A, B, +1>(r, t, lds, TX, c, tw);
  if (t < B && live) {
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) {
      const size_t idx = ((size_t)(t + B * j1) * n2 + y) * n1 + x;
      if (oa) oa[idx] = r[j1].x;
      if (ob) ob[idx] = r[j1].y;
      if (ov) ov[idx] = (T)(scale * r[j1].x);
    }
  }
}

// x-FFT exp(-i) of direct and mirror rows, unpack two real fields:
// A = (F(G) + conj F(-G)) / 2, B = (F(G) - conj F(-G)) / 2i, times scale.
template <int A, int B, int HR>
__global__ void __launch_bounds__(2 * HR * (A > B ? A : B))
    k_rows_fwd_unpack_r(PassGeom g, const double2* __restrict__ W1, double2* __restrict__ out, int ldc, double scale,
                        const double2* __restrict__ twg) {
  constexpr int N = A * B, TM = MaxAB<A, B>::v, ST = N + 1;
  extern __shared__ unsigned char smem_raw[];
  double2* tw = reinterpret_cast<double2*>(smem_raw);
  double2* lds = tw + N;
  const int t = threadIdx.x % TM, fr = threadIdx.x / TM;
  const int hr = fr >> 1, mirror = fr & 1;
  const int row = blockIdx.x * HR + hr;
  W1 += (size_t)blockIdx.y * g.nslot * g.p1;  // field pair blockIdx.y: its rows and its two output fields
  out += (size_t)blockIdx.y * 2 * ldc;
  const bool live = row < g.nrow;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
  RowInfo ri{};
  if (live) ri = g.rows[row];
  const bool zero_row = live && ri.slot_m < 0;
  __syncthreads();
  double2 r[TM];
#pragma unroll
  for (int k = 0; k < TM; ++k) r[k] = make_double2(0.0, 0.0);
  if (live && t < A && !(zero_row && mirror)) {
    const double2* in = W1 + (size_t)(mirror ? ri.slot_m : ri.slot_d) * g.p1;
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) r[k2] = in[t + A * k2];
  }
  col_fft<double2, A, B, -1>(r, t, lds, 1, fr * ST, tw);
  if (t < B) {
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) lds[fr * ST + t + B * j1] = r[j1];
  }
  __syncthreads();
  if (live) {
    const double2* rd = lds + (2 * hr) * ST;
    const double2* rm = zero_row ? rd : rd + ST;
    for (int e = t + mirror * TM; e < ri.cnt; e += 2 * TM) {
      const int m1 = ri.m1_0 + e;
      const double2 P = rd[wrap(m1, N)], M = rm[wrap(-m1, N)];
      out[ri.start + e] = make_double2(0.5 * scale * (P.x + M.x), 0.5 * scale * (P.y - M.y));
      out[(size_t)ldc + ri.start + e] = make_double2(0.5 * scale * (P.y + M.y), -0.5 * scale * (P.x - M.x));
    }
  }
}

}  // namespace amdk
