// Final FP64 reduction of the split-K overlap partials (k_ovl_mfma) of the
// Gamma-point real form
//    A(m,n) = 2 sum_G Re(conj(X_m(G)) Y_n(G)) - Re X_m(0) Re Y_n(0)
// (OVLAP_D, fnl, D_FNL; ordered, so the result is deterministic).
// X and Y are band-major complex arrays (band b at b*ld); G = 0 is element 0
// of the internal order.  A is FP64 column-major (Fortran layout).
#pragma once
#include "fft_core.h"

namespace amdk {

// A = alpha * (2 sum_s part_s - x0 y0) + beta * A, then A(m, n) = 0 where
// occ (if given) is zero for band m or n.  A block sums the split-K partials
// of 64 consecutive outputs with OVR_G interleaved groups of splits (one wave
// each, so that up to 16 loads per output are in flight at once), combined in
// a fixed order (deterministic).  sym: the partials hold only the tile x tile
// tiles on and above the diagonal (X = Y); an element of an off-diagonal
// tile sets both A(m, n) and A(n, m), and the elements below skip.
constexpr int OVR_G = 16;
template <typename V>
__global__ void __launch_bounds__(64 * OVR_G) k_ovl_reduce(const double* __restrict__ part, int nsplit, int M, int N,
                                                            const V* __restrict__ X, int ldx, const V* __restrict__ Y,
                                                            int ldy, double alpha, double* __restrict__ A, int lda,
                                                            int g0, double beta, const double* __restrict__ occ,
                                                            int sym, int tile) {
  __shared__ double sh[OVR_G][64];
  const int l = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const size_t MN = (size_t)M * N;
  const size_t o = (size_t)blockIdx.x * 64 + l;
  const int m = (int)(o % M), n = (int)(o / M);
  const bool live = o < MN && !(sym && m / tile > n / tile);
  double s = 0;
  if (live) {
#pragma unroll 4
    for (int p = grp; p < nsplit; p += OVR_G) s += part[(size_t)p * MN + o];
  }
  sh[grp][l] = s;
  __syncthreads();
  if (grp == 0 && live) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < OVR_G; ++q) v += sh[q][l];
    v *= 2.0;
    if (g0) v -= (double)X[(size_t)m * ldx].x * (double)Y[(size_t)n * ldy].x;
    const bool empty = occ && (occ[m] == 0.0 || occ[n] == 0.0);  // empty bands take no part
    double* a = A + (size_t)n * lda + m;
    *a = empty ? 0.0 : alpha * v + (beta != 0.0 ? beta * *a : 0.0);
    if (sym && m / tile < n / tile) {
      double* at = A + (size_t)m * lda + n;
      *at = empty ? 0.0 : alpha * v + (beta != 0.0 ? beta * *at : 0.0);
    }
  }
}

}  // namespace amdk
