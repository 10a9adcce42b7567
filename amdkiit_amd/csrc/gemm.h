// Band-space GEMMs of the Gamma-point real form, all accumulating in FP64.
//
//  ovl  A(m,n) = 2 sum_G Re(conj(X_m(G)) Y_n(G)) - Re X_m(0) Re Y_n(0)
//       (OVLAP_D, fnl, D_FNL: K = 2*ngpw reals, split over K, ordered FP64
//       reduction of the partial tiles, so the result is deterministic)
//  upd  Y_n(G) = beta Y_n(G) + alpha sum_m X_m(G) Q(m,n)
//       (FNONLOC, the Lagrange update of forces, the TRMM of ortho_gs)
//
// X and Y are band-major complex arrays (band b at b*ld); G = 0 is element 0
// of the internal order.  Q and A are FP64 column-major (Fortran layout).
#pragma once
#include "fft_core.h"

namespace amdk {

constexpr int OVL_TM = 32, OVL_TN = 32, OVL_KT = 16;

template <typename V>
__global__ void __launch_bounds__(256) k_ovl_partial(const V* __restrict__ X, int ldx, int M,
                                                     const V* __restrict__ Y, int ldy, int N, int K,
                                                     int kchunk, double* __restrict__ part) {
  __shared__ double2 xs[OVL_KT][OVL_TM + 1];
  __shared__ double2 ys[OVL_KT][OVL_TN + 1];
  const int m0 = blockIdx.x * OVL_TM, n0 = blockIdx.y * OVL_TN;
  const int split = blockIdx.z;
  const int kb = split * kchunk, ke = min(K, kb + kchunk);
  const int tid = threadIdx.x;
  const int tm = tid % 16, tn = tid / 16;  // 2x2 outputs per thread
  double acc[2][2] = {{0, 0}, {0, 0}};
  for (int k0 = kb; k0 < ke; k0 += OVL_KT) {
    for (int e = tid; e < OVL_KT * OVL_TM; e += 256) {
      const int kk = e % OVL_KT, i = e / OVL_KT;
      const int k = k0 + kk;
      double2 v = make_double2(0, 0);
      if (k < ke && m0 + i < M) { const V x = X[(size_t)(m0 + i) * ldx + k]; v.x = x.x; v.y = x.y; }
      xs[kk][i] = v;
      double2 w = make_double2(0, 0);
      if (k < ke && n0 + i < N) { const V y = Y[(size_t)(n0 + i) * ldy + k]; w.x = y.x; w.y = y.y; }
      ys[kk][i] = w;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < OVL_KT; ++kk) {
      const double2 a0 = xs[kk][tm], a1 = xs[kk][tm + 16];
      const double2 b0 = ys[kk][tn], b1 = ys[kk][tn + 16];
      acc[0][0] += a0.x * b0.x + a0.y * b0.y;
      acc[0][1] += a0.x * b1.x + a0.y * b1.y;
      acc[1][0] += a1.x * b0.x + a1.y * b0.y;
      acc[1][1] += a1.x * b1.x + a1.y * b1.y;
    }
    __syncthreads();
  }
  double* P = part + (size_t)split * M * N;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int m = m0 + tm + 16 * a, n = n0 + tn + 16 * b;
      if (m < M && n < N) P[(size_t)n * M + m] = acc[a][b];
    }
}

// A = alpha * (2 sum_s part_s - x0 y0).  A block sums the split-K partials
// of 64 consecutive outputs with 4 interleaved groups of splits, combined in a
// fixed order (deterministic).
template <typename V>
__global__ void __launch_bounds__(256) k_ovl_reduce(const double* __restrict__ part, int nsplit, int M, int N,
                                                    const V* __restrict__ X, int ldx, const V* __restrict__ Y, int ldy,
                                                    double alpha, double* __restrict__ A, int lda, int g0) {
  __shared__ double sh[4][64];
  const int l = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const size_t MN = (size_t)M * N;
  const size_t o = (size_t)blockIdx.x * 64 + l;
  double s = 0;
  if (o < MN) {
#pragma unroll 4
    for (int p = grp; p < nsplit; p += 4) s += part[(size_t)p * MN + o];
  }
  sh[grp][l] = s;
  __syncthreads();
  if (grp == 0 && o < MN) {
    const int m = (int)(o % M), n = (int)(o / M);
    double v = 2.0 * (((sh[0][l] + sh[1][l]) + sh[2][l]) + sh[3][l]);
    if (g0) v -= (double)X[(size_t)m * ldx].x * (double)Y[(size_t)n * ldy].x;
    A[(size_t)n * lda + m] = alpha * v;
  }
}

constexpr int UPD_KC = 64, UPD_NG = 4, UPD_TN = 16, UPD_MT = 32;

// Y_n = beta*Y_n + alpha * sum_m X_m Q(m, n), complex X/Y, real FP64 Q.
template <typename V>
__global__ void __launch_bounds__(256) k_upd(const V* __restrict__ X, int ldx, int M,
                                             const double* __restrict__ Q, int ldq,
                                             V* __restrict__ Y, int ldy, int N, int K,
                                             double alpha, double beta) {
  __shared__ double2 xs[UPD_MT][UPD_KC];
  __shared__ double qs[UPD_MT][UPD_TN];
  const int k0 = blockIdx.x * UPD_KC, n0 = blockIdx.y * UPD_TN;
  const int tid = threadIdx.x;
  const int kl = tid % UPD_KC, ng = tid / UPD_KC;
  double2 acc[UPD_TN / UPD_NG];
#pragma unroll
  for (int c = 0; c < UPD_TN / UPD_NG; ++c) acc[c] = make_double2(0, 0);
  for (int mb = 0; mb < M; mb += UPD_MT) {
    for (int e = tid; e < UPD_MT * UPD_KC; e += 256) {
      const int kk = e % UPD_KC, i = e / UPD_KC;
      double2 v = make_double2(0, 0);
      if (mb + i < M && k0 + kk < K) { const V x = X[(size_t)(mb + i) * ldx + k0 + kk]; v.x = x.x; v.y = x.y; }
      xs[i][kk] = v;
    }
    for (int e = tid; e < UPD_MT * UPD_TN; e += 256) {
      const int j = e % UPD_TN, i = e / UPD_TN;
      qs[i][j] = (mb + i < M && n0 + j < N) ? Q[(size_t)(n0 + j) * ldq + mb + i] : 0.0;
    }
    __syncthreads();
    const int mm_end = min(UPD_MT, M - mb);
    for (int mm = 0; mm < mm_end; ++mm) {
      const double2 x = xs[mm][kl];
#pragma unroll
      for (int c = 0; c < UPD_TN / UPD_NG; ++c) {
        const double q = qs[mm][ng * (UPD_TN / UPD_NG) + c];
        acc[c].x += x.x * q;
        acc[c].y += x.y * q;
      }
    }
    __syncthreads();
  }
  const int k = k0 + kl;
  if (k < K) {
#pragma unroll
    for (int c = 0; c < UPD_TN / UPD_NG; ++c) {
      const int n = n0 + ng * (UPD_TN / UPD_NG) + c;
      if (n < N) {
        V* y = Y + (size_t)n * ldy + k;
        double yx = 0, yy = 0;
        if (beta != 0.0) { yx = beta * (double)y->x; yy = beta * (double)y->y; }
        V o;
        o.x = yx + alpha * acc[c].x;
        o.y = yy + alpha * acc[c].y;
        *y = o;
      }
    }
  }
}

}  // namespace amdk
