// Band-space GEMMs on the matrix cores: FP64 (v_mfma_f64_16x16x4f64) or
// FP32 (v_mfma_f32_16x16x4f32) arithmetic on FP64 band arrays.
//
// Complex band arrays are read as real arrays of 2*ldc doubles per band (the
// Gamma-point real form of OVLAP_D: K = 2*N_pw reals), so every product is a
// real GEMM:
//   ovl:  part_s(m, n) = sum_{k in split s} X_m[k] Y_n[k]     (TN, split over K)
//   upd:  Y_n[k] = beta Y_n[k] + alpha sum_m Q(m, n) X_m[k]    (NN)
// Both 16x16x4 instructions take the same operands (wave64): lane l supplies
// A(l%16, l/16) and B(l/16, l%16).  They spread the 16x16 result differently
// over the four accumulator registers v of lane l: the FP32 form holds
// D(4*(l/16) + v, l%16), the FP64 form D(4*v + l/16, l%16) (one row group per
// register), see acc_row().  The FP32 variant rounds the operands to FP32
// while staging them, accumulates one 32-deep K tile in FP32 and adds every
// tile into FP64 accumulators, so the rounding error stays at FP32 level
// relative to the tile sums whatever the total K.  Accumulation order is
// fixed, so results are deterministic.
#pragma once
#include <utility>

#include "fft_core.h"
#include "wave.h"

namespace amdk {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma16(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma16(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// row of the 16x16 result held by accumulator register v of a lane with l/16 = lk
template <typename TC>
__device__ __forceinline__ int acc_row(int lk, int v) {
  return sizeof(TC) == 8 ? 4 * v + lk : 4 * lk + v;
}

template <typename TC> struct Acc4;
template <> struct Acc4<double> { typedef dbl4 type; };
template <> struct Acc4<float> { typedef flt4 type; };

__device__ __forceinline__ void acc_flush(dbl4& hi, dbl4& lo) {}  // FP64: accumulate in place
__device__ __forceinline__ void acc_flush(dbl4& hi, flt4& lo) {
  hi[0] += (double)lo[0];
  hi[1] += (double)lo[1];
  hi[2] += (double)lo[2];
  hi[3] += (double)lo[3];
  lo = (flt4){0.f, 0.f, 0.f, 0.f};
}

constexpr int OVM_T = 64, OVM_KT = 64;

// MI355X hands consecutive workgroups round-robin to its 8 XCDs, each with
// its own L2.  Physical block p runs logical block xcd_block(p, per) so that
// every XCD works through one contiguous range of `per` logical blocks: the
// tiles that share operand rows are then resident in the same L2.  Launch
// 8 * per blocks; logical ids >= the real count exit at once.
__device__ __forceinline__ int xcd_block(int p, int per) { return (p & 7) * per + (p >> 3); }

// Block tile 64 x 64 of the result, 4 waves of 32 x 32 (2 x 2 MFMA tiles);
// logical block = (split, n-tile, m-tile), m fastest.  The next 64-deep K
// tile is loaded into registers (double2, 512 B per row) while the current
// one is multiplied from LDS.  sym (X = Y, M = N) computes only the tiles
// on and above the diagonal (m-tile <= n-tile, numbered column by column);
// k_ovl_reduce mirrors them.  A diagonal tile's two operands are the same
// rows, loaded and staged once.
template <typename TC>
__global__ void __launch_bounds__(256) k_ovl_mfma(const double* __restrict__ X, int ldx, int M,
                                                  const double* __restrict__ Y, int ldy, int N, int K,
                                                  int kchunk, int nsplit, int per, int sym,
                                                  double* __restrict__ part) {
  typedef typename Acc4<TC>::type AccT;
  constexpr bool F64 = sizeof(TC) == 8;
  constexpr int KH = OVM_KT / 2;                      // double2 per row and tile
  constexpr int LD = (OVM_T * KH) / 256;              // double2 loads per thread and operand
  __shared__ TC xs[OVM_T][OVM_KT + 1];
  __shared__ TC ys[OVM_T][OVM_KT + 1];
  const int tm = (M + OVM_T - 1) / OVM_T, tn = (N + OVM_T - 1) / OVM_T;
  const int tiles = sym ? tn * (tn + 1) / 2 : tm * tn;
  const int L = xcd_block(blockIdx.x, per);
  if (L >= tiles * nsplit) return;
  const int split = L / tiles, t = L % tiles;
  int m0, n0;
  if (sym) {  // t = j (j + 1) / 2 + i, i <= j
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= t) ++j;
    m0 = (t - j * (j + 1) / 2) * OVM_T;
    n0 = j * OVM_T;
  } else {
    m0 = (t % tm) * OVM_T;
    n0 = (t / tm) * OVM_T;
  }
  const int kb = split * kchunk, ke = min(K, kb + kchunk);
  const bool diag = sym && m0 == n0;  // uniform per block
  const TC(*yt)[OVM_KT + 1] = diag ? xs : ys;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  dbl4 acc[2][2];
  AccT tac[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};
      tac[a][b] = (AccT){0, 0, 0, 0};
    }
  double2 px[LD], py[LD];
  auto load = [&](int k0) {
#pragma unroll
    for (int it = 0; it < LD; ++it) {
      const int e = threadIdx.x + 256 * it, r = e / KH, k = k0 + 2 * (e % KH);
      // K and every chunk start are even, so a pair never straddles ke
      px[it] = (m0 + r < M && k < ke) ? *(const double2*)(X + (size_t)(m0 + r) * ldx + k) : make_double2(0.0, 0.0);
      if (!diag)
        py[it] = (n0 + r < N && k < ke) ? *(const double2*)(Y + (size_t)(n0 + r) * ldy + k) : make_double2(0.0, 0.0);
    }
  };
  load(kb);
  for (int k0 = kb; k0 < ke; k0 += OVM_KT) {
#pragma unroll
    for (int it = 0; it < LD; ++it) {
      const int e = threadIdx.x + 256 * it, r = e / KH, c = 2 * (e % KH);
      xs[r][c] = (TC)px[it].x;
      xs[r][c + 1] = (TC)px[it].y;
      if (!diag) {
        ys[r][c] = (TC)py[it].x;
        ys[r][c + 1] = (TC)py[it].y;
      }
    }
    __syncthreads();
    if (k0 + OVM_KT < ke) load(k0 + OVM_KT);
#pragma unroll
    for (int ks = 0; ks < OVM_KT; ks += 4) {
      const int kk = ks + lk;
      const TC a0 = xs[wm + li][kk], a1 = xs[wm + 16 + li][kk];
      const TC b0 = yt[wn + li][kk], b1 = yt[wn + 16 + li][kk];
      if constexpr (F64) {
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
      } else {
        tac[0][0] = mfma16(a0, b0, tac[0][0]);
        tac[0][1] = mfma16(a0, b1, tac[0][1]);
        tac[1][0] = mfma16(a1, b0, tac[1][0]);
        tac[1][1] = mfma16(a1, b1, tac[1][1]);
      }
      if constexpr (!F64) {
        if ((ks & 31) == 28) {  // FP64 sum of every 32-deep FP32 tile
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc_flush(acc[a][b], tac[a][b]);
        }
      }
    }
    __syncthreads();
  }
  double* P = part + (size_t)split * M * N;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + 16 * a + acc_row<TC>(lk, v);
        const int n = n0 + wn + 16 * b + li;
        if (m < M && n < N) P[(size_t)n * M + m] = acc[a][b][v];
      }
}

// The same product with a 128 x 128 block tile: 8 waves in 2 (m) x 4 (n),
// each 64 x 32 (4 x 2 MFMA tiles), 32-deep K tiles staged in LDS with the
// next one prefetched into registers.  Every operand row a block loads feeds
// twice the output of the 64 x 64 tile, which halves the operand traffic
// from L2 per flop.  Otherwise as k_ovl_mfma (split-K partials, XCD-aware
// block order, sym: tiles on and above the diagonal, diagonal tiles load
// their rows once; FP32 tiles flushed into FP64 every 32 K).
constexpr int OVB_T = 128, OVB_KT = 32;
template <typename TC>
__global__ void __launch_bounds__(512) k_ovl_big(const double* __restrict__ X, int ldx, int M,
                                                 const double* __restrict__ Y, int ldy, int N, int K,
                                                 int kchunk, int nsplit, int per, int sym,
                                                 double* __restrict__ part) {
  typedef typename Acc4<TC>::type AccT;
  constexpr bool F64 = sizeof(TC) == 8;
  constexpr int KH = OVB_KT / 2;                       // double2 per row and tile
  constexpr int LD = (OVB_T * KH) / 512;               // double2 loads per thread and operand
  __shared__ TC xs[OVB_T][OVB_KT + 1];
  __shared__ TC ys[OVB_T][OVB_KT + 1];
  const int tm = (M + OVB_T - 1) / OVB_T, tn = (N + OVB_T - 1) / OVB_T;
  const int tiles = sym ? tn * (tn + 1) / 2 : tm * tn;
  const int L = xcd_block(blockIdx.x, per);
  if (L >= tiles * nsplit) return;
  const int split = L / tiles, t = L % tiles;
  int m0, n0;
  if (sym) {  // t = j (j + 1) / 2 + i, i <= j
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= t) ++j;
    m0 = (t - j * (j + 1) / 2) * OVB_T;
    n0 = j * OVB_T;
  } else {
    m0 = (t % tm) * OVB_T;
    n0 = (t / tm) * OVB_T;
  }
  const int kb = split * kchunk, ke = min(K, kb + kchunk);
  const bool diag = sym && m0 == n0;  // uniform per block
  const TC(*yt)[OVB_KT + 1] = diag ? xs : ys;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  dbl4 acc[4][2];
  AccT tac[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};
      tac[a][b] = (AccT){0, 0, 0, 0};
    }
  double2 px[LD], py[LD];
  auto load = [&](int k0) {
#pragma unroll
    for (int it = 0; it < LD; ++it) {
      const int e = threadIdx.x + 512 * it, r = e / KH, k = k0 + 2 * (e % KH);
      // K and every chunk start are even, so a pair never straddles ke
      px[it] = (m0 + r < M && k < ke) ? *(const double2*)(X + (size_t)(m0 + r) * ldx + k) : make_double2(0.0, 0.0);
      if (!diag)
        py[it] = (n0 + r < N && k < ke) ? *(const double2*)(Y + (size_t)(n0 + r) * ldy + k) : make_double2(0.0, 0.0);
    }
  };
  load(kb);
  for (int k0 = kb; k0 < ke; k0 += OVB_KT) {
#pragma unroll
    for (int it = 0; it < LD; ++it) {
      const int e = threadIdx.x + 512 * it, r = e / KH, c = 2 * (e % KH);
      xs[r][c] = (TC)px[it].x;
      xs[r][c + 1] = (TC)px[it].y;
      if (!diag) {
        ys[r][c] = (TC)py[it].x;
        ys[r][c + 1] = (TC)py[it].y;
      }
    }
    __syncthreads();
    if (k0 + OVB_KT < ke) load(k0 + OVB_KT);
#pragma unroll
    for (int ks = 0; ks < OVB_KT; ks += 4) {
      const int kk = ks + lk;
      TC av[4], bv[2];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = xs[wm + 16 * a + li][kk];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = yt[wn + 16 * b + li][kk];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          if constexpr (F64) acc[a][b] = mfma16(av[a], bv[b], acc[a][b]);
          else tac[a][b] = mfma16(av[a], bv[b], tac[a][b]);
        }
    }
    if constexpr (!F64) {  // FP64 sum of every 32-deep FP32 tile
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc_flush(acc[a][b], tac[a][b]);
    }
    __syncthreads();
  }
  double* P = part + (size_t)split * M * N;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + 16 * a + acc_row<TC>(lk, v);
        const int n = n0 + wn + 16 * b + li;
        if (m < M && n < N) P[(size_t)n * M + m] = acc[a][b][v];
      }
}

// BF16 arithmetic with one refinement term (the BF16 recipe of the
// projector and Lagrange-multiplier products): every FP64 operand x is split
// into hi = bf16(x) and lo = bf16(x - hi), and a product is formed as
// hi*hi' + (hi*lo' + lo*hi') on v_mfma_f32_16x16x32_bf16, summed in FP32 over
// one 32-deep K tile and in FP64 across tiles.  The dropped lo*lo' and the
// rounding of lo leave ~2^-16 of each term (FP32-level sums need 2^-24).
// 16x16x32 operands (wave64): lane l holds A(l % 16, 8 (l / 16) + j) and
// B(8 (l / 16) + j, l % 16), j = 0..7; the result is spread as the FP32
// 16x16x4 form's (D(4 (l / 16) + v, l % 16)).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ flt4 mfma_bf16(bf16x8 a, bf16x8 b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void split_bf16(double x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)(float)x;
  lo = (__bf16)(float)(x - (double)(float)hi);
}

// Overlap partials (as k_ovl_big: 128 x 128 tile, 8 waves of 64 x 32,
// split-K, sym) in the BF16 recipe above.  32-deep K tiles; hi and lo planes
// of both operands in LDS with a 40-element row pitch (80 B: the 16-byte
// reads of the 16 rows of an MFMA operand fall on distinct bank groups).
constexpr int OVH_KT = 32, OVH_P = OVH_KT + 8;
__global__ void __launch_bounds__(512) k_ovl_bf16(const double* __restrict__ X, int ldx, int M,
                                                  const double* __restrict__ Y, int ldy, int N, int K,
                                                  int kchunk, int nsplit, int per, int sym,
                                                  double* __restrict__ part) {
  constexpr int KH = OVH_KT / 2;
  constexpr int LD = (OVB_T * KH) / 512;
  __shared__ __attribute__((aligned(16))) __bf16 xh[OVB_T][OVH_P];
  __shared__ __attribute__((aligned(16))) __bf16 xl[OVB_T][OVH_P];
  __shared__ __attribute__((aligned(16))) __bf16 yh[OVB_T][OVH_P];
  __shared__ __attribute__((aligned(16))) __bf16 yl[OVB_T][OVH_P];
  const int tm = (M + OVB_T - 1) / OVB_T, tn = (N + OVB_T - 1) / OVB_T;
  const int tiles = sym ? tn * (tn + 1) / 2 : tm * tn;
  const int L = xcd_block(blockIdx.x, per);
  if (L >= tiles * nsplit) return;
  const int split = L / tiles, t = L % tiles;
  int m0, n0;
  if (sym) {
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= t) ++j;
    m0 = (t - j * (j + 1) / 2) * OVB_T;
    n0 = j * OVB_T;
  } else {
    m0 = (t % tm) * OVB_T;
    n0 = (t / tm) * OVB_T;
  }
  const int kb = split * kchunk, ke = min(K, kb + kchunk);
  const bool diag = sym && m0 == n0;
  const __bf16(*yth)[OVH_P] = diag ? xh : yh;
  const __bf16(*ytl)[OVH_P] = diag ? xl : yl;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  dbl4 acc[4][2];
  flt4 tac[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = (dbl4){0.0, 0.0, 0.0, 0.0};
      tac[a][b] = (flt4){0.f, 0.f, 0.f, 0.f};
    }
  double2 px[LD], py[LD];
  auto load = [&](int k0) {
#pragma unroll
    for (int it = 0; it < LD; ++it) {
      const int e = threadIdx.x + 512 * it, r = e / KH, k = k0 + 2 * (e % KH);
      px[it] = (m0 + r < M && k < ke) ? *(const double2*)(X + (size_t)(m0 + r) * ldx + k) : make_double2(0.0, 0.0);
      if (!diag)
        py[it] = (n0 + r < N && k < ke) ? *(const double2*)(Y + (size_t)(n0 + r) * ldy + k) : make_double2(0.0, 0.0);
    }
  };
  load(kb);
  for (int k0 = kb; k0 < ke; k0 += OVH_KT) {
#pragma unroll
    for (int it = 0; it < LD; ++it) {
      const int e = threadIdx.x + 512 * it, r = e / KH, c = 2 * (e % KH);
      split_bf16(px[it].x, xh[r][c], xl[r][c]);
      split_bf16(px[it].y, xh[r][c + 1], xl[r][c + 1]);
      if (!diag) {
        split_bf16(py[it].x, yh[r][c], yl[r][c]);
        split_bf16(py[it].y, yh[r][c + 1], yl[r][c + 1]);
      }
    }
    __syncthreads();
    if (k0 + OVH_KT < ke) load(k0 + OVH_KT);
    const int kk = 8 * lk;
    bf16x8 ah[4], al[4], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      ah[a] = *(const bf16x8*)&xh[wm + 16 * a + li][kk];
      al[a] = *(const bf16x8*)&xl[wm + 16 * a + li][kk];
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      bh[b] = *(const bf16x8*)&yth[wn + 16 * b + li][kk];
      bl[b] = *(const bf16x8*)&ytl[wn + 16 * b + li][kk];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        tac[a][b] = mfma_bf16(ah[a], bl[b], tac[a][b]);
        tac[a][b] = mfma_bf16(al[a], bh[b], tac[a][b]);
        tac[a][b] = mfma_bf16(ah[a], bh[b], tac[a][b]);
      }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc_flush(acc[a][b], tac[a][b]);
    __syncthreads();
  }
  double* P = part + (size_t)split * M * N;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + 16 * a + 4 * lk + v;
        const int n = n0 + wn + 16 * b + li;
        if (m < M && n < N) P[(size_t)n * M + m] = acc[a][b][v];
      }
}

constexpr int UPM_MC = 32, UPM_NB = 128, UPM_K = 64;

// Band norms of the updated Y, reduced by the update kernel that writes it
// (what k_grad_norms computes in a separate pass over C2): for bands
// n < nocc and complex elements g < npw, with DOTP weights (1 for Re at
// G = 0, 0 for Im at G = 0, 2 otherwise),
//   part[L]              = sum w_n DOTP(Y_n, Y_n)                 (CNORM)
//   part[stride + L]     = max |Y_n(G)|                            (GEMAX)
//   part[2 stride + L]   = sum w_n (finv_n vpp_g)^2 DOTP(Y_n, Y_n) (PCG ||P||^2)
// for logical block L; k_fold_norms_pcg turns these into the step's reduction
// slots in a fixed order.  part == nullptr: nothing is computed.
struct UpdNorms {
  const double* vpp;
  const double* finv;
  const double* w;
  int nocc, npw;
  double* part;
  int stride;
};

// The block's band weights for the norms, staged in LDS before the main
// loop (so that the epilogue reads no global memory in data-dependent
// order): wn_s[i] = w_n, fn_s[i] = finv_n for n = n0 + i < nocc, else 0.
__device__ __forceinline__ void upd_norms_stage(const UpdNorms& nm, int n0, double* wn_s, double* fn_s) {
  if (nm.part == nullptr) return;
  for (int i = threadIdx.x; i < UPM_NB; i += blockDim.x) {
    const int n = n0 + i;
    const bool ok = n < nm.nocc;
    wn_s[i] = ok ? nm.w[n] : 0.0;
    fn_s[i] = ok ? nm.finv[n] : 0.0;
  }
}

// x of the neighbouring lane (l ^ 1): a quad permute (1, 0, 3, 2) of both
// halves; every lane of the wave must be active.
__device__ __forceinline__ double lane_xor1(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0xB1, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0xB1, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Store Y(n, k) = beta Y + acc (0 at the imaginary part of G = 0 when
// zero_g0_im) for the NT tiles of this lane, rows n0 + 16 t + row(lk, v),
// and reduce the block's norm partials when asked (weights staged by
// upd_norms_stage).  Called by every thread of the block (the norm
// reduction synchronises it).
template <int NT, typename Row>
__device__ __forceinline__ void upd_store(const dbl4* acc, Row row, double* __restrict__ Y, int ldy, int N, int K,
                                          int k, int n0, int lk, double beta, int zero_g0_im, const UpdNorms& nm,
                                          const double* wn_s, const double* fn_s, int L) {
  const bool kin = k < K;
  const bool norms = nm.part != nullptr;
  const int g = k >> 1;
  const bool gin = kin && g < nm.npw;
  const double wd = k == 0 ? 1.0 : (k == 1 ? 0.0 : 2.0);
  const double vg = norms ? nm.vpp[gin ? g : 0] : 0.0;
  double s5 = 0.0, s7 = 0.0, m2 = 0.0;
  // two tiles at a time: their old Y first, unconditionally (out-of-range
  // elements read element 0 and are not stored), so that eight loads are in
  // flight together instead of one branch and one wait each
  constexpr int TB = 2;
  static_assert(NT % TB == 0, "tiles in pairs");
#pragma unroll
  for (int t0 = 0; t0 < NT; t0 += TB) {
    double yo[TB][4];
    if (beta != 0.0) {
#pragma unroll
      for (int u = 0; u < TB; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int n = n0 + 16 * (t0 + u) + row(lk, v);
          yo[u][v] = Y[(kin && n < N) ? (size_t)n * ldy + k : 0];
        }
    }
#pragma unroll
    for (int u = 0; u < TB; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int t = t0 + u;
      const int n = n0 + 16 * t + row(lk, v);
      const double val = (zero_g0_im && k == 1) ? 0.0 : (beta != 0.0 ? beta * yo[u][v] : 0.0) + acc[t][v];
      if (kin && n < N) Y[(size_t)n * ldy + k] = val;
      if (norms) {
        const double y2 = (gin && n < nm.nocc) ? val * val : 0.0;
        m2 = fmax(m2, y2 + lane_xor1(y2));  // |Y_n(G)|^2: the other component is in lane l ^ 1
        const double wn = wn_s[n - n0], f = fn_s[n - n0] * vg;
        s5 += wn * wd * y2;
        s7 += wn * f * f * wd * y2;
      }
    }
  }
  if (!norms) return;  // uniform over the block
  __shared__ double red[3][16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s5 += __shfl_down(s5, o, 64);
    s7 += __shfl_down(s7, o, 64);
    m2 = fmax(m2, __shfl_down(m2, o, 64));
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s5;
    red[1][w] = s7;
    red[2][w] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < nw; ++i) {
      a += red[0][i];
      b += red[1][i];
      c = fmax(c, red[2][i]);
    }
    nm.part[L] = a;
    nm.part[nm.stride + L] = sqrt(c);
    nm.part[2 * nm.stride + L] = b;
  }
}

// Fold the n block partials of upd_store into RED_BLOCKS partials per slot
// (out + q * RED_BLOCKS for q = 0 CNORM, 1 GEMAX (max), 2 ||P||^2), each a
// fixed strided subset in a fixed order, in one block of RED_BLOCKS threads;
// then, in the same block, the PCG scalars of k_pcg_scalar from the folded
// ||P||^2 partials: one launch instead of two between the band update and
// the PCG update.
__global__ void __launch_bounds__(RED_BLOCKS) k_fold_norms_pcg(const double* __restrict__ part, int stride, int n,
                                                              double* __restrict__ out, double* st, int istep) {
  __shared__ double sh[RED_BLOCKS / 64];
  const int r = threadIdx.x;  // one output per slot and thread (blockDim.x = RED_BLOCKS)
  double v[3] = {0.0, 0.0, 0.0};
  // slot by slot in the inner loop, so that a thread's loads of one j are in
  // flight together
  for (int j = r; j < n; j += RED_BLOCKS) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double x = part[(size_t)q * stride + j];
      v[q] = q == 1 ? fmax(v[q], x) : v[q] + x;
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) out[(size_t)q * RED_BLOCKS + r] = v[q];
  const double p2 = block_sum(v[2], sh);
  if (threadIdx.x == 0) pcg_scalar_set(p2, st, istep);
}

// Block: 64 reals of k (4 waves x 16) x 128 columns n (8 MFMA tiles per
// wave); logical block = (k-block, n-block), n fastest, so the n-blocks that
// read the same rows of X share an XCD's L2.  Chunks of 32 m: X (32 x 64)
// and Q (32 x 128) are staged in LDS, the next chunk prefetched into
// registers while the current one is multiplied.
// A second operand pair (X2: M2 rows, Q2) may follow the first: the sum runs
// over the stacked rows (part 1 padded with zero rows to a whole 32-row
// chunk), with Q1 scaled by a1 and Q2 by a2, so that two updates of Y share
// one pass over it.  zero_g0_im stores 0 into real row 1 of every column
// (Im c(G = 0) of a Gamma-point band, which must vanish).  TRI declares Q1
// upper triangular (Q1(m, n) = 0 for m > n): the MFMA tiles whose columns
// all lie left of a chunk's first row are skipped (their terms are zero).
template <typename TC, bool TRI>
__global__ void __launch_bounds__(256) k_upd_mfma(const double* __restrict__ X1, int ldx, int M1,
                                                  const double* __restrict__ Q1, int ldq1, double a1,
                                                  const double* __restrict__ X2, int M2,
                                                  const double* __restrict__ Q2, int ldq2, double a2,
                                                  double* __restrict__ Y, int ldy, int N, int K,
                                                  double beta, int per, int zero_g0_im, UpdNorms nm) {
  typedef typename Acc4<TC>::type AccT;
  constexpr bool F64 = sizeof(TC) == 8;
  constexpr int XL = (UPM_MC * UPM_K / 2) / 256;   // double2 of X per thread
  constexpr int QL = (UPM_MC * UPM_NB / 2) / 256;  // double2 of Q per thread
  __shared__ TC qs[UPM_MC][UPM_NB + 1];
  __shared__ TC xs[UPM_MC][UPM_K + 1];
  const int nnb = (N + UPM_NB - 1) / UPM_NB, nkb = (K + UPM_K - 1) / UPM_K;
  const int L = xcd_block(blockIdx.x, per);
  if (L >= nnb * nkb) return;
  const int k0 = (L / nnb) * UPM_K, n0 = (L % nnb) * UPM_NB;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int kl = wave * 16 + li, k = k0 + kl;
  __shared__ double wn_s[UPM_NB], fn_s[UPM_NB];
  upd_norms_stage(nm, n0, wn_s, fn_s);  // visible after the main loop's first barrier
  dbl4 acc[8];
  AccT tac[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    acc[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
    tac[t] = (AccT){0, 0, 0, 0};
  }
  // stacked rows: part 1 padded to whole 32-row chunks, then part 2, so each
  // chunk comes from one part (uniform per block)
  const int M1p = (M1 + UPM_MC - 1) / UPM_MC * UPM_MC;
  const int M = M1p + M2;
  double2 px[XL], pq[QL];
  double qa = a1;  // scale of the chunk being staged
  auto load = [&](int mb) {
    const bool two = mb >= M1p;
    const double* X = two ? X2 : X1;
    const double* Q = two ? Q2 : Q1;
    const int ldq = two ? ldq2 : ldq1, Ms = two ? M2 : M1, m0 = two ? mb - M1p : mb;
#pragma unroll
    for (int it = 0; it < XL; ++it) {  // X rows m0.., 32 double2 per row
      const int e = threadIdx.x + 256 * it, r = e / (UPM_K / 2), kk = k0 + 2 * (e % (UPM_K / 2));
      px[it] = (m0 + r < Ms && kk < K) ? *(const double2*)(X + (size_t)(m0 + r) * ldx + kk) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int it = 0; it < QL; ++it) {  // Q columns n0.., 16 double2 (m) per column
      const int e = threadIdx.x + 256 * it, j = e / (UPM_MC / 2), i = 2 * (e % (UPM_MC / 2));
      const bool ok = n0 + j < N;
      const double* q = Q + (size_t)(n0 + j) * ldq + m0 + i;
      pq[it] = make_double2(ok && m0 + i < Ms ? q[0] : 0.0, ok && m0 + i + 1 < Ms ? q[1] : 0.0);
    }
  };
  load(0);
  for (int mb = 0; mb < M; mb += UPM_MC) {
#pragma unroll
    for (int it = 0; it < XL; ++it) {
      const int e = threadIdx.x + 256 * it, r = e / (UPM_K / 2), c = 2 * (e % (UPM_K / 2));
      xs[r][c] = (TC)px[it].x;
      xs[r][c + 1] = (TC)px[it].y;
    }
    qa = mb >= M1p ? a2 : a1;
#pragma unroll
    for (int it = 0; it < QL; ++it) {
      const int e = threadIdx.x + 256 * it, j = e / (UPM_MC / 2), i = 2 * (e % (UPM_MC / 2));
      qs[i][j] = (TC)(qa * pq[it].x);
      qs[i + 1][j] = (TC)(qa * pq[it].y);
    }
    __syncthreads();
    if (mb + UPM_MC < M) load(mb + UPM_MC);
    const int mend = min(UPM_MC, M - mb);
    // tiles t < tskip: columns n0 + 16 t + 15 < mb, Q1 = 0 on this chunk
    const int tskip = (TRI && mb < M1p) ? min(8, max(0, (mb - n0) / 16)) : 0;
    for (int ms = 0; ms < mend; ms += 4) {
      const TC b = xs[ms + lk][kl];
      if constexpr (F64) {
#pragma unroll
        for (int t = 0; t < 8; ++t)
          if (!TRI || t >= tskip) acc[t] = mfma16(qs[ms + lk][16 * t + li], b, acc[t]);
      } else {
#pragma unroll
        for (int t = 0; t < 8; ++t)
          if (!TRI || t >= tskip) tac[t] = mfma16(qs[ms + lk][16 * t + li], b, tac[t]);
      }
    }
    if constexpr (!F64) {
#pragma unroll
      for (int t = 0; t < 8; ++t) acc_flush(acc[t], tac[t]);
    }
    __syncthreads();
  }
  if constexpr (TRI) {  // C U^-1 (no norms): the plain store
    if (k < K) {
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int n = n0 + 16 * t + acc_row<TC>(lk, v);
          if (n < N) {
            double* y = Y + (size_t)n * ldy + k;
            *y = (zero_g0_im && k == 1) ? 0.0 : (beta != 0.0 ? beta * *y : 0.0) + acc[t][v];
          }
        }
    }
  } else {
    upd_store<8>(acc, [](int lk, int v) { return acc_row<TC>(lk, v); }, Y, ldy, N, K, k, n0, lk, beta, zero_g0_im, nm,
                 wn_s, fn_s, L);
  }
}

// k_upd_mfma in the BF16 recipe (hi*hi' + hi*lo' + lo*hi' on
// v_mfma_f32_16x16x32_bf16, one 32-row chunk of m per instruction, FP64
// sums across chunks).  The chunk's X (32 m x 64 k) and Q (32 m x 128 n) are
// staged transposed, m fastest, as hi and lo planes, so that each lane reads
// its 8 consecutive m of an operand in one 16-byte load.
constexpr int UPH_P = UPM_MC + 8;
template <bool TRI>
__global__ void __launch_bounds__(256) k_upd_bf16(const double* __restrict__ X1, int ldx, int M1,
                                                  const double* __restrict__ Q1, int ldq1, double a1,
                                                  const double* __restrict__ X2, int M2,
                                                  const double* __restrict__ Q2, int ldq2, double a2,
                                                  double* __restrict__ Y, int ldy, int N, int K,
                                                  double beta, int per, int zero_g0_im, UpdNorms nm) {
  constexpr int XL = (UPM_MC * UPM_K / 2) / 256;
  constexpr int QL = (UPM_MC * UPM_NB / 2) / 256;
  __shared__ __attribute__((aligned(16))) __bf16 qh[UPM_NB][UPH_P];
  __shared__ __attribute__((aligned(16))) __bf16 ql[UPM_NB][UPH_P];
  __shared__ __attribute__((aligned(16))) __bf16 xh[UPM_K][UPH_P];
  __shared__ __attribute__((aligned(16))) __bf16 xl[UPM_K][UPH_P];
  const int nnb = (N + UPM_NB - 1) / UPM_NB, nkb = (K + UPM_K - 1) / UPM_K;
  const int L = xcd_block(blockIdx.x, per);
  if (L >= nnb * nkb) return;
  const int k0 = (L / nnb) * UPM_K, n0 = (L % nnb) * UPM_NB;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int kl = wave * 16 + li, k = k0 + kl;
  __shared__ double wn_s[UPM_NB], fn_s[UPM_NB];
  upd_norms_stage(nm, n0, wn_s, fn_s);  // visible after the main loop's first barrier
  dbl4 acc[8];
  flt4 tac[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    acc[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
    tac[t] = (flt4){0.f, 0.f, 0.f, 0.f};
  }
  const int M1p = (M1 + UPM_MC - 1) / UPM_MC * UPM_MC;
  const int M = M1p + M2;
  double2 px[XL], pq[QL];
  auto load = [&](int mb) {
    const bool two = mb >= M1p;
    const double* X = two ? X2 : X1;
    const double* Q = two ? Q2 : Q1;
    const int ldq = two ? ldq2 : ldq1, Ms = two ? M2 : M1, m0 = two ? mb - M1p : mb;
#pragma unroll
    for (int it = 0; it < XL; ++it) {
      const int e = threadIdx.x + 256 * it, r = e / (UPM_K / 2), kk = k0 + 2 * (e % (UPM_K / 2));
      px[it] = (m0 + r < Ms && kk < K) ? *(const double2*)(X + (size_t)(m0 + r) * ldx + kk) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int it = 0; it < QL; ++it) {
      const int e = threadIdx.x + 256 * it, j = e / (UPM_MC / 2), i = 2 * (e % (UPM_MC / 2));
      const bool ok = n0 + j < N;
      const double* q = Q + (size_t)(n0 + j) * ldq + m0 + i;
      pq[it] = make_double2(ok && m0 + i < Ms ? q[0] : 0.0, ok && m0 + i + 1 < Ms ? q[1] : 0.0);
    }
  };
  load(0);
  for (int mb = 0; mb < M; mb += UPM_MC) {
#pragma unroll
    for (int it = 0; it < XL; ++it) {
      const int e = threadIdx.x + 256 * it, r = e / (UPM_K / 2), c = 2 * (e % (UPM_K / 2));
      split_bf16(px[it].x, xh[c][r], xl[c][r]);
      split_bf16(px[it].y, xh[c + 1][r], xl[c + 1][r]);
    }
    const double qa = mb >= M1p ? a2 : a1;
#pragma unroll
    for (int it = 0; it < QL; ++it) {
      const int e = threadIdx.x + 256 * it, j = e / (UPM_MC / 2), i = 2 * (e % (UPM_MC / 2));
      split_bf16(qa * pq[it].x, qh[j][i], ql[j][i]);
      split_bf16(qa * pq[it].y, qh[j][i + 1], ql[j][i + 1]);
    }
    __syncthreads();
    if (mb + UPM_MC < M) load(mb + UPM_MC);
    // tiles t < tskip: columns n0 + 16 t + 15 < mb, Q1 = 0 on this chunk
    const int tskip = (TRI && mb < M1p) ? min(8, max(0, (mb - n0) / 16)) : 0;
    const bf16x8 bhv = *(const bf16x8*)&xh[kl][8 * lk];
    const bf16x8 blv = *(const bf16x8*)&xl[kl][8 * lk];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (TRI && t < tskip) continue;
      const bf16x8 ahv = *(const bf16x8*)&qh[16 * t + li][8 * lk];
      const bf16x8 alv = *(const bf16x8*)&ql[16 * t + li][8 * lk];
      tac[t] = mfma_bf16(ahv, blv, tac[t]);
      tac[t] = mfma_bf16(alv, bhv, tac[t]);
      tac[t] = mfma_bf16(ahv, bhv, tac[t]);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) acc_flush(acc[t], tac[t]);
    __syncthreads();
  }
  upd_store<8>(acc, [](int lk, int v) { return 4 * lk + v; }, Y, ldy, N, K, k, n0, lk, beta, zero_g0_im, nm, wn_s,
               fn_s, L);
}

// value of v in lane l (l uniform across the wave)
__device__ __forceinline__ double bcast_lane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// S = U^T U for the upper triangle of the n x n block at A (lda; read
// only) and X = U^-1 into rows [0, nz) of columns [0, n) of X (ldx; zero
// below the diagonal), by one block of NT waves, n <= 16 NT.
// Elimination step k subtracts (s_ki / s_kk) times row k from each row
// i > k, both in the upper triangle (s, diagonal included) and in the
// strictly lower triangle, which starts as the zeros of the identity (with
// its unit diagonal implicit) and ends as M = L^-1 of S = L D L^T.  Then
// d_c = s_cc and U^-1(r, c) = M(c, r) / sqrt(d_c) for r < c, 1 / sqrt(d_c)
// for r = c.  The array lives in MFMA accumulators as 16 x 16 tiles; wave J
// holds tile column J: tiles (I, J) above the diagonal hold s, below it M,
// and the diagonal tile has one copy of each (each half read from its own).
// Steps run in panels of PB rows, two barriers per panel.  The owners
// publish the panel's rows to LDS.  The waves that hold columns (thread
// t < n: column t) eliminate inside the panel: lane l < PB repeats it on the
// panel's own column k0 + l, where row k_q's lower part reads 1 at k_q and 0
// beyond, with pivots and multipliers taken from those lanes by readlane
// (the same values in every such wave); every other column takes the same
// row operations.  The eliminated rows R go to LDS, with the lower-part view
// RM (R before k_q, 1 at k_q, 0 after).  Then each tile below the panel
// takes the panel's PB row operations as one rank-PB MFMA update,
// T -= F R (upper) or T -= F RM (lower), F(i, q) = R(q, i) / d_q, and tiles
// that it leaves unchanged are skipped.  A non-positive pivot k sets
// info = info_base + k + 1 and the whole block returns; the block with
// info_base 0 first clears info, and a later block (info_base > 0) does
// nothing when an earlier one has already failed, so info names the first
// failing leading minor.
template <int NT, int PB>
__global__ void __launch_bounds__(64 * NT) k_chol_inv(const double* __restrict__ A, int lda, int n,
                                                     double* __restrict__ X, int ldx, int nz,
                                                     int* __restrict__ info, int info_base) {
  constexpr int NM = 16 * NT, NMP = NM + 2;
  static_assert(PB % 4 == 0 && 16 % PB == 0, "panels of 4, 8 or 16 rows inside one tile row");
  __shared__ double raw[PB][NM];
  __shared__ double R[PB][NMP];
  __shared__ double RM[PB][NMP];
  __shared__ double dinv_p[PB];
  __shared__ double dinv[NM];
  __shared__ int fail;
  const int tid = threadIdx.x, lane = tid & 63, J = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lk = lane >> 4, xj = 16 * J + li;  // this lane's column in its tiles
  const bool jlive = 16 * J < n;
  // earlier leaves ran before this launch on the same stream: the same value in every thread
  if (info_base > 0 && *(volatile const int*)info != 0) return;
  if (tid == 0) {
    fail = 0;
    if (info_base == 0) *info = 0;
  }
  for (int e = tid; e < PB * NMP; e += 64 * NT) {  // columns >= n stay zero
    (&R[0][0])[e] = 0.0;
    (&RM[0][0])[e] = 0.0;
  }
  dbl4 acc[NT], accd;  // tile (I, J) for I = 0 .. NT-1; accd: lower copy of (J, J)
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * I + 4 * v + lk;
      acc[I][v] = (I <= J && i <= xj && xj < n) ? A[i + (size_t)xj * lda] : 0.0;
    }
  accd = (dbl4){0.0, 0.0, 0.0, 0.0};
  // raw[p][x] = row k0 + p of the combined array, from tile row k0 / 16
  auto publish = [&](int k0) {
    if (!jlive) return;
#pragma unroll
    for (int I = 0; I < NT; ++I) {
      if (I != k0 / 16) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * I + 4 * v + lk, p = i - k0;
        if (p >= 0 && p < PB) raw[p][xj] = (I == J && xj < i) ? accd[v] : acc[I][v];
      }
    }
  };
  __syncthreads();
  publish(0);
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += PB) {
    const int pn = min(PB, n - k0);
    if (tid - lane < n) {  // waves holding columns
      const int x = tid, y = k0 + lane;
      double col[PB], pc[PB];
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        col[c] = (x < n && c < pn) ? raw[c][x] : 0.0;
        pc[c] = (lane < pn && c < pn) ? raw[c][min(y, NM - 1)] : 0.0;
      }
      int bad = PB;  // first non-positive pivot of the panel
      double dinv_q[PB];
      // rows q >= pn (past n) are zero and run with pivot 1, so the loops have
      // no data-dependent exits and the independent updates interleave
#pragma unroll
      for (int q = 0; q < PB; ++q) {
        const double dp = bcast_lane(pc[q], q);  // R_q(k_q), lane q
        if (q < pn && !(dp > 0.0) && bad == PB) bad = q;
        const double dq = q < pn ? dp : 1.0;
        // 1 / d_q: hardware estimate and two Newton steps
        double inv = __builtin_amdgcn_rcp(dq);
        inv = fma(inv, fma(-dq, inv, 1.0), inv);
        inv = fma(inv, fma(-dq, inv, 1.0), inv);
        dinv_q[q] = inv;
        const int kq = k0 + q;
#pragma unroll
        for (int c = q + 1; c < PB; ++c) {
          const double f = bcast_lane(pc[q], c) * inv;  // R_q(k_c) / d_q, lane c
          const int kc = k0 + c;
          pc[c] -= f * ((y < kq || y >= kc) ? pc[q] : (y == kq ? 1.0 : 0.0));
          col[c] -= f * col[q];  // exact for columns outside the panel
        }
      }
      if (bad < PB) {  // the same in every column wave: the block returns after the barrier
        if (tid == 0) {
          *info = info_base + k0 + bad + 1;
          fail = 1;
        }
      } else {
        const bool in_panel = x >= k0 && x < k0 + PB;
#pragma unroll
        for (int c = 0; c < PB; ++c) {
          if (c >= pn) continue;
          if (x < n && !in_panel) {
            R[c][x] = col[c];
            RM[c][x] = x < k0 ? col[c] : 0.0;
          }
          if (tid < pn) {  // wave 0 writes the panel's columns
            R[c][k0 + tid] = pc[c];
            RM[c][k0 + tid] = tid < c ? pc[c] : (tid == c ? 1.0 : 0.0);
          }
          if (tid == 0) dinv_p[c] = dinv_q[c];
        }
      }
    }
    __syncthreads();
    if (fail) return;  // the whole block
    if (jlive) {
      // operand rows of F for tile row I: i = 16 I + li, panel step q = q0 + lk
      const bool lower_live = 16 * J <= k0 + PB - 1;  // lower tiles right of the panel stay unchanged
#pragma unroll
      for (int I = 0; I < NT; ++I) {
        if (16 * I + 15 < k0 + pn || 16 * I >= n) continue;  // rows above or past the panel's reach
        if (I > J && !lower_live) continue;
        const int i = 16 * I + li;
        const bool below = i >= k0 + pn && i < n;
#pragma unroll
        for (int q0 = 0; q0 < PB; q0 += 4) {
          const int q = q0 + lk;
          const double a = (below && q < pn) ? -R[q][i] * dinv_p[q] : 0.0;
          if (I < J) acc[I] = mfma16(a, R[q][xj], acc[I]);
          else if (I > J) acc[I] = mfma16(a, RM[q][xj], acc[I]);
          else {
            acc[I] = mfma16(a, R[q][xj], acc[I]);
            if (lower_live) accd = mfma16(a, RM[q][xj], accd);
          }
        }
      }
      // the panel's rows take their eliminated values
#pragma unroll
      for (int I = 0; I < NT; ++I) {
        if (I != k0 / 16) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * I + 4 * v + lk, p = i - k0;
          if (p >= 0 && p < pn && xj < n) {
            acc[I][v] = R[p][xj];
            if (I == J) accd[v] = R[p][xj];
          }
        }
      }
    }
    if (k0 + PB < n) publish(k0 + PB);
    __syncthreads();
  }
  if (jlive) {
#pragma unroll
    for (int I = 0; I < NT; ++I) {
      if (I != J) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * I + 4 * v + lk;
        if (xj == i && i < n) dinv[i] = 1.0 / sqrt(acc[I][v]);
      }
    }
  }
  __syncthreads();
  if (jlive && xj < n) {
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int c = 16 * I + 4 * v + lk, r = xj;  // X(r, c) from element (c, r)
        if (c >= n) continue;
        double o;
        if (r > c) o = 0.0;
        else if (r == c) o = dinv[c];
        else o = (I == J ? accd[v] : acc[I][v]) * dinv[c];
        X[r + (size_t)c * ldx] = o;
      }
  }
  const int zr = nz - n;
  for (int e = tid; e < zr * n; e += 64 * NT) X[n + e % zr + (size_t)(e / zr) * ldx] = 0.0;
}

// f(integral_constant<int, 0>), f(..., 1), ... while f returns true
template <class F, int... P>
__device__ __forceinline__ void for_each_index(F& f, std::integer_sequence<int, P...>) {
  (void)(f(std::integral_constant<int, P>{}) && ...);
}

// k_chol_inv with look-ahead: the same elimination, row operations, tile
// ownership and per-tile order of updates, with the panel elimination on
// NT/4 waves of its own (waves NT/4 .. NT/2 - 1; lane l of the e-th holds
// column 64 e + l), so that it runs
// while the tile waves apply the previous panel's MFMA updates.  Panel p's
// pivot rows (raw), eliminated rows (R, RM) and 1/d (dinv_p) sit in LDS
// buffers p & 1.  Tile wave J at panel p: waits until panel p is eliminated
// (count `elim`), updates the tile row that holds the rows of panel p + 1
// and publishes them (count `pub`), then updates its other tiles.  The
// elimination waves wait until every tile wave has published panel p + 1,
// then eliminate it.  A tile wave publishes panel p + 1 after its last read
// of buffer (p + 1) & 1 (panel p - 1), so no buffer is overwritten while in
// use.  Every wait polls at most CHOL_SPINS times; a non-positive pivot
// (info as in k_chol_inv) or an exhausted wait (info = -1) sets `fail`, and
// every wave then returns.
constexpr int CHOL_SPINS = 1 << 22;

template <int NT, int PB>
__global__ void __launch_bounds__(80 * NT) k_chol_inv_la(const double* __restrict__ A, int lda, int n,
                                                        double* __restrict__ X, int ldx, int nz,
                                                        int* __restrict__ info, int info_base) {
  constexpr int NM = 16 * NT, NMP = NM + 2, NE = NT / 4;
  static_assert(PB % 4 == 0 && 16 % PB == 0, "panels of 4, 8 or 16 rows inside one tile row");
  static_assert(NT % 4 == 0, "one elimination wave per 64 columns");
  __shared__ double raw[2][PB][NM];
  __shared__ double R[2][PB][NMP];
  __shared__ double RM[2][PB][NMP];
  __shared__ double dinv_p[2][PB];
  __shared__ double dinv[NM];
  __shared__ int pub, elim, fail;
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  // earlier leaves ran before this launch on the same stream: the same value in every thread
  if (info_base > 0 && *(volatile const int*)info != 0) return;
  if (tid == 0) {
    fail = 0;
    pub = 0;
    elim = 0;
    if (info_base == 0) *info = 0;
  }
  for (int e = tid; e < 2 * PB * NMP; e += blockDim.x) {  // columns >= n stay zero
    (&R[0][0][0])[e] = 0.0;
    (&RM[0][0][0])[e] = 0.0;
  }
  // poll *cnt until it reaches target; false when the block has failed
  auto wait_for = [&](int* cnt, int target) -> bool {
    for (int s = 0;; ++s) {
      if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(&fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
        return false;
      if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) >=
          target)
        return true;
      if (s >= CHOL_SPINS) {
        if (lane == 0) {
          *info = -1;
          __hip_atomic_store(&fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        return false;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  };
  auto signal = [&](int* cnt) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const int np = (n + PB - 1) / PB;
  // The two wave kinds run separate code; each passes the same four block
  // barriers (set-up, first panel published, loops done, 1/sqrt(d) ready)
  // unless the block fails, in which case every wave returns before the third.
  // waves NE .. 2 NE - 1 eliminate (each on its own SIMD, beside fewer tile
  // waves than SIMDs 0 and 1 carry), the others hold tile columns
  const int ew = wv - NE;
  if (ew >= 0 && ew < NE) {
    // elimination waves: column x; lane l < PB also carries the panel's own column k0 + l
    const int x = 64 * ew + lane;
    __syncthreads();
    __syncthreads();
    for (int p = 0; p < np; ++p) {
      const int k0 = p * PB, pn = min(PB, n - k0), b = p & 1, y = k0 + lane;
      if (p > 0 && !wait_for(&pub, NT * p)) return;
      double col[PB], pc[PB];
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        col[c] = (x < n && c < pn) ? raw[b][c][x] : 0.0;
        pc[c] = (lane < pn && c < pn) ? raw[b][c][min(y, NM - 1)] : 0.0;
      }
      int bad = PB;  // first non-positive pivot of the panel
      double dinv_q[PB];
      // rows q >= pn (past n) are zero and run with pivot 1
#pragma unroll
      for (int q = 0; q < PB; ++q) {
        const double dp = bcast_lane(pc[q], q);  // R_q(k_q), lane q
        if (q < pn && !(dp > 0.0) && bad == PB) bad = q;
        const double dq = q < pn ? dp : 1.0;
        // 1 / d_q: hardware estimate and two Newton steps
        double inv = __builtin_amdgcn_rcp(dq);
        inv = fma(inv, fma(-dq, inv, 1.0), inv);
        inv = fma(inv, fma(-dq, inv, 1.0), inv);
        dinv_q[q] = inv;
        const int kq = k0 + q;
#pragma unroll
        for (int c = q + 1; c < PB; ++c) {
          const double f = bcast_lane(pc[q], c) * inv;  // R_q(k_c) / d_q, lane c
          const int kc = k0 + c;
          pc[c] -= f * ((y < kq || y >= kc) ? pc[q] : (y == kq ? 1.0 : 0.0));
          col[c] -= f * col[q];  // exact for columns outside the panel
        }
      }
      if (bad < PB) {  // the same in every elimination wave
        if (lane == 0) {
          if (ew == 0) *info = info_base + k0 + bad + 1;
          __hip_atomic_store(&fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        return;
      }
      const bool in_panel = x >= k0 && x < k0 + PB;
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        if (c >= pn) continue;
        if (x < n && !in_panel) {
          R[b][c][x] = col[c];
          RM[b][c][x] = x < k0 ? col[c] : 0.0;
        }
        if (ew == 0 && lane < pn) {  // the first elimination wave writes the panel's columns
          R[b][c][k0 + lane] = pc[c];
          RM[b][c][k0 + lane] = lane < c ? pc[c] : (lane == c ? 1.0 : 0.0);
        }
        if (ew == 0 && lane == 0) dinv_p[b][c] = dinv_q[c];
      }
      signal(&elim);
    }
    __syncthreads();
    __syncthreads();
  } else {
    // tile waves: wave J holds tile column J
    const int J = wv < NE ? wv : wv - NE;
    const int li = lane & 15, lk = lane >> 4, xj = 16 * J + li;  // this lane's column in its tiles
    const bool jlive = 16 * J < n;
    dbl4 acc[NT], accd;  // tile (I, J) for I = 0 .. NT-1; accd: lower copy of (J, J)
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * I + 4 * v + lk;
        acc[I][v] = (I <= J && i <= xj && xj < n) ? A[i + (size_t)xj * lda] : 0.0;
      }
    accd = (dbl4){0.0, 0.0, 0.0, 0.0};
    // raw[b][p][x] = row k0 + p of the combined array, from tile row k0 / 16
    auto publish = [&](int k0, int b) {
#pragma unroll
      for (int I = 0; I < NT; ++I) {
        if (I != k0 / 16) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * I + 4 * v + lk, p = i - k0;
          if (p >= 0 && p < PB) raw[b][p][xj] = (I == J && xj < i) ? accd[v] : acc[I][v];
        }
      }
    };
    // the rank-PB update of tile row I by panel buffer b
    auto upd_tile = [&](int I, int k0, int pn, int b, bool lower_live) {
      // operand rows of F for tile row I: i = 16 I + li, panel step q = q0 + lk
      const int i = 16 * I + li;
      const bool below = i >= k0 + pn && i < n;
#pragma unroll
      for (int q0 = 0; q0 < PB; q0 += 4) {
        const int q = q0 + lk;
        const double a = (below && q < pn) ? -R[b][q][i] * dinv_p[b][q] : 0.0;
#pragma unroll
        for (int T = 0; T < NT; ++T) {
          if (T != I) continue;
          if (T < J) acc[T] = mfma16(a, R[b][q][xj], acc[T]);
          else if (T > J) acc[T] = mfma16(a, RM[b][q][xj], acc[T]);
          else {
            acc[T] = mfma16(a, R[b][q][xj], acc[T]);
            if (lower_live) accd = mfma16(a, RM[b][q][xj], accd);
          }
        }
      }
    };
    // the panel's rows take their eliminated values
    auto take_panel = [&](int I, int k0, int pn, int b) {
#pragma unroll
      for (int T = 0; T < NT; ++T) {
        if (T != I) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * T + 4 * v + lk, pp = i - k0;
          if (pp >= 0 && pp < pn && xj < n) {
            acc[T][v] = R[b][pp][xj];
            if (T == J) accd[v] = R[b][pp][xj];
          }
        }
      }
    };
    __syncthreads();
    if (jlive) publish(0, 0);
    __syncthreads();
    // one copy of the panel step per panel index (compile-time p), so that
    // the tile rows it touches, and the register tiles that hold them, are
    // static; a step returns false to end the walk (past n, or failed)
    bool failed = false;
    auto step = [&](auto pc) -> bool {
      constexpr int p = decltype(pc)::value;
      constexpr int k0 = p * PB, b = p & 1, In = (k0 + PB) / 16, Ip = k0 / 16;
      if (k0 >= n) return false;
      const int pn = min(PB, n - k0);
      if (!wait_for(&elim, NE * (p + 1))) {
        failed = true;
        return false;
      }
      const bool nxt = k0 + PB < n;
      const bool lower_live = 16 * J <= k0 + PB - 1;  // lower tiles right of the panel stay unchanged
      // tile rows above or past the panel's reach, and below the diagonal
      // while their lower part is not live, are left unchanged
      auto live_row = [&](int I) { return !(16 * I + 15 < k0 + pn || 16 * I >= n) && !(I > J && !lower_live); };
      if (jlive && nxt && In < NT) {
        if (live_row(In)) upd_tile(In, k0, pn, b, lower_live);
        if (In == Ip) take_panel(Ip, k0, pn, b);
        publish(k0 + PB, b ^ 1);
      }
      if (nxt) signal(&pub);
      if (jlive) {
#pragma unroll
        for (int I = 0; I < NT; ++I) {
          if (nxt && I == In) continue;
          if (live_row(I)) upd_tile(I, k0, pn, b, lower_live);
          if (I == Ip) take_panel(Ip, k0, pn, b);
        }
      }
      return true;
    };
    for_each_index(step, std::make_integer_sequence<int, NT * 16 / PB>{});
    if (failed) return;
    __syncthreads();
    if (jlive) {
#pragma unroll
      for (int I = 0; I < NT; ++I) {
        if (I != J) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * I + 4 * v + lk;
          if (xj == i && i < n) dinv[i] = 1.0 / sqrt(acc[I][v]);
        }
      }
    }
    __syncthreads();
    if (jlive && xj < n) {
#pragma unroll
      for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = 16 * I + 4 * v + lk, r = xj;  // X(r, c) from element (c, r)
          if (c >= n) continue;
          double o;
          if (r > c) o = 0.0;
          else if (r == c) o = dinv[c];
          else o = (I == J ? accd[v] : acc[I][v]) * dinv[c];
          X[r + (size_t)c * ldx] = o;
        }
    }
  }
  const int zr = nz - n;
  for (int e = tid; e < zr * n; e += blockDim.x) X[n + e % zr + (size_t)(e / zr) * ldx] = 0.0;
}

// W <- I (n x n)
__global__ void k_identity(double* __restrict__ W, int n, int ld) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x)
    W[e % n + (size_t)(e / n) * ld] = (e % n == e / n) ? 1.0 : 0.0;
}

// C = alpha op(A) op(B) + beta C (+ diag on the diagonal) for the small
// band-space products (blocked Cholesky, Newton-Schulz, the projector
// correction of the Lagrange matrix), FP64 on v_mfma_f64_16x16x4f64: one
// block per 16 x 16 tile of C, its four waves each summing a quarter of K,
// the quarters added in a fixed order.  op(X) = X or X^T (column-major
// storage); beta == 0 ignores C's contents.
__global__ void __launch_bounds__(256) k_gemm_small_mfma(int ta, int tb, int M, int N, int K, double alpha,
                                                         const double* __restrict__ A, int lda,
                                                         const double* __restrict__ B, int ldb, double beta,
                                                         double* __restrict__ C, int ldc, double diag) {
  __shared__ double red[3][4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  const int kq = ((K + 15) / 16) * 4;  // K per wave, a multiple of 4
  const int kb = wave * kq, ke = min(K, kb + kq);
  const int i = i0 + li, j = j0 + li;
  dbl4 acc = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k0 = kb; k0 < ke; k0 += 4) {
    const int k = k0 + lk;
    const bool kin = k < ke;
    const double a = (kin && i < M) ? (ta ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda]) : 0.0;
    const double b = (kin && j < N) ? (tb ? B[j + (size_t)k * ldb] : B[k + (size_t)j * ldb]) : 0.0;
    acc = mfma16(a, b, acc);
  }
  if (wave > 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) red[wave - 1][v][lane] = acc[v];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double sum = ((acc[v] + red[0][v][lane]) + red[1][v][lane]) + red[2][v][lane];
      const int r = i0 + 4 * v + lk, cidx = j0 + li;  // element (r, cidx) of this lane
      if (r < M && cidx < N) {
        double* cp = C + r + (size_t)cidx * ldc;
        *cp = alpha * sum + (beta != 0.0 ? beta * *cp : 0.0) + (r == cidx ? diag : 0.0);
      }
    }
  }
}

// C = alpha op(A) op(B) + beta C for the small band-block products of the
// blocked Cholesky (FP64, 16 x 16 output tiles, K in LDS chunks of 16);
// op(X) = X or X^T (column-major storage).  beta == 0 ignores C's contents.
__global__ void __launch_bounds__(256) k_gemm_small(int ta, int tb, int M, int N, int K, double alpha,
                                                    const double* __restrict__ A, int lda,
                                                    const double* __restrict__ B, int ldb, double beta,
                                                    double* __restrict__ C, int ldc) {
  __shared__ double as[16][17], bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i = blockIdx.x * 16 + tx, j = blockIdx.y * 16 + ty;
  double acc = 0.0;
  for (int k0 = 0; k0 < K; k0 += 16) {
    {  // as[r][kk] = op(A)(blockIdx.x*16 + r, k0 + kk)
      const int r = tx, kk = ty, ii = blockIdx.x * 16 + r, k = k0 + kk;
      as[r][kk] = (ii < M && k < K) ? (ta ? A[k + (size_t)ii * lda] : A[ii + (size_t)k * lda]) : 0.0;
    }
    {  // bs[kk][c] = op(B)(k0 + kk, blockIdx.y*16 + c)
      const int kk = tx, cc = ty, jj = blockIdx.y * 16 + cc, k = k0 + kk;
      bs[kk][cc] = (jj < N && k < K) ? (tb ? B[jj + (size_t)k * ldb] : B[k + (size_t)jj * ldb]) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) acc += as[tx][kk] * bs[kk][ty];
    __syncthreads();
  }
  if (i < M && j < N) {
    double* c = C + i + (size_t)j * ldc;
    *c = alpha * acc + (beta != 0.0 ? beta * *c : 0.0);
  }
}

}  // namespace amdk
