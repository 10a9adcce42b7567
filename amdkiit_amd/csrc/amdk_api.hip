// libamdk: context creation, plane-wave maps, transfers and the H|psi> /
// density drivers behind the C ABI declared in include/amdk.h.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <tuple>

#include "amdk_ctx.h"
#include "fft_reg.h"

using namespace amdk;

namespace amdk {

FftPlan1D make_plan(int n) {
  FftPlan1D p;
  p.n = n;
  p.nst = 0;
  if (n < 2) throw Error(AMDK_ERR_ARG, "FFT length must be >= 2");
  int m = n;
  while (m % 4 == 0 && m > 4) { p.rad[p.nst++] = 4; m /= 4; }
  const int rs[4] = {5, 3, 4, 2};
  for (int r : rs) {
    while (m % r == 0 && m > 1) {
      if (p.nst >= AMDK_MAX_STAGES) throw Error(AMDK_ERR_ARG, "FFT length has too many factors");
      p.rad[p.nst++] = r;
      m /= r;
    }
  }
  if (m != 1) throw Error(AMDK_ERR_ARG, "FFT length " + std::to_string(n) + " is not 2^a 3^b 5^c");
  return p;
}

void* dev_alloc(amdk_ctx* c, size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) throw Error(AMDK_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  c->owned.push_back(p);
  AMDK_TRY_HIP(hipMemsetAsync(p, 0, bytes, c->stream));
  return p;
}

// Device buffers handed in by the caller may have been written on any
// stream; the library stream does not order after them, so wait for the device.
static void caller_ready() { AMDK_TRY_HIP(hipDeviceSynchronize()); }

// Host <-> device copies go through the library stream (which does not
// synchronise with the null stream), after the zero fill of dev_alloc and any
// queued kernels, and are complete when these return.
void upload(amdk_ctx* c, void* dst, const void* src, size_t bytes) {
  AMDK_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  AMDK_TRY_HIP(hipStreamSynchronize(c->stream));
}
void download(amdk_ctx* c, void* dst, const void* src, size_t bytes) {
  AMDK_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  AMDK_TRY_HIP(hipStreamSynchronize(c->stream));
}

size_t cbytes(const amdk_ctx* c) { return c->prec == AMDK_PREC_FP64 ? sizeof(double2) : sizeof(float2); }
size_t sbytes(const amdk_ctx*) { return sizeof(double2); }

int prof_begin(amdk_ctx* c, const char* name) {
  if (!c->prof) return -1;
  ProfSlot s;
  s.name = name;
  for (hipEvent_t* e : {&s.a, &s.b}) {
    if (!c->event_pool.empty()) {
      *e = c->event_pool.back();
      c->event_pool.pop_back();
    } else {
      AMDK_TRY_HIP(hipEventCreate(e));
    }
  }
  AMDK_TRY_HIP(hipEventRecord(s.a, c->stream));
  c->pending.push_back(s);
  return (int)c->pending.size() - 1;
}

void prof_end(amdk_ctx* c, int slot) {
  if (slot < 0) return;
  AMDK_TRY_HIP(hipEventRecord(c->pending[slot].b, c->stream));
}

void sync(amdk_ctx* c) {
  AMDK_TRY_HIP(hipStreamSynchronize(c->stream));
  for (auto& s : c->pending) {
    float ms = 0;
    AMDK_TRY_HIP(hipEventElapsedTime(&ms, s.a, s.b));
    auto& acc = c->prof_acc[s.name];
    acc.first += ms;
    acc.second += 1;
    c->event_pool.push_back(s.a);
    c->event_pool.push_back(s.b);
  }
  c->pending.clear();
}

}  // namespace amdk

// ---------------------------------------------------------------------------
// Sphere construction (host).  Canonicalise every caller G to the z-half,
// sort by (m3, m2, m1) and cut into half rows.
static void build_sphere(amdk_ctx* c, const int* mill, int ncaller, Sphere& S) {
  if (ncaller <= 0 || !mill) throw Error(AMDK_ERR_ARG, "empty Miller list");
  std::vector<std::array<int, 3>> m(ncaller);
  S.h_conj_caller.assign(ncaller, 0);
  for (int i = 0; i < ncaller; ++i) {
    int a = mill[3 * i], b = mill[3 * i + 1], d = mill[3 * i + 2];
    const bool zhalf = d > 0 || (d == 0 && (b > 0 || (b == 0 && a >= 0)));
    if (!zhalf) { a = -a; b = -b; d = -d; S.h_conj_caller[i] = 1; }
    if (2 * std::abs(a) >= c->n1 || 2 * std::abs(b) >= c->n2 || 2 * std::abs(d) >= c->n3)
      throw Error(AMDK_ERR_ARG, "Miller index outside the FFT grid (|m| must be < n/2)");
    m[i] = {a, b, d};
  }
  std::vector<int> ord(ncaller);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) {
    return std::make_tuple(m[x][2], m[x][1], m[x][0]) < std::make_tuple(m[y][2], m[y][1], m[y][0]);
  });
  S.npw = ncaller;
  S.ncaller = ncaller;
  S.ldc = (ncaller + 63) / 64 * 64;
  S.h_from_caller.assign(ncaller, 0);
  S.h_mill.assign(3 * ncaller, 0);
  for (int k = 0; k < ncaller; ++k) {
    const int i = ord[k];
    if (k > 0 && m[ord[k - 1]] == m[i]) throw Error(AMDK_ERR_ARG, "duplicate plane wave (G or -G listed twice)");
    S.h_from_caller[i] = k;
    for (int t = 0; t < 3; ++t) S.h_mill[3 * k + t] = m[i][t];
  }
  // half rows
  S.h_rows.clear();
  for (int k = 0; k < ncaller;) {
    RowInfo r{};
    r.start = k;
    r.m1_0 = S.h_mill[3 * k];
    r.m2 = S.h_mill[3 * k + 1];
    r.m3 = S.h_mill[3 * k + 2];
    int j = k + 1;
    while (j < ncaller && S.h_mill[3 * j + 1] == r.m2 && S.h_mill[3 * j + 2] == r.m3) {
      if (S.h_mill[3 * j] != S.h_mill[3 * (j - 1)] + 1)
        throw Error(AMDK_ERR_ARG, "plane-wave set is not convex along m1 (row has a gap)");
      ++j;
    }
    r.cnt = j - k;
    S.h_rows.push_back(r);
    k = j;
  }
  S.nrow = (int)S.h_rows.size();
  // full rows -> slots sorted by (m3, m2)
  struct FR { int m3, m2, row, mirror; };
  std::vector<FR> fr;
  S.R = 0;
  for (int r = 0; r < S.nrow; ++r) {
    const RowInfo& ri = S.h_rows[r];
    fr.push_back({ri.m3, ri.m2, r, 0});
    if (!(ri.m2 == 0 && ri.m3 == 0)) fr.push_back({-ri.m3, -ri.m2, r, 1});
    S.R = std::max(S.R, std::abs(ri.m3));
  }
  std::sort(fr.begin(), fr.end(), [](const FR& x, const FR& y) { return std::make_pair(x.m3, x.m2) < std::make_pair(y.m3, y.m2); });
  S.nslot = (int)fr.size();
  S.nplane = 2 * S.R + 1;
  S.h_slot_m2.assign(S.nslot, 0);
  S.h_plane_off.assign(S.nplane + 1, 0);
  for (int s = 0; s < S.nslot; ++s) {
    S.h_slot_m2[s] = fr[s].m2;
    RowInfo& ri = S.h_rows[fr[s].row];
    if (fr[s].mirror) ri.slot_m = s; else ri.slot_d = s;
    S.h_plane_off[fr[s].m3 + S.R + 1] += 1;
  }
  for (int r = 0; r < S.nrow; ++r)
    if (S.h_rows[r].m2 == 0 && S.h_rows[r].m3 == 0) S.h_rows[r].slot_m = -1;
  for (int p = 0; p < S.nplane; ++p) S.h_plane_off[p + 1] += S.h_plane_off[p];
  if (S.h_plane_off[S.nplane] != S.nslot) throw Error(AMDK_ERR_ARG, "internal: slot count mismatch");
  // the register passes address a plane's rows as m2_lo + offset
  S.planes_contiguous = true;
  for (int p = 0; p < S.nplane; ++p)
    for (int s = S.h_plane_off[p] + 1; s < S.h_plane_off[p + 1]; ++s)
      if (S.h_slot_m2[s] != S.h_slot_m2[s - 1] + 1) S.planes_contiguous = false;
  // upload
  S.rows = (RowInfo*)dev_alloc(c, sizeof(RowInfo) * S.nrow);
  S.slot_m2 = (int*)dev_alloc(c, sizeof(int) * S.nslot);
  S.plane_off = (int*)dev_alloc(c, sizeof(int) * (S.nplane + 1));
  S.d_from_caller = (int*)dev_alloc(c, sizeof(int) * ncaller);
  S.d_conj_caller = (int*)dev_alloc(c, sizeof(int) * ncaller);
  upload(c, S.rows, S.h_rows.data(), sizeof(RowInfo) * S.nrow);
  upload(c, S.slot_m2, S.h_slot_m2.data(), sizeof(int) * S.nslot);
  upload(c, S.plane_off, S.h_plane_off.data(), sizeof(int) * (S.nplane + 1));
  upload(c, S.d_from_caller, S.h_from_caller.data(), sizeof(int) * ncaller);
  upload(c, S.d_conj_caller, S.h_conj_caller.data(), sizeof(int) * ncaller);
}

// Row pitch of the W1/W2 work arrays: x rows padded to 16 elements (128 B in
// FP32, 256 B in FP64) so that the x-tiles of the y and z passes start on
// cache-line boundaries.
static int row_pitch(int n1) { return (n1 + 15) / 16 * 16; }

static PassGeom geom_of(const amdk_ctx* c, const Sphere& S) {
  PassGeom g;
  g.n1 = c->n1; g.n2 = c->n2; g.n3 = c->n3;
  g.p1 = row_pitch(c->n1);
  g.nrow = S.nrow; g.nslot = S.nslot; g.nplane = S.nplane; g.R = S.R;
  g.rows = S.rows; g.slot_m2 = S.slot_m2; g.plane_off = S.plane_off;
  return g;
}

// ---------------------------------------------------------------------------
// transfer kernels between caller-order FP64 arrays and internal arrays
template <typename V>
__global__ void k_from_caller(const double2* __restrict__ src, int ldsrc, V* __restrict__ dst, int ldc,
                              const int* __restrict__ pos, const int* __restrict__ cj, int n, int nb) {
  const int b = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double2 v = src[(size_t)b * ldsrc + i];
    if (cj[i]) v.y = -v.y;
    V o;
    o.x = v.x;
    o.y = v.y;
    dst[(size_t)b * ldc + pos[i]] = o;
  }
}
template <typename V>
__global__ void k_to_caller(const V* __restrict__ src, int ldc, double2* __restrict__ dst, int lddst,
                            const int* __restrict__ pos, const int* __restrict__ cj, int n, int nb) {
  const int b = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const V v = src[(size_t)b * ldc + pos[i]];
    double2 o;
    o.x = v.x;
    o.y = cj[i] ? -(double)v.y : (double)v.y;
    dst[(size_t)b * lddst + i] = o;
  }
}

template <typename T>
__global__ void k_scale_copy(const double* __restrict__ src, T* __restrict__ dst, size_t n, double s) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (T)(src[i] * s);
}

// Caller-layout FP64 host arrays <-> internal device layout.  The caller's
// first npw coefficients of each band are copied with their pitch into the
// context's staging array (pitch ldc), then permuted on the device; nothing
// is allocated per call.  Entries of the caller's array beyond npw are
// neither read nor written.
static void bands_in(amdk_ctx* c, const double* host, int ldsrc, int nb, void* dst) {
  const Sphere& S = c->spw;
  AMDK_TRY_HIP(hipMemcpy2DAsync(c->d_stage, sizeof(double2) * S.ldc, host, sizeof(double2) * ldsrc,
                                sizeof(double2) * S.npw, nb, hipMemcpyHostToDevice, c->stream));
  AMDK_TRY_HIP(hipMemsetAsync(dst, 0, sbytes(c) * (size_t)S.ldc * nb, c->stream));
  dim3 grid((S.npw + 255) / 256, nb);
  k_from_caller<double2><<<grid, 256, 0, c->stream>>>((const double2*)c->d_stage, S.ldc, (double2*)dst, S.ldc,
                                                      S.d_from_caller, S.d_conj_caller, S.npw, nb);
  AMDK_TRY_HIP(hipGetLastError());
  AMDK_TRY_HIP(hipStreamSynchronize(c->stream));
}

static void bands_out(amdk_ctx* c, const void* src, int nb, double* host, int lddst) {
  const Sphere& S = c->spw;
  dim3 grid((S.npw + 255) / 256, nb);
  k_to_caller<double2><<<grid, 256, 0, c->stream>>>((const double2*)src, S.ldc, (double2*)c->d_stage, S.ldc,
                                                    S.d_from_caller, S.d_conj_caller, S.npw, nb);
  AMDK_TRY_HIP(hipGetLastError());
  AMDK_TRY_HIP(hipMemcpy2DAsync(host, sizeof(double2) * lddst, c->d_stage, sizeof(double2) * S.ldc,
                                sizeof(double2) * S.npw, nb, hipMemcpyDeviceToHost, c->stream));
  AMDK_TRY_HIP(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// tile shapes per precision
template <typename T> struct Tiles;
template <> struct Tiles<float> { static constexpr int RPB = 8, TX = 16, PG = 4; };
template <> struct Tiles<double> { static constexpr int RPB = 4, TX = 8, PG = 2; };

// ---------------------------------------------------------------------------
// register-FFT dispatch: lengths with an (A, B) split get the kernels of
// fft_reg.h; other lengths use the LDS Stockham passes above.
#define AMDK_REG_SIZES(X) \
  X(20, 5, 4) X(24, 6, 4) X(30, 6, 5) X(36, 6, 6) X(48, 8, 6) X(60, 10, 6) X(72, 9, 8) X(100, 10, 10) \
  X(108, 12, 9) X(144, 12, 12) X(180, 15, 12)

static bool reg_ok(const amdk_ctx* c, int n) {
  if (c->fft_generic) return false;
  switch (n) {
#define AMDK_C(N, A, B) case N: return true;
    AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
    default: return false;
  }
}

// Launch shapes of the register passes: TX x columns per block (y passes,
// x tiles within one plane row: walking the columns across planes instead
// measured slower, these passes being bound by memory), ZTX columns per
// block for the z passes, which walk the n1 x n2 columns in one x-fastest
// sequence, so no x tile is partly past the row end: 16 (at TM = 10 the
// third wave of the 160 threads is half empty, yet 16 columns measured
// faster than 32 or 64: smaller blocks leave fewer waves waiting at each of
// the transform's barriers; 8 was slower).
// HR half rows per block (x passes), PG band pairs sharing one V(r) tile in
// the z pass of H|psi> (2, 4 and 8 measured equal within 1 %).
template <typename T, int TM> struct RegTiles {
  static constexpr int TX = 16;
  static constexpr int ZTX = 16;
  static constexpr int HR = 240 / (2 * TM);
  static constexpr int PG = sizeof(T) == 4 ? 4 : 2;
};

template <typename T, int A, int B>
static void reg_z_vpsi(amdk_ctx* c, const Sphere& S, int np) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, TX = RegTiles<T, TM>::ZTX, PG = RegTiles<T, TM>::PG;
  const size_t lds = (size_t)(N + N * TX) * sizeof(V);
  const int ncb = (c->n1 * c->n2 + TX - 1) / TX, ngrp = (np + PG - 1) / PG, per = (ncb * ngrp + 7) / 8;
  k_z_vpsi_r<T, A, B, TX><<<8 * per, TX * TM, lds, c->stream>>>(geom_of(c, S), (V*)c->W2, (const T*)c->vr, np, PG,
                                                               (const V*)(sizeof(T) == 4 ? c->tw_f[2] : c->tw_d[2]),
                                                               ngrp, per);
}
template <typename T, int A, int B>
static void reg_z_rho(amdk_ctx* c, const Sphere& S, int np, double* rho, int accumulate, const double* wpair,
                      int groups, size_t part_stride) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, TX = RegTiles<T, TM>::ZTX;
  const size_t lds = (size_t)(N + N * TX) * sizeof(V);
  const int ncb = (c->n1 * c->n2 + TX - 1) / TX, per = (ncb * groups + 7) / 8;
  k_z_rho_r<T, A, B, TX><<<8 * per, TX * TM, lds, c->stream>>>(geom_of(c, S), (const V*)c->W2, rho, np, wpair, accumulate,
                                                              (const V*)(sizeof(T) == 4 ? c->tw_f[2] : c->tw_d[2]),
                                                              part_stride, groups, per);
}
template <typename T, int A, int B>
static void reg_cols(amdk_ctx* c, const Sphere& S, int np, bool inverse) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, TX = RegTiles<T, TM>::TX;
  const size_t lds = (size_t)(N + N * TX) * sizeof(V);
  dim3 grid((c->n1 + TX - 1) / TX, S.nplane, np);
  const V* tw = (const V*)(sizeof(T) == 4 ? c->tw_f[1] : c->tw_d[1]);
  if (inverse)
    k_cols_inv_r<T, A, B, TX><<<grid, TX * TM, lds, c->stream>>>(geom_of(c, S), (const V*)c->W1, (V*)c->W2, tw);
  else
    k_cols_fwd_r<T, A, B, TX><<<grid, TX * TM, lds, c->stream>>>(geom_of(c, S), (const V*)c->W2, (V*)c->W1, tw);
}
template <typename T, int A, int B>
static void reg_rows_inv(amdk_ctx* c, const Sphere& S, const void* C, int ldc, int nb, int pair0, int np) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, HR = RegTiles<T, TM>::HR;
  const size_t lds = (size_t)(N + 2 * HR * (N + 1)) * sizeof(V);
  dim3 grid((S.nrow + HR - 1) / HR, np);
  k_rows_inv_r<T, A, B, HR><<<grid, 2 * HR * TM, lds, c->stream>>>(geom_of(c, S), (const double2*)C, ldc, nb, pair0, (V*)c->W1,
                                                                  (const V*)(sizeof(T) == 4 ? c->tw_f[0] : c->tw_d[0]));
}
template <typename T, int A, int B>
static void reg_rows_fwd_hpsi(amdk_ctx* c, const Sphere& S, const void* C, void* C2, int nb, int pair0, int np) {
  using V = typename cplx_of<T>::type;
  constexpr int N = A * B, TM = MaxAB<A, B>::v, HR = RegTiles<T, TM>::HR;
  const size_t lds = (size_t)(N + 2 * HR * (N + 1)) * sizeof(V);
  dim3 grid((S.nrow + HR - 1) / HR, np);
  k_rows_fwd_hpsi_r<T, A, B, HR><<<grid, 2 * HR * TM, lds, c->stream>>>(geom_of(c, S), (const V*)c->W1, (const double2*)C,
                                                                       (double2*)C2, S.ldc, nb, pair0, c->d_occ, c->rc,
                                                                       (const V*)(sizeof(T) == 4 ? c->tw_f[0] : c->tw_d[0]));
}

template <typename T>
static void launch_rows_inv(amdk_ctx* c, const Sphere& S, const void* C, int ldc, int nb, int pair0, int np) {
  using V = typename cplx_of<T>::type;
  constexpr int RPB = Tiles<T>::RPB;
  const size_t lds = (size_t)(c->n1 + 4 * RPB * c->n1) * sizeof(V);
  dim3 grid((S.nrow + RPB - 1) / RPB, np);
  const int ps = prof_begin(c, "rows_inv");
  if (reg_ok(c, c->n1) && S.planes_contiguous) {
    switch (c->n1) {
#define AMDK_C(N, A, B) case N: reg_rows_inv<T, A, B>(c, S, C, ldc, nb, pair0, np); break;
      AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
    }
    AMDK_TRY_HIP(hipGetLastError());
    prof_end(c, ps);
    return;
  }
  k_rows_inv<T, RPB><<<grid, 256, lds, c->stream>>>(geom_of(c, S), (const double2*)C, ldc, nb, pair0, (V*)c->W1, c->p1,
                                                     (const V*)(sizeof(T) == 4 ? c->tw_f[0] : c->tw_d[0]));
  AMDK_TRY_HIP(hipGetLastError());
  prof_end(c, ps);
}

template <typename T>
static void launch_cols(amdk_ctx* c, const Sphere& S, int np, bool inverse) {
  using V = typename cplx_of<T>::type;
  constexpr int TX = Tiles<T>::TX;
  const size_t lds = (size_t)(c->n2 + 2 * c->n2 * TX) * sizeof(V);
  dim3 grid((c->n1 + TX - 1) / TX, S.nplane, np);
  const V* tw = (const V*)(sizeof(T) == 4 ? c->tw_f[1] : c->tw_d[1]);
  if (inverse) {
    const int ps = prof_begin(c, "cols_inv");
    if (reg_ok(c, c->n2) && S.planes_contiguous) {
      switch (c->n2) {
#define AMDK_C(N, A, B) case N: reg_cols<T, A, B>(c, S, np, true); break;
        AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
      }
      AMDK_TRY_HIP(hipGetLastError());
      prof_end(c, ps);
      return;
    }
    k_cols_inv<T, TX><<<grid, 256, lds, c->stream>>>(geom_of(c, S), (const V*)c->W1, (V*)c->W2, c->p2, tw);
    AMDK_TRY_HIP(hipGetLastError());
    prof_end(c, ps);
  } else {
    const int ps = prof_begin(c, "cols_fwd");
    if (reg_ok(c, c->n2) && S.planes_contiguous) {
      switch (c->n2) {
#define AMDK_C(N, A, B) case N: reg_cols<T, A, B>(c, S, np, false); break;
        AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
      }
      AMDK_TRY_HIP(hipGetLastError());
      prof_end(c, ps);
      return;
    }
    k_cols_fwd<T, TX><<<grid, 256, lds, c->stream>>>(geom_of(c, S), (const V*)c->W2, (V*)c->W1, c->p2, tw);
    AMDK_TRY_HIP(hipGetLastError());
    prof_end(c, ps);
  }
}

template <typename T>
static void launch_z_vpsi(amdk_ctx* c, const Sphere& S, int np) {
  using V = typename cplx_of<T>::type;
  constexpr int TX = Tiles<T>::TX, PG = Tiles<T>::PG;
  const size_t lds = (size_t)(c->n3 + 2 * c->n3 * TX) * sizeof(V) + (size_t)c->n3 * TX * sizeof(T);
  dim3 grid((c->n1 + TX - 1) / TX, c->n2, (np + PG - 1) / PG);
  const int ps = prof_begin(c, "z_vpsi");
  if (reg_ok(c, c->n3)) {
    switch (c->n3) {
#define AMDK_C(N, A, B) case N: reg_z_vpsi<T, A, B>(c, S, np); break;
      AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
    }
    AMDK_TRY_HIP(hipGetLastError());
    prof_end(c, ps);
    return;
  }
  k_z_vpsi<T, TX><<<grid, 256, lds, c->stream>>>(geom_of(c, S), (V*)c->W2, (const T*)c->vr, np, PG, c->p3,
                                                 (const V*)(sizeof(T) == 4 ? c->tw_f[2] : c->tw_d[2]));
  AMDK_TRY_HIP(hipGetLastError());
  prof_end(c, ps);
}

template <typename T>
static void launch_rows_fwd_hpsi(amdk_ctx* c, const Sphere& S, const void* C, void* C2, int nb, int pair0, int np) {
  using V = typename cplx_of<T>::type;
  constexpr int RPB = Tiles<T>::RPB;
  const size_t lds = (size_t)(c->n1 + 4 * RPB * c->n1) * sizeof(V);
  dim3 grid((S.nrow + RPB - 1) / RPB, np);
  const int ps = prof_begin(c, "rows_fwd_hpsi");
  if (reg_ok(c, c->n1) && S.planes_contiguous) {
    switch (c->n1) {
#define AMDK_C(N, A, B) case N: reg_rows_fwd_hpsi<T, A, B>(c, S, C, C2, nb, pair0, np); break;
      AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
    }
    AMDK_TRY_HIP(hipGetLastError());
    prof_end(c, ps);
    return;
  }
  k_rows_fwd_hpsi<T, RPB><<<grid, 256, lds, c->stream>>>(geom_of(c, S), (const V*)c->W1, (const double2*)C, (double2*)C2, S.ldc, nb,
                                                          pair0, c->d_occ, c->rc, c->p1,
                                                          (const V*)(sizeof(T) == 4 ? c->tw_f[0] : c->tw_d[0]));
  AMDK_TRY_HIP(hipGetLastError());
  prof_end(c, ps);
}

template <typename T>
// groups > 1 (register path only): pair groups accumulate into `groups` grids
// spaced part_stride apart, starting at rho.
static void launch_z_rho(amdk_ctx* c, const Sphere& S, int np, double* rho, int accumulate, const double* wpair,
                         int groups = 1, size_t part_stride = 0) {
  using V = typename cplx_of<T>::type;
  constexpr int TX = Tiles<T>::TX;
  const size_t lds = (size_t)(c->n3 + 2 * c->n3 * TX) * sizeof(V);
  dim3 grid((c->n1 + TX - 1) / TX, c->n2);
  const int ps = prof_begin(c, "z_rho");
  if (reg_ok(c, c->n3)) {
    switch (c->n3) {
#define AMDK_C(N, A, B) case N: reg_z_rho<T, A, B>(c, S, np, rho, accumulate, wpair, groups, part_stride); break;
      AMDK_REG_SIZES(AMDK_C)
#undef AMDK_C
    }
    AMDK_TRY_HIP(hipGetLastError());
    prof_end(c, ps);
    return;
  }
  if (groups != 1) throw Error(AMDK_ERR_STATE, "grouped density pass needs the register FFT path");
  k_z_rho<T, TX, 16><<<grid, 256, lds, c->stream>>>(geom_of(c, S), (const V*)c->W2, rho, np, wpair, accumulate, c->p3,
                                                     (const V*)(sizeof(T) == 4 ? c->tw_f[2] : c->tw_d[2]));
  AMDK_TRY_HIP(hipGetLastError());
  prof_end(c, ps);
}

template <typename T>
// planes_ready: W2 already holds the y-transformed planes of all pairs of
// these bands (left by run_density of the same C0 in one chunk), so the x and
// y inverse passes are skipped.
static void run_vpsi(amdk_ctx* c, const void* C0, void* C2, int nb, bool planes_ready = false) {
  const Sphere& S = c->spw;
  const int npairs = (nb + 1) / 2;
  if (planes_ready && npairs > c->cap_pairs) throw Error(AMDK_ERR_STATE, "planes of all pairs do not fit the work arrays");
  for (int p0 = 0; p0 < npairs; p0 += c->cap_pairs) {
    const int np = std::min(c->cap_pairs, npairs - p0);
    if (!planes_ready) {
      launch_rows_inv<T>(c, S, C0, S.ldc, nb, p0, np);
      launch_cols<T>(c, S, np, true);
    }
    launch_z_vpsi<T>(c, S, np);
    launch_cols<T>(c, S, np, false);
    launch_rows_fwd_hpsi<T>(c, S, C0, C2, nb, p0, np);
  }
}

template <typename T>
static void renormalise_density(amdk_ctx* c, const void* C0, int nb, double* rho, bool norms_ready);

// Band weights of the density on the device: d_wpair[b] = f_b / Omega (pairs
// read two consecutive entries), d_wdens[b] = f_b.  Set once per occupation.
static void prepare_density_weights(amdk_ctx* c, const double* occ, int nb) {
  c->occ_state_loaded = false;
  std::vector<double> w(nb + 2, 0.0);
  for (int b = 0; b < nb; ++b) w[b] = occ[b] / c->omega;
  AMDK_TRY_HIP(hipMemcpyAsync(c->d_wpair, w.data(), sizeof(double) * (nb + 1), hipMemcpyHostToDevice, c->stream));
  AMDK_TRY_HIP(hipMemcpyAsync(c->d_wdens, occ, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
  AMDK_TRY_HIP(hipStreamSynchronize(c->stream));
}

// rho(r) for bands [0, nb) with the weights of prepare_density_weights
// Density of the pairs of C0 into rho.  Into grid[0] with many pairs, the z
// pass runs in RHO_GROUPS pair groups (more blocks in flight) that accumulate
// into grid[1 ..] and are summed into rho by k_rho_combine, which also leaves
// the grid-sum partials in slot 8.
static constexpr int RHO_GROUPS = 4;
template <typename T>
// norms_after: event after which slot 10 holds sum_b f_b <c_b|c_b> (side stream)
static void run_density(amdk_ctx* c, const void* C0, int nb, double* rho, bool norms_ready = false,
                        hipEvent_t norms_after = nullptr) {
  const Sphere& S = c->spw;
  const int npairs = (nb + 1) / 2;
  const size_t nr = (size_t)c->n1 * c->n2 * c->n3;
  const bool grouped = rho == c->grid[0] && reg_ok(c, c->n3) && std::min(npairs, c->cap_pairs) >= 2 * RHO_GROUPS;
  for (int p0 = 0; p0 < npairs; p0 += c->cap_pairs) {
    const int np = std::min(c->cap_pairs, npairs - p0);
    launch_rows_inv<T>(c, S, C0, S.ldc, nb, p0, np);
    launch_cols<T>(c, S, np, true);
    if (grouped) launch_z_rho<T>(c, S, np, c->grid[1], p0 > 0, c->d_wpair + 2 * p0, RHO_GROUPS, nr);
    else launch_z_rho<T>(c, S, np, rho, p0 > 0, c->d_wpair + 2 * p0);
  }
  if (grouped) {
    k_rho_combine<RHO_GROUPS><<<RED_BLOCKS, 256, 0, c->stream>>>(c->grid[1], nr, nr, rho, c->d_part + 8 * RED_BLOCKS);
    AMDK_TRY_HIP(hipGetLastError());
  }
  if (norms_after) AMDK_TRY_HIP(hipStreamWaitEvent(c->stream, norms_after, 0));
  if (sizeof(T) == 4) renormalise_density<T>(c, C0, nb, rho, norms_ready, grouped);
}

// ---------------------------------------------------------------------------
#define API_BEGIN \
  if (!ctx) return AMDK_ERR_ARG; \
  try {
#define API_END                                  \
  }                                              \
  catch (const amdk::Error& e) {                 \
    ctx->err = e.what();                         \
    return e.code;                               \
  }                                              \
  catch (const std::exception& e) {              \
    ctx->err = e.what();                         \
    return AMDK_ERR_DEVICE;                      \
  }                                              \
  return AMDK_OK;

static void check_nb(amdk_ctx* c, int nb) {
  if (nb < 1 || nb > c->nbmax) throw Error(AMDK_ERR_ARG, "band count outside 1..nbmax");
}

static void check_finite(const double* p, size_t n, const char* what) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) throw Error(AMDK_ERR_NONFINITE, std::string("non-finite value in ") + what);
}

static void alloc_scf_buffers(amdk_ctx* c);
static void destroy_scf(amdk_ctx* c);
static double* noalloc_scratch(int dev);
static void grid_in(amdk_ctx* c, const double* host, double* dev);
static void grid_out(amdk_ctx* c, const double* dev, double* host);
static void check_finite_grid(const amdk_ctx* c, const double* host, const char* what);

extern "C" {

int amdk_ctx_create(const amdk_desc* d, amdk_ctx** out) {
  if (!out) return AMDK_ERR_ARG;
  *out = nullptr;
  amdk_ctx* ctx = new amdk_ctx();
  int rc = AMDK_OK;
  try {
    if (!d) throw Error(AMDK_ERR_ARG, "null descriptor");
    if (d->precision != AMDK_PREC_FP64 && d->precision != AMDK_PREC_MIXED && d->precision != AMDK_PREC_BF16)
      throw Error(AMDK_ERR_ARG, "bad precision mode");
    if (d->nbmax < 1) throw Error(AMDK_ERR_ARG, "nbmax must be >= 1");
    if (d->n1 > 192 || d->n2 > 192 || d->n3 > 192) throw Error(AMDK_ERR_ARG, "grid dimension above 192 is not supported");
    ctx->n1 = d->n1; ctx->n2 = d->n2; ctx->n3 = d->n3;
    ctx->p1 = make_plan(d->n1);
    ctx->p2 = make_plan(d->n2);
    ctx->p3 = make_plan(d->n3);
    ctx->prec = d->precision;
    ctx->nbmax = d->nbmax;
    {
      const char* f = std::getenv("AMDK_FFT");
      ctx->fft_generic = f && std::string(f) == "stockham";
      const char* o = std::getenv("AMDK_OVL_BIG");
      ctx->ovl_big = !(o && std::string(o) == "0");
    }
    std::memcpy(ctx->acell, d->a_cell, sizeof(ctx->acell));
    // reciprocal: g_j = 2 pi * row j of inv(A), A columns = a_j
    const double* a = d->a_cell;  // a[3*j+i] = (a_j)_i
    auto A = [&](int i, int j) { return a[3 * j + i]; };
    const double det = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                       A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    if (!(std::fabs(det) > 1e-12)) throw Error(AMDK_ERR_ARG, "singular cell");
    ctx->omega = std::fabs(det);
    double inv[3][3];
    inv[0][0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) / det;
    inv[0][1] = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) / det;
    inv[0][2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) / det;
    inv[1][0] = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) / det;
    inv[1][1] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) / det;
    inv[1][2] = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) / det;
    inv[2][0] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) / det;
    inv[2][1] = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) / det;
    inv[2][2] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) / det;
    const double tpi = 2.0 * M_PI;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) ctx->rc.g[3 * j + i] = tpi * inv[j][i];
    // minimum image of E_SR
    {
      const bool cubic = std::fabs(A(0, 1)) + std::fabs(A(0, 2)) + std::fabs(A(1, 0)) + std::fabs(A(1, 2)) +
                                 std::fabs(A(2, 0)) + std::fabs(A(2, 1)) < 1e-12 &&
                         std::fabs(A(0, 0) - A(1, 1)) < 1e-12 && std::fabs(A(0, 0) - A(2, 2)) < 1e-12;
      if (d->symmetry < 0) throw Error(AMDK_ERR_ARG, "bad lattice symmetry code");
      PbcInfo& pb = ctx->pbc;
      for (int k = 0; k < 9; ++k) pb.a[k] = a[k];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) pb.ainv[3 * j + i] = inv[j][i];
      if (d->symmetry == 2) {
        pb.mode = 2;  // conventional constant from |a_1| = A / sqrt2
        pb.A = std::sqrt(2.0 * (A(0, 0) * A(0, 0) + A(1, 0) * A(1, 0) + A(2, 0) * A(2, 0)));
      } else if ((d->symmetry == 0 || d->symmetry == 1) && cubic) {
        pb.mode = 1;
        pb.A = A(0, 0);
      } else {
        pb.mode = 3;
        pb.A = 0.0;
      }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error(AMDK_ERR_DEVICE, "no HIP device available");
    AMDK_TRY_HIP(hipGetDevice(&ctx->device));
    ctx->noalloc_scratch = noalloc_scratch(ctx->device);
    AMDK_TRY_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    AMDK_TRY_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    for (hipEvent_t* e : {&ctx->ev_fork, &ctx->ev_nl})
      AMDK_TRY_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    build_sphere(ctx, d->miller_pw, d->ngpw, ctx->spw);
    build_sphere(ctx, d->miller_rho, d->ngrho, ctx->srho);
    // twiddles
    const int ns[3] = {d->n1, d->n2, d->n3};
    for (int ax = 0; ax < 3; ++ax) {
      const int n = ns[ax];
      std::vector<double2> td(n);
      std::vector<float2> tf(n);
      for (int t = 0; t < n; ++t) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * t / n;
        td[t].x = (double)cosl(ang);
        td[t].y = (double)sinl(ang);
        tf[t].x = (float)td[t].x;
        tf[t].y = (float)td[t].y;
      }
      ctx->tw_d[ax] = dev_alloc(ctx, sizeof(double2) * n);
      ctx->tw_f[ax] = dev_alloc(ctx, sizeof(float2) * n);
      upload(ctx, ctx->tw_d[ax], td.data(), sizeof(double2) * n);
      upload(ctx, ctx->tw_f[ax], tf.data(), sizeof(float2) * n);
    }
    const size_t nr = (size_t)d->n1 * d->n2 * d->n3;
    ctx->vr = dev_alloc(ctx, (ctx->prec == AMDK_PREC_FP64 ? sizeof(double) : sizeof(float)) * nr);
    // FFT work arrays for up to cap_pairs band pairs at once (W1 + W2 within
    // 24 GiB of the 288 GiB of HBM3E; AMDK_PAIR_BUDGET_MB overrides it, e.g.
    // to exercise the chunked path at small sizes)
    const Sphere& S = ctx->spw;
    const size_t w2_pair = (size_t)S.nplane * d->n2 * row_pitch(d->n1) * cbytes(ctx);
    const size_t w1_pair = (size_t)S.nslot * row_pitch(d->n1) * cbytes(ctx);
    const int npairs = (d->nbmax + 1) / 2;
    size_t budget = (size_t)24 << 30;
    if (const char* b = std::getenv("AMDK_PAIR_BUDGET_MB")) {
      const long mb = std::atol(b);
      if (mb > 0) budget = (size_t)mb << 20;
    }
    ctx->cap_pairs = (int)std::max<size_t>(1, std::min<size_t>(npairs, budget / (w2_pair + w1_pair)));
    ctx->W1 = dev_alloc(ctx, w1_pair * ctx->cap_pairs);
    ctx->W2 = dev_alloc(ctx, w2_pair * ctx->cap_pairs);
    ctx->d_occ = (double*)dev_alloc(ctx, sizeof(double) * (d->nbmax + 1));
    ctx->d_wpair = (double*)dev_alloc(ctx, sizeof(double) * (d->nbmax + 2));
    ctx->d_wdens = (double*)dev_alloc(ctx, sizeof(double) * (d->nbmax + 1));
    alloc_scf_buffers(ctx);
    AMDK_TRY_HIP(hipStreamSynchronize(ctx->stream));
  } catch (const amdk::Error& e) {
    ctx->err = e.what();
    rc = e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    rc = AMDK_ERR_DEVICE;
  }
  *out = ctx;  // returned even on failure so that amdk_last_error works
  return rc;
}

void amdk_ctx_destroy(amdk_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  destroy_scf(ctx);
  for (void* p : ctx->owned) (void)hipFree(p);
  for (auto& s : ctx->pending) { (void)hipEventDestroy(s.a); (void)hipEventDestroy(s.b); }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : {ctx->ev_fork, ctx->ev_nl})
    if (e) (void)hipEventDestroy(e);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* amdk_last_error(const amdk_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int amdk_ldc(const amdk_ctx* ctx) { return ctx ? ctx->spw.ldc : -1; }

int amdk_pw_order(const amdk_ctx* ctx, int* pos, int* conj) {
  if (!ctx || !pos || !conj) return AMDK_ERR_ARG;
  for (int i = 0; i < ctx->spw.ncaller; ++i) {
    pos[i] = ctx->spw.h_from_caller[i];
    conj[i] = ctx->spw.h_conj_caller[i];
  }
  return AMDK_OK;
}

int amdk_set_vks(amdk_ctx* ctx, const double* v_r) {
  API_BEGIN
  if (!v_r) throw Error(AMDK_ERR_ARG, "null v_r");
  const size_t nr = (size_t)ctx->n1 * ctx->n2 * ctx->n3;
  check_finite_grid(ctx, v_r, "v_r");
  double* stage = ctx->grid[1];  // dense scratch grid, free between calls
  grid_in(ctx, v_r, stage);
  const double s = 1.0 / (double)nr;  // forward FFT normalisation folded into V
  if (ctx->prec == AMDK_PREC_FP64)
    k_scale_copy<double><<<1024, 256, 0, ctx->stream>>>(stage, (double*)ctx->vr, nr, s);
  else
    k_scale_copy<float><<<1024, 256, 0, ctx->stream>>>(stage, (float*)ctx->vr, nr, s);
  AMDK_TRY_HIP(hipGetLastError());
  AMDK_TRY_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_vks = true;
  API_END
}

int amdk_vpsi_dev(amdk_ctx* ctx, const void* c0, void* c2, const double* occ, int nb) {
  API_BEGIN
  caller_ready();
  check_nb(ctx, nb);
  if (!c0 || !c2 || !occ) throw Error(AMDK_ERR_ARG, "null pointer");
  if (!ctx->have_vks) throw Error(AMDK_ERR_STATE, "V_KS not set (amdk_set_vks)");
  AMDK_TRY_HIP(hipMemcpyAsync(ctx->d_occ, occ, sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
  ctx->occ_state_loaded = false;  // d_occ now holds this call's occupations
  if (ctx->prec == AMDK_PREC_FP64) run_vpsi<double>(ctx, c0, c2, nb);
  else run_vpsi<float>(ctx, c0, c2, nb);
  sync(ctx);
  API_END
}

int amdk_vpsi(amdk_ctx* ctx, const double* c0, int ldc, double* c2, const double* occ, int nb) {
  API_BEGIN
  check_nb(ctx, nb);
  if (!c0 || !c2 || !occ) throw Error(AMDK_ERR_ARG, "null pointer");
  if (ldc < ctx->spw.npw) throw Error(AMDK_ERR_ARG, "ldc < ngpw");
  if (!ctx->have_vks) throw Error(AMDK_ERR_STATE, "V_KS not set (amdk_set_vks)");
  check_finite(c0, (size_t)2 * ldc * nb, "c0");
  // band scratch of the context (C2 and Ct carry no state between calls)
  void* dc0 = ctx->Ct;
  void* dc2 = ctx->C2;
  bands_in(ctx, c0, ldc, nb, dc0);
  AMDK_TRY_HIP(hipMemcpyAsync(ctx->d_occ, occ, sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
  ctx->occ_state_loaded = false;  // d_occ now holds this call's occupations
  if (ctx->prec == AMDK_PREC_FP64) run_vpsi<double>(ctx, dc0, dc2, nb);
  else run_vpsi<float>(ctx, dc0, dc2, nb);
  sync(ctx);
  bands_out(ctx, dc2, nb, c2, ldc);
  API_END
}

int amdk_density_dev(amdk_ctx* ctx, const void* c0, const double* occ, int nb, double* rho_r_dev) {
  API_BEGIN
  caller_ready();
  check_nb(ctx, nb);
  if (!c0 || !occ || !rho_r_dev) throw Error(AMDK_ERR_ARG, "null pointer");
  prepare_density_weights(ctx, occ, nb);
  if (ctx->prec == AMDK_PREC_FP64) run_density<double>(ctx, c0, nb, rho_r_dev);
  else run_density<float>(ctx, c0, nb, rho_r_dev);
  sync(ctx);
  API_END
}

int amdk_density(amdk_ctx* ctx, const double* c0, int ldc, const double* occ, int nb, double* rho_r) {
  API_BEGIN
  check_nb(ctx, nb);
  if (!c0 || !occ || !rho_r) throw Error(AMDK_ERR_ARG, "null pointer");
  if (ldc < ctx->spw.npw) throw Error(AMDK_ERR_ARG, "ldc < ngpw");
  check_finite(c0, (size_t)2 * ldc * nb, "c0");
  const size_t nr = (size_t)ctx->n1 * ctx->n2 * ctx->n3;
  void* dc0 = ctx->Ct;          // band scratch
  double* drho = ctx->grid[0];  // density grid (rewritten before every use)
  bands_in(ctx, c0, ldc, nb, dc0);
  prepare_density_weights(ctx, occ, nb);
  if (ctx->prec == AMDK_PREC_FP64) run_density<double>(ctx, dc0, nb, drho);
  else run_density<float>(ctx, dc0, nb, drho);
  sync(ctx);
  grid_out(ctx, drho, rho_r);
  API_END
}

int amdk_bands_to_dev(amdk_ctx* ctx, const double* c, int ldc, int nb, void* dst) {
  API_BEGIN
  check_nb(ctx, nb);
  if (!c || !dst) throw Error(AMDK_ERR_ARG, "null pointer");
  if (ldc < ctx->spw.npw) throw Error(AMDK_ERR_ARG, "ldc < ngpw");
  bands_in(ctx, c, ldc, nb, dst);
  API_END
}

int amdk_bands_from_dev(amdk_ctx* ctx, const void* src, int nb, double* c, int ldc) {
  API_BEGIN
  caller_ready();
  check_nb(ctx, nb);
  if (!c || !src) throw Error(AMDK_ERR_ARG, "null pointer");
  if (ldc < ctx->spw.npw) throw Error(AMDK_ERR_ARG, "ldc < ngpw");
  bands_out(ctx, src, nb, c, ldc);
  API_END
}

int amdk_profile_enable(amdk_ctx* ctx, int on) {
  API_BEGIN
  ctx->prof = on != 0;
  API_END
}

int amdk_profile_query(amdk_ctx* ctx, const char* kernel, double* total_ms, int* count) {
  API_BEGIN
  if (!kernel || !total_ms || !count) throw Error(AMDK_ERR_ARG, "null pointer");
  auto it = ctx->prof_acc.find(kernel);
  *total_ms = it == ctx->prof_acc.end() ? 0.0 : it->second.first;
  *count = it == ctx->prof_acc.end() ? 0 : it->second.second;
  API_END
}

int amdk_sphere_info(const amdk_ctx* ctx, int which, int* out6) {
  if (!ctx || !out6 || (which != 0 && which != 1)) return AMDK_ERR_ARG;
  const Sphere& S = which == 0 ? ctx->spw : ctx->srho;
  out6[0] = S.npw;
  out6[1] = S.ldc;
  out6[2] = S.nrow;
  out6[3] = S.nslot;
  out6[4] = S.nplane;
  out6[5] = S.R;
  return AMDK_OK;
}

int amdk_profile_names(amdk_ctx* ctx, char* buf, int len) {
  API_BEGIN
  if (!buf || len < 1) throw Error(AMDK_ERR_ARG, "null buffer");
  std::string all;
  for (const auto& kv : ctx->prof_acc) all += kv.first + "\n";
  if ((int)all.size() + 1 > len) throw Error(AMDK_ERR_ARG, "buffer too small");
  std::memcpy(buf, all.c_str(), all.size() + 1);
  API_END
}

int amdk_profile_reset(amdk_ctx* ctx) {
  API_BEGIN
  ctx->prof_acc.clear();
  API_END
}

}  // extern "C"

#include "scf_api.inc"

// ---------------------------------------------------------------------------
// Allocation-free OVLAP_D.  Its split-K partials live in `scratch`, a fixed
// array of the code object: every GPU that loads the library holds one copy
// (32 MiB) until the process ends.  In host code `scratch` names a host-side
// stand-in, so kernels get the device address that hipGetSymbolAddress
// returns, looked up once per device (at the first amdk_ctx_create with that
// device current) and kept in g_scratch_ptr.
#define AMDK_NOALLOC_SCRATCH_DOUBLES (4u << 20)
static __device__ __align__(256) double scratch[AMDK_NOALLOC_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {};

static double* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= 64) throw Error(AMDK_ERR_DEVICE, "device index outside 0..63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(AMDK_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  return g_scratch_ptr[dev];
}

extern "C" int amdk_overlap_noalloc(amdk_ctx* ctx, const double* ca, const double* cb, int ldc, int n1, int n2,
                                    double* A) {
  API_BEGIN
  check_nb(ctx, n1);
  check_nb(ctx, n2);
  if (!ca || !cb || !A) throw Error(AMDK_ERR_ARG, "null pointer");
  if ((size_t)n1 * n2 > AMDK_NOALLOC_SCRATCH_DOUBLES)
    throw Error(AMDK_ERR_ARG, "n1*n2 = " + std::to_string((size_t)n1 * n2) + " exceeds the allocation-free limit of " +
                                  std::to_string(AMDK_NOALLOC_SCRATCH_DOUBLES) + " (use amdk_overlap)");
  stage_in(ctx, ca, ldc, n1, ctx->Ct);
  stage_in(ctx, cb, ldc, n2, ctx->C2);
  const Sphere& S = ctx->spw;
  // every split writes all n1*n2 partials it owns before k_ovl_reduce reads
  // them: nothing in the array has to start at zero
  ovl_into<double2>(ctx, (const double2*)ctx->Ct, S.ldc, n1, (const double2*)ctx->C2, S.ldc, n2, S.npw, 1.0, ctx->d_S,
                    n1, nullptr, 0.0, nullptr, false, ctx->noalloc_scratch, AMDK_NOALLOC_SCRATCH_DOUBLES);
  AMDK_TRY_HIP(hipMemcpyAsync(A, ctx->d_S, sizeof(double) * n1 * n2, hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  API_END
}
