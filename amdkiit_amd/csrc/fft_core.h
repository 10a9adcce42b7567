// LDS-resident batched 1D FFT used by every pass of the sphere-sparse 3D FFT.
//
// Mixed-radix (2, 3, 4, 5) Stockham autosort, ping-pong between two LDS
// buffers.  Element i of transform b lives at buf[b*bstr + i*istr].  When
// bfast is set, consecutive threads take consecutive transforms (column
// tiles, where transforms are adjacent in memory); otherwise consecutive
// threads walk along one transform (row tiles).
//
// Sign convention: DIR = -1 computes y_k = sum_j x_j exp(-2 pi i jk/n),
// DIR = +1 the same with exp(+...).  No scaling is applied.
#pragma once
#include "amdk_internal.h"

namespace amdk {

template <typename T> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by s*i, s = +-1
template <int S, typename V>
__device__ __forceinline__ V mul_si(V a) { V r; r.x = -S * a.y; r.y = S * a.x; return r; }

template <typename V, int R, int DIR> struct Butterfly;

template <typename V, int DIR> struct Butterfly<V, 2, DIR> {
  __device__ static __forceinline__ void run(V* v) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename V, int DIR> struct Butterfly<V, 3, DIR> {
  __device__ static __forceinline__ void run(V* v) {
    using T = decltype(v[0].x);
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    V t = cadd(v[1], v[2]);
    V d = csub(v[1], v[2]);
    V a0 = v[0];
    V m;  // a0 - t/2
    m.x = a0.x - (T)0.5 * t.x;
    m.y = a0.y - (T)0.5 * t.y;
    V u = mul_si<DIR>(d);  // s*i*(a1-a2)
    u.x *= h;
    u.y *= h;
    v[0] = cadd(a0, t);
    v[1] = cadd(m, u);
    v[2] = csub(m, u);
  }
};

template <typename V, int DIR> struct Butterfly<V, 4, DIR> {
  __device__ static __forceinline__ void run(V* v) {
    V t0 = cadd(v[0], v[2]);
    V t1 = csub(v[0], v[2]);
    V t2 = cadd(v[1], v[3]);
    V t3 = mul_si<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <typename V, int DIR> struct Butterfly<V, 5, DIR> {
  __device__ static __forceinline__ void run(V* v) {
    using T = decltype(v[0].x);
    const T c1 = (T)0.30901699437494742410;   // cos(2pi/5)
    const T c2 = (T)-0.80901699437494742410;  // cos(4pi/5)
    const T s1 = (T)0.95105651629515357212;   // sin(2pi/5)
    const T s2 = (T)0.58778525229247312917;   // sin(4pi/5)
    V a0 = v[0];
    V t1 = cadd(v[1], v[4]);
    V t2 = cadd(v[2], v[3]);
    V t3 = csub(v[1], v[4]);
    V t4 = csub(v[2], v[3]);
    V b1, b2, e1, e2;
    b1.x = a0.x + c1 * t1.x + c2 * t2.x;
    b1.y = a0.y + c1 * t1.y + c2 * t2.y;
    b2.x = a0.x + c2 * t1.x + c1 * t2.x;
    b2.y = a0.y + c2 * t1.y + c1 * t2.y;
    e1.x = s1 * t3.x + s2 * t4.x;
    e1.y = s1 * t3.y + s2 * t4.y;
    e2.x = s2 * t3.x - s1 * t4.x;
    e2.y = s2 * t3.y - s1 * t4.y;
    V d1 = mul_si<DIR>(e1);
    V d2 = mul_si<DIR>(e2);
    v[0].x = a0.x + t1.x + t2.x;
    v[0].y = a0.y + t1.y + t2.y;
    v[1] = cadd(b1, d1);
    v[4] = csub(b1, d1);
    v[2] = cadd(b2, d2);
    v[3] = csub(b2, d2);
  }
};

// One Stockham stage of radix R over `batch` transforms of length n.
// tw[t] = exp(-2 pi i t / n), t in [0, n).
template <typename V, int R, int DIR>
__device__ __forceinline__ void fft_stage(const V* __restrict__ src, V* __restrict__ dst,
                                          const V* __restrict__ tw, int n, int Ns, int batch,
                                          int bstr, int istr, bool bfast) {
  const int m = n / R;
  const int total = batch * m;
  const int tstep = n / (Ns * R);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    int b, j;
    if (bfast) {
      b = t % batch;
      j = t / batch;
    } else {
      j = t % m;
      b = t / m;
    }
    const int k = j % Ns;
    V v[R];
    const V* s = src + b * bstr;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[(j + r * m) * istr];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V w = tw[k * r * tstep];
        if (DIR > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Butterfly<V, R, DIR>::run(v);
    const int o = (j / Ns) * Ns * R + k;
    V* d = dst + b * bstr;
#pragma unroll
    for (int r = 0; r < R; ++r) d[(o + r * Ns) * istr] = v[r];
  }
}

// Full transform.  Caller must __syncthreads() after filling `a`.  Returns the
// buffer (a or b) holding the result; a barrier has been passed on return.
template <typename V, int DIR>
__device__ V* lds_fft(V* a, V* b, const V* tw, const FftPlan1D& p, int batch, int bstr,
                      int istr, bool bfast) {
  V* src = a;
  V* dst = b;
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.rad[s];
    switch (R) {
      case 2: fft_stage<V, 2, DIR>(src, dst, tw, p.n, Ns, batch, bstr, istr, bfast); break;
      case 3: fft_stage<V, 3, DIR>(src, dst, tw, p.n, Ns, batch, bstr, istr, bfast); break;
      case 4: fft_stage<V, 4, DIR>(src, dst, tw, p.n, Ns, batch, bstr, istr, bfast); break;
      default: fft_stage<V, 5, DIR>(src, dst, tw, p.n, Ns, batch, bstr, istr, bfast); break;
    }
    __syncthreads();
    Ns *= R;
    V* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

}  // namespace amdk
