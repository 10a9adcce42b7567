// Internal declarations shared by the HIP translation units of libamdk.
//
// Layout conventions (device side):
//  * A "sphere" is a set of half-sphere plane waves G = m1*b1 + m2*b2 + m3*b3
//    stored in z-half order: (m3 > 0) or (m3 == 0 && m2 > 0) or
//    (m3 == 0 && m2 == 0 && m1 >= 0), sorted by (m3, m2, m1).  Each fixed
//    (m2, m3) forms a contiguous "half row" of consecutive m1 values.  The
//    missing half is the complex conjugate (real fields, Gamma point).
//  * Complex arrays are interleaved (re, im) pairs; band b of a coefficient
//    array starts at b * ldc elements.
//  * Real-space grids are stored x-fastest: r = x + n1*(y + n2*z).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#define AMDK_MAX_STAGES 10

namespace amdk {

// Factorisation of one FFT length into radix 2/3/4/5 stages.
struct FftPlan1D {
  int n;
  int nst;
  int rad[AMDK_MAX_STAGES];
};

FftPlan1D make_plan(int n);  // host; throws std::runtime_error on bad n

// Per half-row metadata (device): packed to keep loads few.
struct RowInfo {
  int start;   // offset of the first element in the sphere array
  int cnt;     // number of consecutive m1 values
  int m1_0;    // m1 of the first element
  int m2, m3;  // row coordinates
  int slot_d;  // slot of the direct row (m2, m3) in the row buffer
  int slot_m;  // slot of the mirrored row (-m2, -m3); -1 for the (0,0) row
  int pad;
};

// One sphere (psi or rho) as seen by the FFT kernels.
struct Sphere {
  int npw = 0;        // half-sphere plane waves
  int ldc = 0;        // padded leading dimension (elements) used for bands
  int nrow = 0;       // half rows
  int nslot = 0;      // full rows (direct + mirrored)
  int R = 0;          // max |m3|; planes are m3 = -R..R
  int nplane = 0;     // 2R+1
  RowInfo* rows = nullptr;      // device [nrow]
  int* slot_m2 = nullptr;       // device [nslot]
  int* plane_off = nullptr;     // device [nplane+1], slots sorted by (m3, m2)
  std::vector<RowInfo> h_rows;  // host copies
  std::vector<int> h_slot_m2, h_plane_off;
  std::vector<int> h_mill;      // host [3*npw] internal order
  // caller order -> internal: idx[i] (>=0), conj flag sgn[i] (+1 direct, -1 conj)
  std::vector<int> h_from_caller, h_conj_caller;
  int* d_from_caller = nullptr;  // device [ncaller]
  int* d_conj_caller = nullptr;  // device [ncaller]
  int ncaller = 0;
  bool planes_contiguous = false;  // each plane's rows are consecutive m2
};

}  // namespace amdk
