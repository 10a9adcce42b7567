// Context object behind the C ABI and the helpers every entry point uses.
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/amdk.h"
#include "amdk_internal.h"
#include "fft_passes.h"

#define AMDK_TRY_HIP(expr)                                                         \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      throw amdk::Error(AMDK_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                              \
  } while (0)

namespace amdk {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Device buffer owned by the context (freed in the destructor).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct ProfSlot {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace amdk

struct amdk_ctx {
  int n1 = 0, n2 = 0, n3 = 0;
  double acell[9];
  double omega = 0;
  amdk::Recip rc;
  int prec = AMDK_PREC_FP64;
  int nbmax = 0;
  hipStream_t stream = nullptr;
  amdk::Sphere spw, srho;
  amdk::FftPlan1D p1, p2, p3;
  // twiddles exp(-2 pi i t/n) in float2 and double2
  void* tw_f[3] = {nullptr, nullptr, nullptr};
  void* tw_d[3] = {nullptr, nullptr, nullptr};
  // V_KS(r)/N in the compute precision; set flag
  void* vr = nullptr;
  bool have_vks = false;
  // FFT workspaces sized for `cap_pairs` band pairs
  void* W1 = nullptr;
  void* W2 = nullptr;
  int cap_pairs = 0;
  double* d_occ = nullptr;   // [nbmax]
  double* d_wpair = nullptr; // [nbmax+1]
  // allocations to free
  std::vector<void*> owned;
  std::string err;
  // profiling
  bool prof = false;
  std::vector<amdk::ProfSlot> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, int>> prof_acc;
};

namespace amdk {

void* dev_alloc(amdk_ctx* c, size_t bytes);  // zero-filled, owned by ctx
void sync(amdk_ctx* c);                      // stream sync + collect profile
int prof_begin(amdk_ctx* c, const char* name);
void prof_end(amdk_ctx* c, int slot);
size_t cbytes(const amdk_ctx* c);            // sizeof(complex) in compute precision

}  // namespace amdk
