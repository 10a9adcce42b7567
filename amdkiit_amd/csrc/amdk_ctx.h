// Context object behind the C ABI and the helpers every entry point uses.
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/amdk.h"
#include "amdk_internal.h"
#include "fft_passes.h"
#include "gemm.h"
#include "dense.h"
#include "gemm_mfma.h"

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
  amdk::PbcInfo pbc;          // minimum-image rule of E_SR (from amdk_desc.symmetry)
  int prec = AMDK_PREC_FP64;
  bool fft_generic = false;   // AMDK_FFT=stockham forces the LDS Stockham passes
  bool ovl_big = true;        // 128 x 128 overlap tiles where they fit (AMDK_OVL_BIG=0: 64 x 64 only)
  bool ortho_lowdin = false;  // SCF step orthonormalises by C S^-1/2 instead of Cholesky
  int nbmax = 0;
  hipStream_t stream = nullptr;
  int device = 0;                     // device current at amdk_ctx_create
  double* noalloc_scratch = nullptr;  // that device's copy of the library's fixed scratch array
  // second stream of the SCF step (work that only reads C0 runs beside the
  // density pass and V_KS) and the events that fork and join it
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_nl = nullptr;
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
  // internal Miller indices (3 ints per G) of both spheres
  int* d_mill_pw = nullptr;
  int* d_mill_rho = nullptr;
  // species tables (amdk_set_species)
  bool have_species = false;
  int nsp = 0, maxproj = 0;
  std::vector<int> h_nproj, h_proj_l;
  std::vector<double> h_dmat, h_zv;
  std::vector<double> h_wsg;  // [nsp][maxproj] PCG preconditioner weights (WSG of hess)
  int maxproj_caller = 0;     // maxproj of the caller's species descriptor (row length of its tables)
  double* d_vps = nullptr;    // [nsp][ldc_rho]
  double* d_rhos = nullptr;   // [nsp][ldc_rho]
  double* d_twnl = nullptr;   // [nsp][maxproj][ldc_pw]
  double* d_zv = nullptr;
  // geometry (amdk_set_positions)
  bool have_pos = false;
  int nat = 0, nproj_tot = 0;
  std::vector<int> h_sp;
  std::vector<double> h_tau;
  std::vector<int4> h_plist;  // (atom, species, k, l)
  int* d_sp = nullptr;
  int* d_order = nullptr;     // atom indices in species order (stable)
  int4* d_agrp = nullptr;     // atom groups: {first index into d_order, count, species, 0}
  int nagrp = 0;
  double* d_tau = nullptr;
  double2* d_tab = nullptr;   // per-atom axis phases
  int tabW = 0, hmax = 0;
  double2* d_sv = nullptr;    // eigrxvps on the rho sphere
  double2* d_sr = nullptr;    // eigrxrhos
  void* d_beta = nullptr;     // [nproj_tot][ldc_pw] compute precision
  void* d_dbeta = nullptr;
  int4* d_plist = nullptr;
  double* d_vpp = nullptr;    // PCG preconditioner K(G)^-1 [ldc_pw]
  double* d_g2pw = nullptr;   // |G|^2 on the wavefunction sphere [ldc_pw], 0 on the padding
  double* d_dmat = nullptr;   // [nsp][maxproj][maxproj]
  int* d_nproj = nullptr;     // [nsp]
  int* d_pstart = nullptr;    // [nat+1] first projector row of each atom
  std::vector<double> h_twnl_int, h_vps0;
  std::vector<int> h_pstart;
  double esr = 0, eself = 0;
  std::vector<double> h_fsr;
  double* d_gm = nullptr;     // [9] reciprocal rows for the phase tables
  double* d_fsr = nullptr;    // [3 nat] E_SR forces
  double* d_esr = nullptr;    // [nat] E_SR per-atom partials
  // dense FP64 workspaces (density sphere)
  double2* W1r = nullptr;
  double2* W2r = nullptr;
  double* grid[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // rho, va, gx, gy, gz
  double2* d_rhog = nullptr;  // rho(G)
  double2* d_vg = nullptr;    // V_KS(G)
  double2* d_fld0 = nullptr;  // 2 x ldc_rho
  double2* d_fld1 = nullptr;
  bool have_rhog = false;
  // d_rhog was formed by the last SCF step (not by a stateless amdk_vks)
  bool rhog_from_state = false;
  // reductions and band-space matrices
  double* d_part = nullptr;   // [8][RED_BLOCKS]
  double* d_npart = nullptr;  // per-block band-norm partials of the Lagrange update
  int npart_stride = 0;
  double* d_red = nullptr;    // [16]
  double* d_gpart = nullptr;  // split-K partials
  size_t gpart_cap = 0;
  double* d_S = nullptr;      // nbmax x nbmax
  double* d_Ut = nullptr;     // nbmax x nbmax: scratch of the blocked Cholesky inverse (block products)
  double* d_X = nullptr;      // nbmax x nbmax: U^-1 of the orthonormalisation
  double* d_nl = nullptr;     // P x nbmax
  double* d_Q = nullptr;      // max(P, nbmax) x nbmax
  double* d_dfnl = nullptr;   // 3 x P x nbmax: <d beta / d R_k | C> per direction
  double* d_finv = nullptr;   // 1/f for PCG
  double* d_wnorm = nullptr;  // 1 for occupied bands (norm weights)
  double* d_wdens = nullptr;  // f_b for the density norm
  double* d_pcg = nullptr;    // PCG scalars
  int* d_info = nullptr;      // directly after d_red[16]
  void* h_stage = nullptr;    // pinned host copy of d_red and d_info
  // resident wavefunction state for amdk_scf_step
  void* C0 = nullptr;
  void* C2 = nullptr;
  void* Hd = nullptr;
  void* Pd = nullptr;
  void* Ct = nullptr;
  void* d_stage = nullptr;  // caller-layout staging of host band arrays (ldc x nbmax)
  int nb_state = 0;
  std::vector<double> h_occ_state;
  // d_occ, d_wnorm, d_finv and d_wpair hold the resident state's occupations
  // (API calls with their own occupations overwrite them)
  bool occ_state_loaded = false;
  // caller layouts (amdk_set_layout): real-space grids with leading
  // dimensions ld1 x ld2 (0: dense n1 x n2), nl(1, nat, nghmax, nb) when
  // nl_nghmax > 0, force(3, NA_MAX, nsp) when force_na_max > 0
  int grid_ld1 = 0, grid_ld2 = 0;
  int nl_nghmax = 0;
  int force_na_max = 0;
  std::vector<int> h_idx_in_sp;  // index of each atom among the atoms of its species
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
size_t cbytes(const amdk_ctx* c);            // sizeof(complex) in FFT compute precision
size_t sbytes(const amdk_ctx* c);            // sizeof(complex) of band storage (FP64)

}  // namespace amdk
