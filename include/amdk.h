/*
 * amdk.h — C ABI of libamdk, the MI355X (gfx950) Kohn–Sham wavefunction
 * hot path of AMDKIIT (Gamma point, norm-conserving, plane waves).
 *
 * Every entry point replaces one reference routine and keeps its argument
 * meaning; the reference interface is cited beside each declaration
 * (paths relative to the AMDKIIT source tree).  The Fortran binding that
 * the reference call sites use is amdkiit_amd/fortran/amdk_hip.f90.
 *
 * Conventions
 *  - Complex coefficient arrays are interleaved (re, im) doubles, band-major
 *    with leading dimension ldc >= ngpw complex elements: element (g, b)
 *    is at [2*(b*ldc + g)], i.e. Fortran COMPLEX*16 C(LDC, NB).
 *  - Plane-wave order is the caller's: the Miller indices passed to
 *    amdk_ctx_create fix it (any half sphere; G and -G must not both appear).
 *  - Real-space grids are dense n1*n2*n3 doubles, x fastest (no padding).
 *  - Return value 0 = success; otherwise an AMDK_ERR_* code and
 *    amdk_last_error() describes it.  The Fortran shims turn a non-zero
 *    status into STOP, as the reference does on LAPACK failures.
 *  - Calls are stream-ordered on the context's HIP stream and synchronous
 *    on return.  One host thread per context.
 *  - *_dev entry points take device pointers in the context's internal
 *    layout (plane-wave order of amdk_pw_order, leading dimension
 *    amdk_ldc(), element type float2 in AMDK_PREC_MIXED, double2 in
 *    AMDK_PREC_FP64) so that wavefunctions stay resident in HBM.
 */
#ifndef AMDK_H
#define AMDK_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AMDK_OK = 0,
  AMDK_ERR_ARG = 1,        /* bad argument / size mismatch */
  AMDK_ERR_DEVICE = 2,     /* no device or HIP runtime failure */
  AMDK_ERR_OOM = 3,        /* device allocation failed */
  AMDK_ERR_NOT_PD = 4,     /* overlap matrix not positive definite (UINV) */
  AMDK_ERR_NONFINITE = 5,  /* NaN/Inf found in an input */
  AMDK_ERR_STATE = 6       /* required table (V_KS, projectors, positions) not set */
};

enum { AMDK_PREC_FP64 = 0, AMDK_PREC_MIXED = 1 };

typedef struct amdk_ctx amdk_ctx;

typedef struct {
  int n1, n2, n3;          /* FFT grid (nrgrids, set_cell.f90:177-212) */
  double a_cell[9];        /* lattice vectors a1,a2,a3 as columns, bohr (a_cell(3,3)) */
  int ngpw;                /* wavefunction half-sphere size (ngpw) */
  const int* miller_pw;    /* [3*ngpw] (m1,m2,m3) per G, caller order (inyh(:,1:ngpw) - nh) */
  int ngrho;               /* density half-sphere size (ngrho) */
  const int* miller_rho;   /* [3*ngrho] */
  int nbmax;               /* largest band count that will be passed */
  int precision;           /* AMDK_PREC_FP64 or AMDK_PREC_MIXED */
} amdk_desc;

/* Context: geometry, maps, FFT plans, device workspaces.
 * Replaces initialize's set_reciprocal/prepare_fft (init_module.f90:105-107,
 * gvectors.f90:111-341, fft_interface.f90:151-363). */
int amdk_ctx_create(const amdk_desc* d, amdk_ctx** out);
void amdk_ctx_destroy(amdk_ctx* ctx);
const char* amdk_last_error(const amdk_ctx* ctx);
/* Internal leading dimension (complex elements) of *_dev band arrays. */
int amdk_ldc(const amdk_ctx* ctx);
/* Internal plane-wave order: for caller index i, pos[i] is the internal
 * index and conj[i] is 1 when the internal entry holds c(-G) = conj(c(G)). */
int amdk_pw_order(const amdk_ctx* ctx, int* pos, int* conj);

/* Local Kohn–Sham potential V_KS(r) (V_POT after cal_vpot_ex, xc.f90:4-434). */
int amdk_set_vks(amdk_ctx* ctx, const double* v_r);

/* H|psi> local + kinetic part: C2 = -(f/2)(|G|^2 c + 2 (V psi)(G)), f==0 -> 1.
 * Replaces VPSI(C_ATOM, C2_nl, FC, NOSTATE) (wfn_initialize.f90:268-365). */
int amdk_vpsi(amdk_ctx* ctx, const double* c0, int ldc, double* c2, const double* occ, int nb);
int amdk_vpsi_dev(amdk_ctx* ctx, const void* c0, void* c2, const double* occ, int nb);

/* rho(r) = sum_i f_i |psi_i(r)|^2 / Omega on the dense grid.
 * Replaces eval_density (density.f90:5-98), closed shell. */
int amdk_density(amdk_ctx* ctx, const double* c0, int ldc, const double* occ, int nb, double* rho_r);
int amdk_density_dev(amdk_ctx* ctx, const void* c0, const double* occ, int nb, double* rho_r_dev);

/* Transfer helpers between caller-layout FP64 host arrays and the internal
 * device layout (no reference counterpart: they replace nothing, they keep
 * C_0/C2 resident across an SCF step). */
int amdk_bands_to_dev(amdk_ctx* ctx, const double* c, int ldc, int nb, void* dst);
int amdk_bands_from_dev(amdk_ctx* ctx, const void* src, int nb, double* c, int ldc);

/* Per-kernel timing with HIP events on the context stream. */
int amdk_profile_enable(amdk_ctx* ctx, int on);
int amdk_profile_query(amdk_ctx* ctx, const char* kernel, double* total_ms, int* count);
int amdk_profile_reset(amdk_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
