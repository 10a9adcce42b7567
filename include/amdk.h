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
 *  - Real-space grids are n1*n2*n3 doubles, x fastest, dense unless
 *    amdk_set_layout gives the reference's padded leading dimensions.
 *  - Return value 0 = success; otherwise an AMDK_ERR_* code and
 *    amdk_last_error() describes it.  The Fortran shims turn a non-zero
 *    status into STOP, as the reference does on LAPACK failures.
 *  - Calls are stream-ordered on the context's HIP stream and synchronous
 *    on return.  One host thread per context.
 *  - *_dev entry points take device pointers in the context's internal
 *    layout (plane-wave order of amdk_pw_order, leading dimension
 *    amdk_ldc(), element type double2) so that wavefunctions stay resident
 *    in HBM.  They wait for the whole device before reading caller buffers,
 *    which may have been written on any stream.  Coefficients are stored in
 *    FP64 in both precision modes; AMDK_PREC_MIXED runs the FFTs, V(r) psi(r)
 *    and the projector and Lagrange-multiplier GEMMs in FP32 (FP64 sums of
 *    32-deep FP32 tiles), while the orthonormalisation (overlap, Cholesky,
 *    C U^-1), energies, forces' sums and the density-grid potential stay FP64.
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

/* AMDK_PREC_BF16: AMDK_PREC_MIXED with the projector and Lagrange-multiplier
 * products in BF16 with one refinement term: operands split into hi =
 * bf16(x) and lo = bf16(x - hi), products hi*hi' + hi*lo' + lo*hi' on the
 * BF16 matrix cores, FP32 sums over 32-deep tiles, FP64 across tiles. */
enum { AMDK_PREC_FP64 = 0, AMDK_PREC_MIXED = 1, AMDK_PREC_BF16 = 2 };

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
  int symmetry;            /* lattice code of set_vec (set_cell.f90:17-42: 1 cubic,
                              2 FCC, 3 BCC, ...), selecting the minimum image of
                              E_SR (apply_pbc.f90); 0 = not given: cubic when the
                              cell is cubic, else nearest image in fractional
                              coordinates */
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

/* Species tables, all on the caller's sphere orders (set once per run).
 * vps/rhos: V_s(G)/Omega and -Z exp(-G^2 r_c^2/4)/Omega on the density
 * sphere (form_factor, exp_igr.f90:142-288); twnl: Y_lm(G^) w(|G|)/sqrt(Omega)
 * on the wavefunction sphere (nlpro, pseudopotential.f90:108-162, computed
 * once instead of every step); dmat: projector coupling per species (the
 * diagonal KB weights wsg of nlppgrid, pseudopotential.f90:434-529, or UPF
 * hlsg blocks).  Replaces the module globals vps, rhos, twnl, wsg/hlsg. */
typedef struct {
  int nsp, maxproj;
  const double* zv;      /* [nsp] valence charge */
  const double* vps;     /* [nsp][ngrho] */
  const double* rhos;    /* [nsp][ngrho] */
  const int* nproj;      /* [nsp] projectors of the species (ngh) */
  const int* proj_l;     /* [nsp][maxproj] angular momentum (nghtol) */
  const double* twnl;    /* [nsp][maxproj][ngpw] */
  const double* dmat;    /* [nsp][maxproj][maxproj] */
} amdk_species;
int amdk_set_species(amdk_ctx* ctx, const amdk_species* s);

/* PCG preconditioner weights WSG(is, iv) of hess (cg.f90:20-40), which adds
 * N_s WSG(s, k) twnl_k(G)^2 per projector to K(G): wsg[s * maxproj + k], with
 * the maxproj of the species descriptor.  amdk_set_species sets them to the
 * diagonal of dmat (the KB weights of the CPMD format); the reference never
 * sets WSG for UPF species (it is filled only by nlppgrid,
 * pseudopotential.f90:480-489), so its callers pass zeros for them.  Takes
 * effect at once if positions are set, else at amdk_set_positions. */
int amdk_set_precon_weights(amdk_ctx* ctx, const double* wsg);

/* Atomic positions (bohr, tau[3*a+k]) and species of each atom.  Builds the
 * structure factors, sum_a V_s e^{-iGR}, sum_a rho_s e^{-iGR}, the projector
 * table beta = (-i)^l twnl e^{-iGR} and the short-range ion energy/forces.
 * Replaces exp_igrxfactor (exp_igr.f90:4-140) and E_SR (total_energy.f90:133-239). */
int amdk_set_positions(amdk_ctx* ctx, int nat, const int* sp_of_atom, const double* tau);
/* Number of projector rows (sum over atoms of nproj) = rows of nl. */
int amdk_nproj(const amdk_ctx* ctx);

/* nl(p, b) = <beta_p|psi_b> (full sphere), column-major (nproj, nb).
 * Replaces fnl(C0, nstat) (pseudopotential.f90:164-226) and upf_fnl
 * (pseudopotential.f90:356-433): the same product with UPF tables. */
int amdk_fnl(amdk_ctx* ctx, const double* c0, int ldc, int nb, double* nl);
/* c2 += -f~ sum_p beta_p (D nl)_p with f~ = f (>= 1e-5) else 1.
 * Replaces FNONLOC(CA, NOSTATE, FC) (wfn_initialize.f90:560-612) and
 * UPF_FNONLOC (wfn_initialize.f90:614-695; D = hlsg over the same (l, m)). */
int amdk_fnonloc(amdk_ctx* ctx, const double* nl, double* c2, int ldc, const double* occ, int nb);
/* A(i,j) = 2 Re <ca_i|cb_j>_half - ca_i(0) cb_j(0), column-major (n1, n2).
 * Replaces OVLAP_D(NG_PW, N1, N2, A, CA, CB) (wfn_initialize.f90:234-263). */
int amdk_overlap(amdk_ctx* ctx, const double* ca, const double* cb, int ldc, int n1, int n2, double* A);
/* amdk_overlap that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its split-K partial
 * sums go to a fixed 32 MiB array of the library, one per GPU, held until the
 * process ends (looked up at amdk_ctx_create); everything else uses buffers
 * the context allocated at set-up.  Limit: n1*n2 <= 4,194,304 (AMDK_ERR_ARG
 * beyond it).  Calls on contexts of the same GPU share that array, so calls
 * that overlap in time on one GPU (several host threads) race; nothing in it
 * needs to start at zero. */
int amdk_overlap_noalloc(amdk_ctx* ctx, const double* ca, const double* cb, int ldc, int n1, int n2, double* A);
/* Gradient projection: G = <c2|c0> masked by occupation, c2 -= c0 G,
 * Im c2(G=0) = 0; returns GEMAX and CNORM.
 * Replaces FORCES + eval_GMAX (gradient.f90:5-61). */
int amdk_lagrange(amdk_ctx* ctx, const double* c0, double* c2, int ldc, const double* occ, int nb,
                  double* gemax, double* cnorm);
/* Cholesky orthonormalisation in FP64: S = U^T U, c <- c U^-1, Im c(0) = 0.
 * Returns AMDK_ERR_NOT_PD where UINV would STOP.
 * Replaces ortho_gs(CP) (wfn_initialize.f90:448-558). */
int amdk_ortho(amdk_ctx* ctx, double* c, int ldc, int nb);

enum { AMDK_XC_TETER93 = 0, AMDK_XC_PZ = 1, AMDK_XC_PBE = 2 };

/* Energies of one SCF step (Hartree), as printed by eprint.f90. */
typedef struct {
  double etot, ekin, estat, eself, esr, eloc, enl, exc, eh;
  double gemax, cnorm, nelec;
} amdk_energies;

/* From rho(r): rho(G) (rhor2g, density.f90:128-150), V_H + V_loc
 * (eval_pot, potential.f90:4-46) and V_xc (cal_vpot_ex, xc.f90:4-497) ->
 * V_KS(r) kept on the device (and copied to v_r when non-null).
 * e receives eh, eloc, exc (others untouched). */
int amdk_vks(amdk_ctx* ctx, const double* rho_r, int xc, double* v_r, amdk_energies* e);

/* Forces -dE/dR (Ha/bohr) [3*nat]: POTFOR + NLFOR/D_FNL (or UPF_NLFOR/
 * UPF_D_FNL, nuclear_grad.f90:90-199) + E_SR (nuclear_grad.f90:4-372),
 * using the rho(G) formed last on the device (by amdk_vks or an SCF step). */
int amdk_forces(amdk_ctx* ctx, const double* c0, int ldc, const double* occ, int nb, double* f);

/* Device-resident SCF (wfn_opt loop body, do_opt.f90:80-129): load C_0,
 * run steps, read C_0 back.  istep is 1-based within one optimisation (it
 * drives the PCG momentum schedule of cg.f90:88-129). */
int amdk_wfn_load(amdk_ctx* ctx, const double* c0, int ldc, const double* occ, int nb);
int amdk_wfn_store(amdk_ctx* ctx, double* c0, int ldc);
int amdk_scf_step(amdk_ctx* ctx, int istep, int xc, amdk_energies* e);
/* As amdk_scf_step, and with f non-null also the nuclear forces [3*nat] of
 * this step's state before its wavefunction update: the TFOR branch of
 * wfn_opt (E_SR + POTFOR + NLFOR, do_opt.f90:97-106). */
int amdk_scf_step_tfor(amdk_ctx* ctx, int istep, int xc, amdk_energies* e, double* f);
/* Forces for the resident C_0 and the rho(G) of the last SCF step;
 * AMDK_ERR_STATE if no step ran since amdk_set_positions or a stateless
 * amdk_vks has replaced that rho(G) since. */
int amdk_scf_forces(amdk_ctx* ctx, double* f);
/* Cholesky-orthonormalise the resident C_0 (ortho_gs(c_0) after a geometry
 * update, molecular_dynamics.f90:108, do_opt.f90:259). */
int amdk_wfn_ortho(amdk_ctx* ctx);

/* Transfer helpers between caller-layout FP64 host arrays and the internal
 * device layout (no reference counterpart: they replace nothing, they keep
 * C_0/C2 resident across an SCF step). */
int amdk_bands_to_dev(amdk_ctx* ctx, const double* c, int ldc, int nb, void* dst);
int amdk_bands_from_dev(amdk_ctx* ctx, const void* src, int nb, double* c, int ldc);

/* Sphere geometry as the FFT kernels see it: which = 0 wavefunction sphere,
 * 1 density sphere; out = {npw, ldc, half rows, full rows, planes, R}. */
int amdk_sphere_info(const amdk_ctx* ctx, int which, int* out6);

/* ---- the reference's own array layouts and the remaining routines of one
 * SCF step, for a caller that keeps the reference's globals on the host ---- */

/* Layouts of the caller's host arrays, applied from this call on to every
 * entry point that takes or returns them.  0 keeps the dense default.
 *  grid_ld1, grid_ld2: real-space grids (RHO, V_POT) with element (x, y, z)
 *    at x + ld1 (y + ld2 z), the reference's nrlead padding (RHO(NNR1),
 *    NNR1 = nrlead(1) nrlead(2) nrlead(3), init_module.f90:159,
 *    gvectors.f90:64,494-497); padding entries are neither read nor written.
 *  nl_nghmax: nl(1, atom_t, nghmax, nstate) (pseudopotential.f90:174):
 *    projector k of atom a (atoms in amdk_set_positions order, grouped by
 *    species as the reference's isa) at a + nat (k + nghmax b).
 *  force_na_max: force(3, NA_MAX, sp_t) (system_data_types.f90:151): atom a,
 *    the ia-th of its species s, at 3 (ia + NA_MAX s) + k; unused atom slots
 *    are zeroed where a routine overwrites the array. */
typedef struct {
  int grid_ld1, grid_ld2;
  int nl_nghmax;
  int force_na_max;
} amdk_layout;
int amdk_set_layout(amdk_ctx* ctx, const amdk_layout* layout);

/* exp_igrxfactor (exp_igr.f90:4-140) from the reference's own arrays:
 * atco(3, NA_MAX, sp_t) bohr with atom_p_sp(sp_t) atoms per species;
 * otherwise as amdk_set_positions (atoms in species order). */
int amdk_set_positions_atco(amdk_ctx* ctx, int nsp, const int* atom_p_sp, int na_max, const double* atco);

/* Parts of the nuclear forces, ADDED into f (the caller's force layout), as
 * E_SR (total_energy.f90:133-239, whose caller zeroes force first), POTFOR
 * (nuclear_grad.f90:4-50, rho(G) of the last amdk_vks or SCF step) and NLFOR
 * / UPF_NLFOR (nuclear_grad.f90:51-89,153-199, for c0 with occupations occ)
 * accumulate into force.  c0/occ/nb are only read for NLFOR. */
enum { AMDK_FORCE_ESR = 1, AMDK_FORCE_POTFOR = 2, AMDK_FORCE_NLFOR = 4 };
int amdk_forces_add(amdk_ctx* ctx, const double* c0, int ldc, const double* occ, int nb, int parts, double* f);

/* PCG(istep) (cg.f90:63-145 with hess/PRECON, cg.f90:3-61,281-311):
 * P = f^-1 K(G)^-1 C2, d = P + gamma d_prev (gamma = forget ||P||^2 /
 * ||P_prev||^2, forget 0.6 / 0.91 / 0.95 for istep <= 10 / <= 20 / later),
 * c0 += (5^2 / 400) d.  hnm1 is HNM1 (d_prev in, d out; not read at istep 1).
 * The ||P||^2 history is kept by the context, as the reference SAVEs it, and
 * apart from the resident SCF's. */
int amdk_pcg(amdk_ctx* ctx, double* c0, const double* c2, double* hnm1, int ldc, const double* occ, int nb, int istep);

/* Energy terms that e_total (total_energy.f90:4-37) sums besides those of
 * amdk_vks: E_kin = sum_i f_i sum_G |G|^2 |c_i(G)|^2 (e_ke, total_energy.f90:38-63),
 * E_nl of pp_energy (pseudopotential.f90:55-99) from nl (caller layout),
 * and E_SR / E_self of the positions (total_energy.f90:133-239). */
int amdk_ekin(amdk_ctx* ctx, const double* c0, int ldc, const double* occ, int nb, double* ekin);
int amdk_enl(amdk_ctx* ctx, const double* nl, const double* occ, int nb, double* enl);
int amdk_ion_energies(amdk_ctx* ctx, double* esr, double* eself);

/* Orthonormalisation inside amdk_scf_step / amdk_scf_step_tfor:
 * AMDK_ORTHO_CHOLESKY (default) is ortho_gs (wfn_initialize.f90:448-558);
 * AMDK_ORTHO_LOWDIN is C <- C S^-1/2 in FP64 (S^-1/2 by a coupled
 * Newton-Schulz iteration, for bands that were orthonormal before the step's
 * PCG update).  No counterpart in the reference, which uses Cholesky only;
 * the converged state is the same, the SCF path differs.  amdk_ortho and
 * amdk_wfn_ortho stay Cholesky. */
enum { AMDK_ORTHO_CHOLESKY = 0, AMDK_ORTHO_LOWDIN = 1 };
int amdk_set_ortho_method(amdk_ctx* ctx, int method);

/* Per-kernel timing with HIP events on the context stream. */
int amdk_profile_enable(amdk_ctx* ctx, int on);
/* Names of the timed regions so far, '\n'-separated, into buf[len]. */
int amdk_profile_names(amdk_ctx* ctx, char* buf, int len);
int amdk_profile_query(amdk_ctx* ctx, const char* kernel, double* total_ms, int* count);
int amdk_profile_reset(amdk_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
