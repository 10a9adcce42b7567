! amdk_ref: the reference's hot-path subroutines under their own names and
! argument lists, on libamdk.  A reference build USEs this module instead of
! the modules that define them (wfn_initialize, pseudopotential, density,
! potential, xc, gradient, cg, nuclear_grad, total_energy, exp_igr) and
! keeps its call sites, e.g. the wfn_opt loop of do_opt.f90:82-112.
!
! The module holds what those routines take implicitly from
! system_data_types: NGPW_l, NSTATE, OCCUPATION, C_0, C2, HNM1, RHO, V_POT,
! nl, force, atco, atom_p_sp, GEMAX, CNORM, the energy terms.  Arrays keep
! the reference's shapes and leading dimensions:
!   RHO(NNR1, 1), V_POT(NNR1, 1) with NNR1 = nrlead(1) nrlead(2) nrlead(3),
!     nrlead = n + mod(n + 1, 2)                     (init_module.f90:159,
!                                                     gvectors.f90:64,494-497)
!   nl(1, atom_t, nghmax, nstate), allocated by fnl (pseudopotential.f90:174)
!   force(3, NA_MAX, sp_t), allocated by E_SR       (total_energy.f90:146,
!                                                     system_data_types.f90:151)
!   C_0, C2, HNM1 (NGPW_l, NSTATE)
! One call, amdk_ref_init, creates the device context from the plane-wave
! setup (Miller indices of both spheres, FFT grid, cell) and sets these
! layouts; amdk_ref_species passes the pseudopotential tables.  A failing
! library call prints its message and STOPs, as the reference does on
! LAPACK failures (wfn_initialize.f90:513-527).
module amdk_ref
  use iso_c_binding
  use amdk_hip
  implicit none

  type(c_ptr), save :: amdk_context = c_null_ptr
  integer, save :: ngpw_l = 0, nstate = 0, atom_t = 0, nghmax = 0, na_max = 0, sp_t = 0
  integer, save :: nrgrids(3) = 0, nrlead(3) = 0
  integer(8), save :: nnr1 = 0
  integer(c_int), save :: xc_code = AMDK_XC_TETER93
  real(c_double), save :: cell_volume = 0.0d0
  logical, save :: tfor = .false.

  complex(c_double_complex), allocatable, target :: c_0(:,:), c2(:,:), hnm1(:,:)
  real(c_double), allocatable, target :: rho(:,:), v_pot(:,:), occupation(:)
  real(c_double), allocatable, target :: atco(:,:,:)
  integer(c_int), allocatable, target :: atom_p_sp(:)
  real(c_double), pointer :: nl(:,:,:,:) => null()
  real(c_double), pointer :: force(:,:,:) => null()

  ! GEMAX/CNORM of FORCES; the energy terms of e_total (eprint.f90 names)
  real(c_double), save :: gemax = 0.0d0, cnorm = 0.0d0
  real(c_double), save :: etotal = 0.0d0, eke = 0.0d0, eht = 0.0d0, epseu = 0.0d0, enl = 0.0d0, exc = 0.0d0
  real(c_double), save :: esr = 0.0d0, eself = 0.0d0, estat = 0.0d0

contains

  ! Context for the grid nr(3), cell a_cell(3,3) (lattice vectors as columns),
  ! set_vec's lattice code, the two half spheres as Miller indices
  ! (inyh - nh: mill_pw(3, ngpw) with G = 0 first, mill_rho(3, ngrho)), the
  ! state count and precision; allocates the module arrays above.
  subroutine amdk_ref_init(nr, a_cell, symmetry, ngpw, mill_pw, ngrho, mill_rho, nstates, precision, xc, &
                           nsp, nat_per_species, namax, ngh_max)
    integer, intent(in) :: nr(3), symmetry, ngpw, ngrho, nstates, precision, xc, nsp, namax, ngh_max
    real(c_double), intent(in) :: a_cell(3, 3)
    integer(c_int), intent(in), target :: mill_pw(3, ngpw), mill_rho(3, ngrho)
    integer, intent(in) :: nat_per_species(nsp)
    type(amdk_desc) :: d
    type(amdk_layout) :: lay
    integer :: i, j
    nrgrids = nr
    nrlead = nr + mod(nr + 1, 2)
    nnr1 = int(nrlead(1), 8) * nrlead(2) * nrlead(3)
    ngpw_l = ngpw
    nstate = nstates
    sp_t = nsp
    na_max = namax
    nghmax = ngh_max
    atom_t = sum(nat_per_species)
    xc_code = int(xc, c_int)
    cell_volume = abs(a_cell(1,1) * (a_cell(2,2) * a_cell(3,3) - a_cell(3,2) * a_cell(2,3)) &
                    - a_cell(1,2) * (a_cell(2,1) * a_cell(3,3) - a_cell(3,1) * a_cell(2,3)) &
                    + a_cell(1,3) * (a_cell(2,1) * a_cell(3,2) - a_cell(3,1) * a_cell(2,2)))
    d%n1 = nr(1); d%n2 = nr(2); d%n3 = nr(3)
    do j = 1, 3
      do i = 1, 3
        d%a_cell(3 * (j - 1) + i) = a_cell(i, j)
      end do
    end do
    d%ngpw = int(ngpw, c_int); d%miller_pw = c_loc(mill_pw)
    d%ngrho = int(ngrho, c_int); d%miller_rho = c_loc(mill_rho)
    d%nbmax = int(nstates, c_int); d%precision = int(precision, c_int); d%symmetry = int(symmetry, c_int)
    call amdk_check(amdk_context, amdk_ctx_create(d, amdk_context), 'amdk_ref_init')
    lay%grid_ld1 = nrlead(1); lay%grid_ld2 = nrlead(2)
    lay%nl_nghmax = int(nghmax, c_int); lay%force_na_max = int(na_max, c_int)
    call amdk_check(amdk_context, amdk_set_layout(amdk_context, lay), 'amdk_ref_init')
    allocate(c_0(ngpw_l, nstate), c2(ngpw_l, nstate), hnm1(ngpw_l, nstate))
    allocate(rho(nnr1, 1), v_pot(nnr1, 1), occupation(nstate))
    allocate(atco(3, na_max, sp_t), atom_p_sp(sp_t))
    atom_p_sp = int(nat_per_species, c_int)
    c_0 = (0.0d0, 0.0d0); c2 = (0.0d0, 0.0d0); hnm1 = (0.0d0, 0.0d0)
    rho = 0.0d0; v_pot = 0.0d0; occupation = 0.0d0; atco = 0.0d0
  end subroutine

  ! Pseudopotential tables of every species on the two spheres (form_factor,
  ! nlpro / nlppgrid and their UPF forms): vps, rhos(ngrho, sp_t),
  ! twnl(ngpw, nghmax, sp_t), ngh(sp_t), nghtol(nghmax, sp_t) and the
  ! projector coupling dmat(nghmax, nghmax, sp_t) (wsg on the diagonal, or
  ! hlsg).
  ! wsg: the reference's WSG(spmax, maxngh) as hess reads it (cg.f90:20-40);
  ! zero for UPF species, whose path never sets it.  Without it the library
  ! takes the diagonal of dmat.
  subroutine amdk_ref_species(zv, vps, rhos, ngh, nghtol, twnl, dmat, wsg)
    real(c_double), intent(in), target :: zv(:), vps(:,:), rhos(:,:), twnl(:,:,:), dmat(:,:,:)
    integer(c_int), intent(in), target :: ngh(:), nghtol(:,:)
    real(c_double), intent(in), optional :: wsg(:,:)
    type(amdk_species) :: s
    real(c_double), allocatable :: w(:,:)
    s%nsp = int(sp_t, c_int); s%maxproj = int(nghmax, c_int)
    s%zv = c_loc(zv); s%vps = c_loc(vps); s%rhos = c_loc(rhos)
    s%nproj = c_loc(ngh); s%proj_l = c_loc(nghtol); s%twnl = c_loc(twnl); s%dmat = c_loc(dmat)
    call amdk_check(amdk_context, amdk_set_species(amdk_context, s), 'amdk_ref_species')
    if (present(wsg)) then
      allocate(w(nghmax, sp_t))          ! projector fastest, as the library reads it
      w = transpose(wsg(1:sp_t, 1:nghmax))
      call amdk_check(amdk_context, amdk_set_precon_weights(amdk_context, w), 'amdk_ref_species (wsg)')
      deallocate(w)
    end if
  end subroutine

  subroutine amdk_ref_finalize()
    if (c_associated(amdk_context)) call amdk_ctx_destroy(amdk_context)
    amdk_context = c_null_ptr
    if (associated(nl)) deallocate(nl)
    if (associated(force)) deallocate(force)
    deallocate(c_0, c2, hnm1, rho, v_pot, occupation, atco, atom_p_sp)
  end subroutine

  ! exp_igr.f90:4-140: structure factors (and E_SR) of atco
  subroutine exp_igrxfactor()
    call amdk_check(amdk_context, amdk_set_positions_atco(amdk_context, int(sp_t, c_int), atom_p_sp, &
                    int(na_max, c_int), atco), 'exp_igrxfactor')
  end subroutine

  ! density.f90:5-98: RHO(:,1) = sum_i f_i |psi_i(r)|^2 / Omega
  subroutine eval_density()
    call amdk_check(amdk_context, amdk_density(amdk_context, c_0, int(ngpw_l, c_int), occupation, &
                    int(nstate, c_int), rho), 'eval_density')
  end subroutine

  ! density.f90:128-150: rho(G) is formed on the device from RHO by eval_pot
  subroutine rhor2g()
  end subroutine

  ! potential.f90:4-46 with xc.f90:4-434: V_POT(:,1) = V_H + V_loc + V_xc of
  ! RHO(:,1); E_H, E_loc and E_xc of the same density
  subroutine eval_pot()
    type(amdk_energies) :: e
    call amdk_check(amdk_context, amdk_vks(amdk_context, rho, xc_code, v_pot, e), 'eval_pot')
    eht = e%eh * cell_volume
    epseu = e%eloc
    exc = e%exc
  end subroutine

  ! xc.f90:4-434: V_xc is added to V_POT by eval_pot's library call
  subroutine cal_vpot_ex()
  end subroutine

  ! pseudopotential.f90:3-103: nl = <beta|C_0> and E_nl (E_loc: eval_pot)
  subroutine pp_energy()
    call fnl(c_0, nstate)
    call amdk_check(amdk_context, amdk_enl(amdk_context, nl, occupation, int(nstate, c_int), enl), 'pp_energy')
  end subroutine

  ! total_energy.f90:4-37
  subroutine e_total()
    call amdk_check(amdk_context, amdk_ekin(amdk_context, c_0, int(ngpw_l, c_int), occupation, int(nstate, c_int), &
                    eke), 'e_ke')
    call amdk_check(amdk_context, amdk_ion_energies(amdk_context, esr, eself), 'e_sr')
    estat = eht + esr - eself
    etotal = enl + epseu + eke + exc + estat
  end subroutine

  ! wfn_initialize.f90:268-365: C2_nl = -(f/2)(|G|^2 c + 2 (V_POT psi)(G))
  subroutine vpsi(c_atom, c2_nl, fc, nostate)
    integer, intent(in) :: nostate
    complex(c_double_complex), intent(in) :: c_atom(ngpw_l, nostate)
    complex(c_double_complex), intent(inout) :: c2_nl(ngpw_l, nostate)
    real(c_double), intent(in) :: fc(nostate)
    call amdk_check(amdk_context, amdk_set_vks(amdk_context, v_pot), 'VPSI')
    call amdk_check(amdk_context, amdk_vpsi(amdk_context, c_atom, int(ngpw_l, c_int), c2_nl, fc, &
                    int(nostate, c_int)), 'VPSI')
  end subroutine

  ! wfn_initialize.f90:560-612: CA += -f~ sum beta D nl, nl of the last fnl
  subroutine fnonloc(ca, nostate, fc)
    integer, intent(in) :: nostate
    complex(c_double_complex), intent(inout) :: ca(ngpw_l, nostate)
    real(c_double), intent(in) :: fc(nostate)
    call amdk_check(amdk_context, amdk_fnonloc(amdk_context, nl, ca, int(ngpw_l, c_int), fc, int(nostate, c_int)), &
                    'FNONLOC')
  end subroutine

  ! wfn_initialize.f90:614-695 (D = hlsg from amdk_ref_species)
  subroutine upf_fnonloc(ca, nostate, fc)
    integer, intent(in) :: nostate
    complex(c_double_complex), intent(inout) :: ca(ngpw_l, nostate)
    real(c_double), intent(in) :: fc(nostate)
    call fnonloc(ca, nostate, fc)
  end subroutine

  ! pseudopotential.f90:164-226: ALLOCATE(nl(1, atom_t, nghmax, nstat)), nl = <beta|C0>
  subroutine fnl(c0, nstat)
    integer, intent(in) :: nstat
    complex(c_double_complex), intent(in) :: c0(ngpw_l, nstat)
    if (associated(nl)) deallocate(nl)
    allocate(nl(1, atom_t, nghmax, nstat))
    call amdk_check(amdk_context, amdk_fnl(amdk_context, c0, int(ngpw_l, c_int), int(nstat, c_int), nl), 'fnl')
  end subroutine

  ! pseudopotential.f90:356-433
  subroutine upf_fnl(c0, nstat)
    integer, intent(in) :: nstat
    complex(c_double_complex), intent(in) :: c0(ngpw_l, nstat)
    call fnl(c0, nstat)
  end subroutine

  ! wfn_initialize.f90:234-263: A = 2 Re <CA|CB> - CA(0) CB(0)
  subroutine ovlap_d(ng_pw, n1, n2, a, ca, cb)
    integer, intent(in) :: ng_pw, n1, n2
    real(c_double), intent(out) :: a(n1, n2)
    complex(c_double_complex), intent(in) :: ca(ng_pw, n1), cb(ng_pw, n2)
    call amdk_check(amdk_context, amdk_overlap(amdk_context, ca, cb, int(ng_pw, c_int), int(n1, c_int), &
                    int(n2, c_int), a), 'OVLAP_D')
  end subroutine

  ! gradient.f90:5-61: C2 -= C_0 Lambda (occupied), Im C2(G=0) = 0, GEMAX, CNORM
  subroutine forces()
    call amdk_check(amdk_context, amdk_lagrange(amdk_context, c_0, c2, int(ngpw_l, c_int), occupation, &
                    int(nstate, c_int), gemax, cnorm), 'FORCES')
  end subroutine

  ! total_energy.f90:133-239 (TFOR part): force = E_SR forces
  subroutine e_sr()
    if (.not. associated(force)) allocate(force(3, na_max, sp_t))
    force = 0.0d0
    call amdk_check(amdk_context, amdk_forces_add(amdk_context, c_0, int(ngpw_l, c_int), occupation, &
                    int(nstate, c_int), AMDK_FORCE_ESR, force), 'E_SR')
  end subroutine

  ! nuclear_grad.f90:4-50: force += local-potential forces
  subroutine potfor()
    call amdk_check(amdk_context, amdk_forces_add(amdk_context, c_0, int(ngpw_l, c_int), occupation, &
                    int(nstate, c_int), AMDK_FORCE_POTFOR, force), 'POTFOR')
  end subroutine

  ! nuclear_grad.f90:51-89 (with D_FNL, :314-372): force += nonlocal forces of C_0
  subroutine nlfor()
    call amdk_check(amdk_context, amdk_forces_add(amdk_context, c_0, int(ngpw_l, c_int), occupation, &
                    int(nstate, c_int), AMDK_FORCE_NLFOR, force), 'NLFOR')
  end subroutine

  ! nuclear_grad.f90:90-199
  subroutine upf_nlfor()
    call nlfor()
  end subroutine

  ! cg.f90:63-145: preconditioned CG update of C_0 along HNM1
  subroutine pcg(istep)
    integer, intent(in) :: istep
    call amdk_check(amdk_context, amdk_pcg(amdk_context, c_0, c2, hnm1, int(ngpw_l, c_int), occupation, &
                    int(nstate, c_int), int(istep, c_int)), 'PCG')
  end subroutine

  ! wfn_initialize.f90:448-558: Cholesky orthonormalisation of CP
  subroutine ortho_gs(cp)
    complex(c_double_complex), intent(inout) :: cp(:,:)
    call amdk_check(amdk_context, amdk_ortho(amdk_context, cp, int(size(cp, 1), c_int), int(size(cp, 2), c_int)), &
                    'ortho_gs')
  end subroutine

end module amdk_ref
