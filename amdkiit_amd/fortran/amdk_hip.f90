! amdk_hip: ISO_C_BINDING interfaces of libamdk (include/amdk.h) and the
! error check.  The reference's routines with their own names and argument
! lists (VPSI, FNONLOC, fnl, OVLAP_D, FORCES, PCG, ortho_gs, ...) are in
! module amdk_ref (amdk_ref.f90), built on these interfaces.
module amdk_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: AMDK_OK = 0, AMDK_ERR_NOT_PD = 4
  integer(c_int), parameter, public :: AMDK_PREC_FP64 = 0, AMDK_PREC_MIXED = 1, AMDK_PREC_BF16 = 2
  integer(c_int), parameter, public :: AMDK_XC_TETER93 = 0, AMDK_XC_PZ = 1, AMDK_XC_PBE = 2

  type, bind(C), public :: amdk_desc
    integer(c_int) :: n1, n2, n3
    real(c_double) :: a_cell(9)
    integer(c_int) :: ngpw
    type(c_ptr) :: miller_pw
    integer(c_int) :: ngrho
    type(c_ptr) :: miller_rho
    integer(c_int) :: nbmax, precision, symmetry
  end type

  type, bind(C), public :: amdk_species
    integer(c_int) :: nsp, maxproj
    type(c_ptr) :: zv, vps, rhos, nproj, proj_l, twnl, dmat
  end type

  type, bind(C), public :: amdk_layout
    integer(c_int) :: grid_ld1, grid_ld2, nl_nghmax, force_na_max
  end type

  integer(c_int), parameter, public :: AMDK_FORCE_ESR = 1, AMDK_FORCE_POTFOR = 2, AMDK_FORCE_NLFOR = 4

  type, bind(C), public :: amdk_energies
    real(c_double) :: etot, ekin, estat, eself, esr, eloc, enl, exc, eh
    real(c_double) :: gemax, cnorm, nelec
  end type

  public :: amdk_ctx_create, amdk_ctx_destroy, amdk_set_vks, amdk_vpsi, amdk_density, &
            amdk_set_species, amdk_set_precon_weights, amdk_set_positions, amdk_nproj, amdk_fnl, amdk_fnonloc, &
            amdk_overlap, amdk_lagrange, amdk_ortho, amdk_vks, amdk_forces, &
            amdk_wfn_load, amdk_wfn_store, amdk_scf_step, amdk_scf_forces, amdk_scf_step_tfor, &
            amdk_wfn_ortho
  public :: amdk_check, amdk_set_layout, amdk_set_positions_atco, amdk_forces_add, amdk_pcg, amdk_ekin, &
            amdk_enl, amdk_ion_energies, amdk_overlap_noalloc

  interface
    integer(c_int) function amdk_ctx_create(d, ctx) bind(C, name="amdk_ctx_create")
      import :: c_int, c_ptr, amdk_desc
      type(amdk_desc), intent(in) :: d
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine amdk_ctx_destroy(ctx) bind(C, name="amdk_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function amdk_last_error(ctx) bind(C, name="amdk_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function amdk_set_vks(ctx, v) bind(C, name="amdk_set_vks")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: v(*)
    end function
    integer(c_int) function amdk_vpsi(ctx, c0, ldc, c2, occ, nb) bind(C, name="amdk_vpsi")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb
      complex(c_double_complex), intent(inout) :: c2(*)
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function amdk_density(ctx, c0, ldc, occ, nb, rho) bind(C, name="amdk_density")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(in) :: occ(*)
      real(c_double), intent(out) :: rho(*)
    end function
    integer(c_int) function amdk_set_species(ctx, s) bind(C, name="amdk_set_species")
      import :: c_int, c_ptr, amdk_species
      type(c_ptr), value :: ctx
      type(amdk_species), intent(in) :: s
    end function
    integer(c_int) function amdk_set_precon_weights(ctx, wsg) bind(C, name="amdk_set_precon_weights")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: wsg(*)
    end function
    integer(c_int) function amdk_set_positions(ctx, nat, sp, tau) bind(C, name="amdk_set_positions")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nat
      integer(c_int), intent(in) :: sp(*)
      real(c_double), intent(in) :: tau(*)
    end function
    integer(c_int) function amdk_nproj(ctx) bind(C, name="amdk_nproj")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function amdk_fnl(ctx, c0, ldc, nb, nl) bind(C, name="amdk_fnl")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(out) :: nl(*)
    end function
    integer(c_int) function amdk_fnonloc(ctx, nl, c2, ldc, occ, nb) bind(C, name="amdk_fnonloc")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: nl(*)
      complex(c_double_complex), intent(inout) :: c2(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function amdk_overlap(ctx, ca, cb, ldc, n1, n2, a) bind(C, name="amdk_overlap")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: ca(*), cb(*)
      integer(c_int), value :: ldc, n1, n2
      real(c_double), intent(out) :: a(*)
    end function
    integer(c_int) function amdk_lagrange(ctx, c0, c2, ldc, occ, nb, gemax, cnorm) bind(C, name="amdk_lagrange")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      complex(c_double_complex), intent(inout) :: c2(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(in) :: occ(*)
      real(c_double), intent(out) :: gemax, cnorm
    end function
    integer(c_int) function amdk_ortho(ctx, c, ldc, nb) bind(C, name="amdk_ortho")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: c(*)
      integer(c_int), value :: ldc, nb
    end function
    integer(c_int) function amdk_vks(ctx, rho, xc, v, e) bind(C, name="amdk_vks")
      import :: c_int, c_ptr, c_double, amdk_energies
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rho(*)
      integer(c_int), value :: xc
      real(c_double), intent(out) :: v(*)
      type(amdk_energies), intent(inout) :: e
    end function
    integer(c_int) function amdk_forces(ctx, c0, ldc, occ, nb, f) bind(C, name="amdk_forces")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(in) :: occ(*)
      real(c_double), intent(out) :: f(*)
    end function
    integer(c_int) function amdk_wfn_load(ctx, c0, ldc, occ, nb) bind(C, name="amdk_wfn_load")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function amdk_wfn_store(ctx, c0, ldc) bind(C, name="amdk_wfn_store")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: c0(*)
      integer(c_int), value :: ldc
    end function
    integer(c_int) function amdk_scf_step(ctx, istep, xc, e) bind(C, name="amdk_scf_step")
      import :: c_int, c_ptr, amdk_energies
      type(c_ptr), value :: ctx
      integer(c_int), value :: istep, xc
      type(amdk_energies), intent(out) :: e
    end function
    integer(c_int) function amdk_scf_step_tfor(ctx, istep, xc, e, f) bind(C, name="amdk_scf_step_tfor")
      import :: c_int, c_ptr, c_double, amdk_energies
      type(c_ptr), value :: ctx
      integer(c_int), value :: istep, xc
      type(amdk_energies), intent(out) :: e
      real(c_double), intent(out) :: f(*)
    end function
    integer(c_int) function amdk_wfn_ortho(ctx) bind(C, name="amdk_wfn_ortho")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function amdk_scf_forces(ctx, f) bind(C, name="amdk_scf_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: f(*)
    end function
    integer(c_int) function amdk_overlap_noalloc(ctx, ca, cb, ldc, n1, n2, a) bind(C, name="amdk_overlap_noalloc")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: ca(*), cb(*)
      integer(c_int), value :: ldc, n1, n2
      real(c_double), intent(out) :: a(*)
    end function
    integer(c_int) function amdk_set_layout(ctx, l) bind(C, name="amdk_set_layout")
      import :: c_int, c_ptr, amdk_layout
      type(c_ptr), value :: ctx
      type(amdk_layout), intent(in) :: l
    end function
    integer(c_int) function amdk_set_positions_atco(ctx, nsp, atom_p_sp, na_max, atco) &
        bind(C, name="amdk_set_positions_atco")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsp, na_max
      integer(c_int), intent(in) :: atom_p_sp(*)
      real(c_double), intent(in) :: atco(*)
    end function
    integer(c_int) function amdk_forces_add(ctx, c0, ldc, occ, nb, parts, f) bind(C, name="amdk_forces_add")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb, parts
      real(c_double), intent(in) :: occ(*)
      real(c_double), intent(inout) :: f(*)
    end function
    integer(c_int) function amdk_pcg(ctx, c0, c2, hnm1, ldc, occ, nb, istep) bind(C, name="amdk_pcg")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: c0(*), hnm1(*)
      complex(c_double_complex), intent(in) :: c2(*)
      integer(c_int), value :: ldc, nb, istep
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function amdk_ekin(ctx, c0, ldc, occ, nb, ekin) bind(C, name="amdk_ekin")
      import :: c_int, c_ptr, c_double_complex, c_double
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: c0(*)
      integer(c_int), value :: ldc, nb
      real(c_double), intent(in) :: occ(*)
      real(c_double), intent(out) :: ekin
    end function
    integer(c_int) function amdk_enl(ctx, nl, occ, nb, enl) bind(C, name="amdk_enl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: nl(*), occ(*)
      integer(c_int), value :: nb
      real(c_double), intent(out) :: enl
    end function
    integer(c_int) function amdk_ion_energies(ctx, esr, eself) bind(C, name="amdk_ion_energies")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: esr, eself
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Non-zero status -> message and STOP (UINV behaviour of the reference).
  subroutine amdk_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == AMDK_OK) return
    n = int(c_strlen(amdk_last_error(ctx)))
    call c_f_pointer(amdk_last_error(ctx), msg, [n])
    write (*, '(A,A,I3,A,*(A))') where, ': amdk error ', rc, ' : ', msg(1:n)
    stop 1
  end subroutine

end module amdk_hip
