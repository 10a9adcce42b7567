! Test driver of module amdk_ref: allocates the reference's arrays with the
! reference's shapes, runs NSTEP iterations of the wfn_opt loop body in the
! order of do_opt.f90:82-112 (TFOR on the last one) and writes the energy
! terms of every step, the forces and the final C_0.
!
! Usage: amdk_wfopt <setup file> <output file>
! Setup (stream, native endianness; integers int32, reals float64):
!   n1 n2 n3 symmetry precision xc nstate nsp na_max nghmax ngpw ngrho nstep
!   a_cell(3,3)  mill_pw(3,ngpw)  mill_rho(3,ngrho)  occupation(nstate)
!   atom_p_sp(nsp)  atco(3,na_max,nsp)
!   zv(nsp)  vps(ngrho,nsp)  rhos(ngrho,nsp)  ngh(nsp)  nghtol(nghmax,nsp)
!   twnl(ngpw,nghmax,nsp)  dmat(nghmax,nghmax,nsp)  wsg(nsp,nghmax)  c_0(ngpw,nstate) complex
! Output (stream): per step etotal eke eht epseu enl exc esr eself gemax cnorm,
!   then force(3,na_max,nsp), then c_0(ngpw,nstate).
program amdk_wfopt
  use iso_c_binding
  use amdk_hip
  use amdk_ref
  implicit none
  integer(c_int) :: hdr(13)
  integer :: n(3), symm, prec, xc, nst, nsp, namax, nghm, ngpw, ngrho, nstep, i, u
  real(c_double) :: a(3, 3)
  integer(c_int), allocatable, target :: mpw(:,:), mrho(:,:), ngh(:), nghtol(:,:), aps(:)
  real(c_double), allocatable, target :: occ(:), xat(:,:,:), zv(:), vps(:,:), rhos(:,:), twnl(:,:,:), dmat(:,:,:)
  real(c_double), allocatable :: wsg(:,:)
  complex(c_double_complex), allocatable :: c_in(:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  n = hdr(1:3); symm = hdr(4); prec = hdr(5); xc = hdr(6); nst = hdr(7); nsp = hdr(8); namax = hdr(9)
  nghm = hdr(10); ngpw = hdr(11); ngrho = hdr(12); nstep = hdr(13)
  allocate(mpw(3, ngpw), mrho(3, ngrho), occ(nst), aps(nsp), xat(3, namax, nsp))
  allocate(zv(nsp), vps(ngrho, nsp), rhos(ngrho, nsp), ngh(nsp), nghtol(nghm, nsp))
  allocate(twnl(ngpw, nghm, nsp), dmat(nghm, nghm, nsp), wsg(nsp, nghm), c_in(ngpw, nst))
  read(u) a, mpw, mrho, occ, aps, xat, zv, vps, rhos, ngh, nghtol, twnl, dmat, wsg, c_in
  close(u)

  call amdk_ref_init(n, a, symm, ngpw, mpw, ngrho, mrho, nst, prec, xc, nsp, aps, namax, nghm)
  call amdk_ref_species(zv, vps, rhos, ngh, nghtol, twnl, dmat, wsg)
  occupation = occ
  atco = xat
  call exp_igrxfactor()
  c_0 = c_in

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do i = 1, nstep
    ! do_opt.f90:82-112
    call eval_density()
    call rhor2g()
    call eval_pot()
    call cal_vpot_ex()
    call pp_energy()
    call e_total()
    c2 = (0.0d0, 0.0d0)
    call vpsi(c_0, c2, occupation, nstate)
    call fnonloc(c2, nstate, occupation)
    call forces()
    tfor = i == nstep
    if (tfor) then
      call e_sr()
      call potfor()
      call nlfor()
    end if
    call pcg(i)
    call ortho_gs(c_0)
    write(u) etotal, eke, eht, epseu, enl, exc, esr, eself, gemax, cnorm
    deallocate(nl)
  end do
  write(u) force, c_0
  close(u)
  call amdk_ref_finalize()
  write (*, '(A,I0,A)') 'amdk_wfopt: ', nstep, ' steps done'
end program amdk_wfopt
