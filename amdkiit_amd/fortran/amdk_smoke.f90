! Fortran smoke driver for the ISO_C_BINDING layer: builds a small cubic
! plane-wave basis, applies VPSI with a constant potential (closed form
! C2 = -(f/2)(|G|^2 + 2 V0) c), orthonormalises with ortho_gs and checks the
! overlap against the identity.  Prints PASS/FAIL and exits non-zero on FAIL.
program amdk_smoke
  use iso_c_binding
  use amdk_hip
  implicit none
  integer, parameter :: n = 24, nb = 5
  real(c_double), parameter :: alat = 10.0d0, ecut = 10.0d0, v0 = -0.37d0
  real(c_double), parameter :: pi = 3.14159265358979323846d0
  type(amdk_desc) :: d
  type(c_ptr) :: ctx
  integer(c_int), allocatable, target :: mill(:,:), millr(:,:)
  complex(c_double_complex), allocatable :: c(:,:), c2(:,:)
  real(c_double), allocatable :: g2(:), v(:), s(:,:), occ(:)
  integer :: ix, iy, iz, ng, ngr, i, b, mx
  real(c_double) :: gsq, tp, err, x
  integer(c_int) :: rc

  tp = 2.0d0 * pi / alat
  mx = n / 2 - 1
  allocate(mill(3, (2*mx+1)**3), millr(3, (2*mx+1)**3), g2((2*mx+1)**3))
  ng = 0; ngr = 0
  do ix = 0, mx
    do iy = -mx, mx
      do iz = -mx, mx
        if (ix == 0 .and. (iy < 0 .or. (iy == 0 .and. iz < 0))) cycle
        gsq = tp * tp * dble(ix*ix + iy*iy + iz*iz)
        if (gsq < 4.0d0 * ecut) then
          ngr = ngr + 1
          millr(:, ngr) = [ix, iy, iz]
        end if
        if (gsq < ecut) then
          ng = ng + 1
          mill(:, ng) = [ix, iy, iz]
          g2(ng) = gsq
        end if
      end do
    end do
  end do
  ! put G = 0 first (the reference orders by |G|^2)
  do i = 1, ng
    if (all(mill(:, i) == 0)) then
      mill(:, i) = mill(:, 1); mill(:, 1) = 0
      gsq = g2(i); g2(i) = g2(1); g2(1) = gsq
    end if
  end do
  d%n1 = n; d%n2 = n; d%n3 = n
  d%a_cell = 0.0d0
  d%a_cell(1) = alat; d%a_cell(5) = alat; d%a_cell(9) = alat
  d%ngpw = ng; d%miller_pw = c_loc(mill)
  d%ngrho = ngr; d%miller_rho = c_loc(millr)
  d%nbmax = nb; d%precision = AMDK_PREC_FP64; d%symmetry = 1
  rc = amdk_ctx_create(d, ctx)
  call amdk_check(ctx, rc, 'amdk_ctx_create')

  allocate(c(ng, nb), c2(ng, nb), v(n*n*n), s(nb, nb), occ(nb))
  occ = [2.0d0, 2.0d0, 2.0d0, 1.0d0, 0.0d0]
  do b = 1, nb
    do i = 1, ng
      x = sin(1.3d0 * i + 0.7d0 * b)
      c(i, b) = cmplx(x, cos(0.9d0 * i * b), c_double_complex) / (1.0d0 + g2(i))
    end do
    c(1, b) = cmplx(real(c(1, b)), 0.0d0, c_double_complex)
  end do
  v = v0
  call amdk_check(ctx, amdk_set_vks(ctx, v), 'set_vks')
  c2 = (0.0d0, 0.0d0)
  call amdk_check(ctx, amdk_vpsi(ctx, c, int(ng, c_int), c2, occ, int(nb, c_int)), 'VPSI')
  err = 0.0d0
  do b = 1, nb
    x = 0.5d0 * occ(b)
    if (x == 0.0d0) x = 1.0d0
    do i = 1, ng
      err = max(err, abs(c2(i, b) + x * (g2(i) + 2.0d0 * v0) * c(i, b)))
    end do
  end do
  write (*, '(A,ES10.3)') 'VPSI constant-potential max error: ', err
  if (err > 1.0d-12) then
    write (*, '(A)') 'FAIL'
    stop 1
  end if
  call amdk_check(ctx, amdk_ortho(ctx, c, int(ng, c_int), int(nb, c_int)), 'ortho_gs')
  call amdk_check(ctx, amdk_overlap(ctx, c, c, int(ng, c_int), int(nb, c_int), int(nb, c_int), s), 'OVLAP_D')
  err = 0.0d0
  do b = 1, nb
    s(b, b) = s(b, b) - 1.0d0
  end do
  err = maxval(abs(s))
  write (*, '(A,ES10.3)') 'ortho_gs max |S - I|: ', err
  if (err > 1.0d-12) then
    write (*, '(A)') 'FAIL'
    stop 1
  end if
  call amdk_ctx_destroy(ctx)
  write (*, '(A)') 'PASS'
end program amdk_smoke
