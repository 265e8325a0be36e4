!==============================================================================
! stab_gpu: drop-in driver for stab's Chebyshev temporal (itype 1) and spatial
! (itype 2) solves with the arithmetic on the GPU.
!
! Reads the same list-directed stdin sequence as the reference (input.f90:14-128
! then stab.f90:46-55: index, and x for spatial), reads profile.<ind>
! (getmean.f90:27-80: '#' lines skipped, v forced to zero) and writes evec.dat
! with the records of temporal.f90:883-890 / spatial.f90:1120-1126.
!
! One extension: the reference computes the whole spectrum, this driver the
! eigenpairs nearest a shift.  An optional line after the reference inputs,
!     shift_r  shift_i  [nev]
! gives that shift (the value the reference's getevec -v reads from its
! heredoc, thesis/TStest/run.sh:13-17) and the number of pairs (default 8).
! Without it the shift is 0.35 alpha (temporal) or omega / 0.35 (spatial).
! itype 7 is the multiple-temporal sweep of stab.f90:84-86 / mtemporal.f90:
! no alpha/beta lines and no index (profile.0), then the lines
! "alpha_min alpha_max alpha_inc" and "beta_min beta_max beta_inc"; the
! points (alpha outer, beta inner, mtemporal.f90:25-38) are solved in one
! batched call and each writes eig.<iver> with the evec.dat records.  For it
! the optional last line is a phase-speed guess "c_r c_i [nev]" (shift =
! alpha c; default c = 0.35).
! Record 4 holds the computed eigenvalues followed by huge placeholders, and
! record 5 their eigenvectors followed by zero columns, so both records keep
! the reference's length and getevec reads the file unchanged.
! Set STAB_MODE=mixed for the FP32-LU path.
!==============================================================================
program stab_gpu
  use iso_c_binding
  use stab_hip
  implicit none

  real(8), parameter :: pt5 = 0.5d0, gamma1 = 0.4d0, huge_ev = 1.0d300
  integer :: mattyp, ny, ievec, itmp, top, wall, wallt, curve, itype, ind, nev, ios
  real(8) :: T0, Te, Ma, Re, Pr, yi, ymax, x, ar, ai, br, bi, orr, oi, sr, si, rnev
  complex(8) :: alpha, beta, omega, shift
  type(stab_point) :: pt(1)
  type(c_ptr) :: ctx
  integer(c_int) :: rc, mode
  integer :: nym, n, nvec, nmax, i, k, na, nb, ia, ib, npts, iver
  real(8) :: amin, amax, ainc, bmin, bmax, binc
  complex(8) :: cguess
  type(stab_point), allocatable :: pts(:)
  complex(8), allocatable :: evs(:), vecs(:)
  real(8), allocatable :: ress(:)
  integer(c_int32_t), allocatable :: infs(:)
  real(8), allocatable, target :: ym(:), vt(:,:), y(:), eta(:), deta(:), d2eta(:)
  complex(8), allocatable :: eval(:), evec(:), evout(:), vecout(:,:)
  real(8), allocatable :: resid(:)
  integer(c_int32_t) :: info(1)
  character(len=80) :: fname
  character(len=32) :: envmode
  logical :: ider

  ! ---- input.f90 sequence ----------------------------------------------------
  T0 = 0.0d0; Te = 0.0d0; x = 0.0d0; Ma = 0.0d0
  read (*, *) mattyp
  if (mattyp /= 0 .and. mattyp /= 1) stop 'mattyp must be 0 or 1'
  if (mattyp == 1) then
    read (*, *) T0
    Te = T0 / (1.0d0 + pt5 * gamma1 * Ma**2)      ! Ma not read yet: Te = T0
  end if
  read (*, *) Ma, Re, Pr
  read (*, *) ny, yi, ymax
  read (*, *) ievec
  read (*, *) itmp
  ider = (itmp /= 0)
  read (*, *) top, wall, wallt, curve
  read (*, *) itype
  alpha = (0.0d0, 0.0d0); beta = (0.0d0, 0.0d0); omega = (0.0d0, 0.0d0)
  if (itype == 1) then
    read (*, *) ar, ai
    read (*, *) br, bi
    alpha = cmplx(ar, ai, kind=8); beta = cmplx(br, bi, kind=8)
  else if (itype == 2) then
    read (*, *) orr, oi
    read (*, *) br, bi
    omega = cmplx(orr, oi, kind=8); beta = cmplx(br, bi, kind=8)
  else if (itype == 7) then
    continue                                     ! input.f90:103-107: nothing more
  else
    write (*, *) 'itype ', itype, ' is outside the Chebyshev path of stab_gpu'
    call exit(2)
  end if
  ind = 0
  if (itype /= 7) read (*, *) ind
  if (itype == 2) read (*, *) x
  if (itype == 7) then
    read (*, *) amin, amax, ainc                 ! mtemporal.f90:20-23
    read (*, *) bmin, bmax, binc
  end if
  if (curve /= 0) then
    write (*, *) 'curve /= 0 (wall curvature) is outside this path'
    call exit(2)
  end if
  ! optional shift line
  nev = 8
  cguess = cmplx(0.35d0, 0.0d0, kind=8)
  if (itype == 7) then
    read (*, *, iostat=ios) sr, si, rnev
    if (ios == 0) then
      cguess = cmplx(sr, si, kind=8)
      nev = int(rnev)
    else
      read (*, *, iostat=ios) sr, si
      if (ios == 0) cguess = cmplx(sr, si, kind=8)
    end if
    nev = max(1, min(nev, STAB_MAX_NEV))
    call makename('profile', ind, fname)
    call read_profile(trim(fname), nym, ym, vt)
    call sweep_temporal()
    stop
  end if
  if (itype == 1) then
    shift = cmplx(0.35d0 * real(alpha), 0.0d0, kind=8)
  else
    shift = cmplx(real(omega) / 0.35d0, 0.0d0, kind=8)
  end if
  read (*, *, iostat=ios) sr, si, rnev
  if (ios == 0) then
    shift = cmplx(sr, si, kind=8)
    nev = int(rnev)
  else
    read (*, *, iostat=ios) sr, si
    if (ios == 0) shift = cmplx(sr, si, kind=8)
  end if
  nev = max(1, min(nev, STAB_MAX_NEV))

  ! ---- profile.<ind> (getmean.f90) --------------------------------------------
  call makename('profile', ind, fname)
  call read_profile(trim(fname), nym, ym, vt)

  ! ---- GPU solve ----------------------------------------------------------------
  rc = stab_ctx_create(ctx, 0)
  if (rc /= STAB_OK) then
    write (*, *) 'stab_gpu: no usable HIP device (status ', rc, ')'
    call exit(3)
  end if
  rc = stab_profile_set(ctx, 0, nym, ym, vt)
  call check(rc)
  pt(1)%itype = itype; pt(1)%ny = ny; pt(1)%mattyp = mattyp; pt(1)%wallt = wallt
  pt(1)%top = top; pt(1)%profile = 0; pt(1)%nev = nev; pt(1)%ider = merge(1, 0, ider)
  pt(1)%Ma = Ma; pt(1)%Re = Re; pt(1)%Pr = Pr; pt(1)%Te = Te; pt(1)%yi = yi; pt(1)%ymax = ymax
  pt(1)%alpha = alpha; pt(1)%beta = beta; pt(1)%omega = omega; pt(1)%shift = shift
  n = 5 * ny
  nvec = n
  if (itype == 2) nvec = 2 * n
  allocate (eval(STAB_MAX_NEV), evec(nvec * STAB_MAX_NEV), resid(STAB_MAX_NEV))
  call get_mode(mode)
  rc = stab_solve(ctx, 1, pt, mode, eval, evec, resid, info, c_null_ptr)
  call check(rc)
  allocate (y(ny), eta(ny), deta(ny), d2eta(ny))
  rc = stab_grid(ctx, pt(1), c_loc(y), c_loc(eta), c_loc(deta), c_loc(d2eta), c_null_ptr, c_null_ptr, c_null_ptr)
  call check(rc)
  rc = stab_ctx_destroy(ctx)

  do k = 1, nev
    write (*, '(i3,2x,1p,2(e22.14,1x),"  resid ",e9.2)') k, eval(k), resid(k)
  end do
  if (info(1) /= STAB_INFO_OK) write (*, *) 'stab_gpu: info = ', info(1)

  ! ---- evec.dat (temporal.f90:883-890) ---------------------------------------------
  nmax = nvec
  allocate (evout(nmax), vecout(nmax, nmax))
  evout = cmplx(huge_ev, 0.0d0, kind=8)
  vecout = (0.0d0, 0.0d0)
  evout(1:nev) = eval(1:nev)
  do k = 1, nev
    vecout(:, k) = evec((k - 1) * nvec + 1:k * nvec)
  end do
  open (unit=77, file='evec.dat', form='unformatted', status='unknown')
  write (77) ind, ny, 5, itype, ievec, curve, top, wall, wallt, ider
  write (77) omega, alpha, beta, Re, Ma, Pr
  write (77) x, y, eta, deta, d2eta, yi, ymax
  write (77) evout
  if (itype == 1 .or. ievec == 1) write (77) vecout
  close (77)

contains

  ! mtemporal.f90:25-38 batched: all (alpha, beta) points in one stab_solve,
  ! then eig.<iver> per point with the records of temporal.f90:883-890.
  subroutine sweep_temporal()
    na = max(nint((amax - amin) / ainc), 1)
    nb = max(nint((bmax - bmin) / binc), 1)
    npts = na * nb
    allocate (pts(npts))
    n = 5 * ny
    iver = 0
    do ia = 1, na
      do ib = 1, nb
        iver = iver + 1
        pts(iver)%itype = 1; pts(iver)%ny = ny; pts(iver)%mattyp = mattyp; pts(iver)%wallt = wallt
        pts(iver)%top = top; pts(iver)%profile = 0; pts(iver)%nev = nev; pts(iver)%ider = merge(1, 0, ider)
        pts(iver)%Ma = Ma; pts(iver)%Re = Re; pts(iver)%Pr = Pr; pts(iver)%Te = Te
        pts(iver)%yi = yi; pts(iver)%ymax = ymax
        pts(iver)%alpha = cmplx(amin + dble(ia - 1) * ainc, 0.0d0, kind=8)
        pts(iver)%beta = cmplx(bmin + dble(ib - 1) * binc, 0.0d0, kind=8)
        pts(iver)%omega = (0.0d0, 0.0d0)
        pts(iver)%shift = pts(iver)%alpha * cguess
      end do
    end do
    rc = stab_ctx_create(ctx, 0)
    if (rc /= STAB_OK) then
      write (*, *) 'stab_gpu: no usable HIP device (status ', rc, ')'
      call exit(3)
    end if
    rc = stab_profile_set(ctx, 0, nym, ym, vt)
    call check(rc)
    allocate (evs(STAB_MAX_NEV * npts), vecs(n * STAB_MAX_NEV * npts), ress(STAB_MAX_NEV * npts), infs(npts))
    call get_mode(mode)
    rc = stab_solve(ctx, npts, pts, mode, evs, vecs, ress, infs, c_null_ptr)
    call check(rc)
    allocate (y(ny), eta(ny), deta(ny), d2eta(ny))
    rc = stab_grid(ctx, pts(1), c_loc(y), c_loc(eta), c_loc(deta), c_loc(d2eta), c_null_ptr, c_null_ptr, c_null_ptr)
    call check(rc)
    rc = stab_ctx_destroy(ctx)
    allocate (evout(n), vecout(n, n))
    do iver = 1, npts
      write (*, '(2(1x,i3,1x),4(1pe13.6,1x))') (iver - 1) / nb + 1, mod(iver - 1, nb) + 1, &
        real(pts(iver)%alpha), aimag(pts(iver)%alpha), real(pts(iver)%beta), aimag(pts(iver)%beta)
      evout = cmplx(huge_ev, 0.0d0, kind=8)
      vecout = (0.0d0, 0.0d0)
      do k = 1, nev
        evout(k) = evs((iver - 1) * STAB_MAX_NEV + k)
        vecout(:, k) = vecs((iver - 1) * n * STAB_MAX_NEV + (k - 1) * n + 1:(iver - 1) * n * STAB_MAX_NEV + k * n)
      end do
      if (infs(iver) /= STAB_INFO_OK) write (*, *) 'stab_gpu: point ', iver, ' info = ', infs(iver)
      call makename('eig', iver, fname)
      open (unit=77, file=trim(fname), form='unformatted', status='unknown')
      write (77) ind, ny, 5, 1, ievec, curve, top, wall, wallt, ider
      write (77) (0.0d0, 0.0d0), pts(iver)%alpha, pts(iver)%beta, Re, Ma, Pr
      write (77) x, y, eta, deta, d2eta, yi, ymax
      write (77) evout
      write (77) vecout
      close (77)
    end do
  end subroutine sweep_temporal

  subroutine check(code)
    integer(c_int), intent(in) :: code
    if (code /= STAB_OK) then
      write (*, *) 'stab_gpu: ', trim(stab_error_text(ctx)), ' (status ', code, ')'
      call exit(1)
    end if
  end subroutine check

  subroutine get_mode(m)
    integer(c_int), intent(out) :: m
    integer :: l
    m = STAB_MODE_PLAIN_F64
    call get_environment_variable('STAB_MODE', envmode, l)
    if (l > 0) then
      if (envmode(1:5) == 'mixed') m = STAB_MODE_MIXED
    end if
  end subroutine get_mode

  ! makename (mtemporal.f90:53-76): base.<iver>
  subroutine makename(base, iver, name)
    character(*), intent(in) :: base
    integer, intent(in) :: iver
    character(len=80), intent(out) :: name
    character(len=16) :: num
    write (num, '(i0)') iver
    name = trim(base)//'.'//trim(num)
  end subroutine makename

  ! Rows "y rho u v w T"; lines starting with '#' skipped; v forced to zero.
  subroutine read_profile(path, nrow, yy, vv)
    character(*), intent(in) :: path
    integer, intent(out) :: nrow
    real(8), allocatable, intent(out) :: yy(:), vv(:,:)
    character(len=512) :: line
    integer :: u, io, j
    real(8) :: row(6)
    open (newunit=u, file=path, form='formatted', status='old', iostat=io)
    if (io /= 0) then
      write (*, *) 'Error reading profile...'
      call exit(1)
    end if
    nrow = 0
    do
      read (u, '(a)', iostat=io) line
      if (io /= 0) exit
      if (line(1:1) /= '#' .and. len_trim(line) > 0) nrow = nrow + 1
    end do
    rewind (u)
    allocate (yy(nrow), vv(nrow, 5))
    j = 0
    do
      read (u, '(a)', iostat=io) line
      if (io /= 0) exit
      if (line(1:1) == '#' .or. len_trim(line) == 0) cycle
      j = j + 1
      read (line, *) row
      yy(j) = row(1)
      vv(j, :) = row(2:6)
      vv(j, 3) = 0.0d0
    end do
    close (u)
  end subroutine read_profile

end program stab_gpu
