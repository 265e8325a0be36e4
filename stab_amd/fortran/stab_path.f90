!==============================================================================
! The Chebyshev path of stab as Fortran subroutines over libstab_hip.
!
! Module `stuff` carries the run parameters under the names and meanings of
! the reference's module (stuff.f90:9-55), so that callers written against
! the reference -- its main program (stab.f90:49,55) and its sweep drivers
! (mtemporal.f90:37, mspatial.f90:90) -- compile and link here unchanged.
! The subroutines keep the reference's call surface:
!   input                  the list-directed stdin sequence of input.f90:14-128
!   temporal(name, ind)    temporal.f90:2   (profile.<ind> in, file `name` out)
!   spatial(name, ind)     spatial.f90:2
!   mtemporal(ind)         mtemporal.f90:2  (every (alpha, beta) point in one batched GPU call)
!   mspatial(ind)          mspatial.f90:2   (delta.dat stations x omega x beta, one batched call)
!   makename(base, iver, fname)   mtemporal.f90:53-76
! The arithmetic runs on the GPU through the C ABI (stab_hip_iface.f90).
!
! What differs from the reference, by design: the reference computes the
! whole spectrum with LAPACK; these routines compute the stab_nev eigenpairs
! nearest a shift.  The shift is module stab_gpu_opts' stab_shift when
! stab_have_shift is set (the value the reference's user gives getevec -v),
! else 0.35 alpha (temporal) or omega / 0.35 (spatial).  The output file keeps
! the reference's record lengths (n or 2n eigenvalues and vectors,
! temporal.f90:883-890, spatial.f90:1120-1126): the computed pairs occupy the
! last stab_nev slots in ascending order of their imaginary parts (the order
! PIKSR2 gives the full spectrum, temporal.f90:844-855), so index nmax is the
! computed pair of largest Im; the slots before them hold (1e300, -1e300),
! which sort first, with zero eigenvectors.
!==============================================================================
module stuff
  implicit none
  ! run parameters (stuff.f90:11-17)
  real(8) :: Ma = 0.0d0, Re = 0.0d0, Pr = 0.0d0, T0 = 0.0d0
  complex(8) :: alpha = (0.0d0, 0.0d0), beta = (0.0d0, 0.0d0), omega = (0.0d0, 0.0d0)
  real(8) :: alphar = 0.0d0, alphai = 0.0d0, betar = 0.0d0, betai = 0.0d0, omegar = 0.0d0, omegai = 0.0d0
  integer :: itype = 1, ievec = 1
  logical :: ider = .true., Navier = .true.
  integer :: top = 0, wall = 0, wallt = 0
  integer :: curve = 0
  ! edge conditions (stuff.f90:26-27); Te as input.f90:24-25 leaves it
  real(8) :: Te = 0.0d0
  ! constants (stuff.f90:31-42)
  real(8), parameter :: zero = 0.0d0, pt5 = 0.5d0, one = 1.0d0, two = 2.0d0
  real(8), parameter :: pi = 3.1415926535897932385d0
  complex(8), parameter :: im = (0.0d0, 1.0d0)
  ! fluid properties (stuff.f90:46-50)
  real(8) :: gamma = 1.4d0, gamma1 = 0.4d0, cv = 716.5d0, cp = 1003.1d0
  integer :: mattyp = 0
  ! problem dimensions (stuff.f90:54-55)
  integer :: ny = 1, nx = 1, nz = 1, nsd = 3, ndof = 5
  real(8) :: x = 0.0d0, ymin = 0.0d0, ymax = 0.0d0, yi = 0.0d0
end module stuff

module stab_gpu_opts
  use iso_c_binding
  implicit none
  ! eigenpairs per point and the shift (see the file header)
  integer :: stab_nev = 8
  logical :: stab_have_shift = .false.
  complex(8) :: stab_shift = (0.0d0, 0.0d0)
  ! phase-speed guess of the sweep drivers: shift = alpha c (temporal) or omega / c (spatial)
  complex(8) :: stab_phase_guess = (0.35d0, 0.0d0)
  ! STAB_MODE_PLAIN_F64 or STAB_MODE_MIXED (environment STAB_MODE=mixed)
  integer(c_int) :: stab_mode = 0
  logical :: stab_mode_read = .false.
  type(c_ptr) :: stab_ctx = c_null_ptr
  real(8), parameter :: stab_placeholder_re = 1.0d300, stab_placeholder_im = -1.0d300
end module stab_gpu_opts

module stab_path_util
  use iso_c_binding
  use stab_hip
  use stab_gpu_opts
  implicit none
contains

  subroutine stab_check(code)
    integer(c_int), intent(in) :: code
    if (code /= STAB_OK) then
      write (*, *) 'stab: ', trim(stab_error_text(stab_ctx)), ' (status ', code, ')'
      call exit(1)
    end if
  end subroutine stab_check

  !> The GPU context of this process, created on first use.
  subroutine stab_context()
    integer(c_int) :: rc
    character(len=32) :: env
    integer :: l
    if (.not. stab_mode_read) then
      call get_environment_variable('STAB_MODE', env, l)
      if (l >= 5) then
        if (env(1:5) == 'mixed') stab_mode = STAB_MODE_MIXED
      end if
      stab_mode_read = .true.
    end if
    if (c_associated(stab_ctx)) return
    rc = stab_ctx_create(stab_ctx, 0)
    if (rc /= STAB_OK) then
      write (*, *) 'stab: no usable HIP device (status ', rc, ')'
      call exit(3)
    end if
  end subroutine stab_context

  !> Rows "y rho u v w T" of a profile file; '#' lines skipped, v forced to
  !> zero (getmean.f90:27-80).
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

  !> Profile table of profile.<ind> (read before any GPU call, so a missing
  !> file is reported as the reference reports it).
  subroutine fetch_profile(ind, nym, ym, vt)
    integer, intent(in) :: ind
    integer, intent(out) :: nym
    real(8), allocatable, intent(out) :: ym(:), vt(:,:)
    character(len=80) :: fname
    call makename('profile', ind, fname)
    call read_profile(trim(fname), nym, ym, vt)
  end subroutine fetch_profile

  !> Send a profile table to slot `slot` of the context.
  subroutine send_profile(slot, nym, ym, vt)
    integer, intent(in) :: slot, nym
    real(8), intent(in) :: ym(:), vt(:,:)
    call stab_check(stab_profile_set(stab_ctx, int(slot, c_int), int(nym, c_int), ym, vt))
  end subroutine send_profile

  !> Sort the first k values by ascending imaginary part, carrying the
  !> permutation (insertion sort; PIKSR2's order, temporal.f90:844-855).
  subroutine sort_by_imag(k, ev, perm)
    integer, intent(in) :: k
    complex(8), intent(inout) :: ev(:)
    integer, intent(out) :: perm(:)
    integer :: i, j, pj
    complex(8) :: v
    do i = 1, k
      perm(i) = i
    end do
    do j = 2, k
      v = ev(j)
      pj = perm(j)
      i = j - 1
      do while (i >= 1)
        if (aimag(ev(i)) <= aimag(v)) exit
        ev(i + 1) = ev(i)
        perm(i + 1) = perm(i)
        i = i - 1
      end do
      ev(i + 1) = v
      perm(i + 1) = pj
    end do
  end subroutine sort_by_imag

  !> Write one eigensystem file with the records of temporal.f90:883-890 /
  !> spatial.f90:1120-1126 from the computed pairs (see the file header).
  subroutine write_eigensystem(name, ind, nyy, itp, nvec, k, ev, vec, om, al, be, xx, y, eta, deta, d2eta, yii, ymx)
    use stuff, only: ievec, curve, top, wall, wallt, ider, Re, Ma, Pr
    character(*), intent(in) :: name
    integer, intent(in) :: ind, nyy, itp, nvec, k
    complex(8), intent(in) :: ev(:), vec(:), om, al, be
    real(8), intent(in) :: xx, y(:), eta(:), deta(:), d2eta(:), yii, ymx
    complex(8), allocatable :: evout(:), vecout(:,:), es(:)
    integer, allocatable :: perm(:)
    integer :: j, u, slot
    allocate (evout(nvec), vecout(nvec, nvec), es(k), perm(k))
    evout = cmplx(stab_placeholder_re, stab_placeholder_im, kind=8)
    vecout = (0.0d0, 0.0d0)
    es = ev(1:k)
    call sort_by_imag(k, es, perm)
    do j = 1, k
      slot = nvec - k + j
      evout(slot) = es(j)
      vecout(:, slot) = vec((perm(j) - 1) * nvec + 1:perm(j) * nvec)
    end do
    open (newunit=u, file=name, form='unformatted', status='unknown')
    write (u) ind, nyy, 5, itp, ievec, curve, top, wall, wallt, ider
    write (u) om, al, be, Re, Ma, Pr
    write (u) xx, y, eta, deta, d2eta, yii, ymx
    write (u) evout
    if (itp == 1 .or. ievec == 1) write (u) vecout
    close (u)
  end subroutine write_eigensystem

  !> A stab_point from module stuff's current values.
  function point_from_stuff(itp, slot, shift, nev) result(p)
    use stuff
    integer, intent(in) :: itp, slot, nev
    complex(8), intent(in) :: shift
    type(stab_point) :: p
    p%itype = itp; p%ny = ny; p%mattyp = mattyp; p%wallt = wallt; p%top = top
    p%profile = slot; p%nev = nev; p%ider = merge(1, 0, ider)
    p%Ma = Ma; p%Re = Re; p%Pr = Pr; p%Te = Te; p%yi = yi; p%ymax = ymax
    p%alpha = alpha; p%beta = beta; p%omega = omega; p%shift = shift
  end function point_from_stuff

  !> One temporal point (itype 1 Chebyshev, 3 finite differences) from module
  !> stuff and profile.<ind>, solved and written to `name`.
  subroutine solve_temporal_point(name, ind, itp)
    use stuff
    character(*), intent(in) :: name
    integer, intent(in) :: ind, itp
    type(stab_point) :: pts(1)
    character(len=80) :: names(1)
    complex(8) :: sh
    integer :: nym
    real(8), allocatable :: ym(:), vt(:,:)
    if (curve /= 0) then
      write (*, *) 'curve /= 0 (wall curvature) is outside this path'
      call exit(2)
    end if
    call fetch_profile(ind, nym, ym, vt)
    call stab_context()
    call send_profile(0, nym, ym, vt)
    sh = cmplx(0.35d0 * real(alpha), 0.0d0, kind=8)
    if (stab_have_shift) sh = stab_shift
    pts(1) = point_from_stuff(itp, 0, sh, max(1, min(stab_nev, STAB_MAX_NEV)))
    names(1) = name
    call solve_and_write(1, pts, names, [ind], [x])
  end subroutine solve_temporal_point

  !> One spatial point (itype 2 Chebyshev, 4 finite differences) from module
  !> stuff and profile.<ind>, solved and written to `name`.
  subroutine solve_spatial_point(name, ind, itp)
    use stuff
    character(*), intent(in) :: name
    integer, intent(in) :: ind, itp
    type(stab_point) :: pts(1)
    character(len=80) :: names(1)
    complex(8) :: sh
    integer :: nym
    real(8), allocatable :: ym(:), vt(:,:)
    if (curve /= 0) then
      write (*, *) 'curve /= 0 (wall curvature) is outside this path'
      call exit(2)
    end if
    call fetch_profile(ind, nym, ym, vt)
    call stab_context()
    call send_profile(0, nym, ym, vt)
    sh = cmplx(real(omega) / 0.35d0, 0.0d0, kind=8)
    if (stab_have_shift) sh = stab_shift
    pts(1) = point_from_stuff(itp, 0, sh, max(1, min(stab_nev, STAB_MAX_NEV)))
    names(1) = name
    call solve_and_write(1, pts, names, [ind], [x])
  end subroutine solve_spatial_point

  !> Solve a batch and write one eigensystem file per point.
  subroutine solve_and_write(npts, pts, names, inds, xs)
    integer, intent(in) :: npts
    type(stab_point), intent(inout) :: pts(:)
    character(len=80), intent(in) :: names(:)
    integer, intent(in) :: inds(:)
    real(8), intent(in) :: xs(:)
    complex(8), allocatable, target :: evs(:), vecs(:)
    real(8), allocatable :: ress(:)
    integer(c_int32_t), allocatable :: infs(:)
    real(8), allocatable, target :: y(:), eta(:), deta(:), d2eta(:)
    integer :: i, nvec, off, off2, nyy
    integer(c_size_t) :: nbytes
    nbytes = stab_evec_bytes(int(npts, c_int), pts)
    allocate (evs(STAB_MAX_NEV * npts), vecs(nbytes / 16), ress(STAB_MAX_NEV * npts), infs(npts))
    call stab_check(stab_solve(stab_ctx, int(npts, c_int), pts, stab_mode, evs, vecs, ress, infs, c_null_ptr))
    off = 0
    do i = 1, npts
      nyy = pts(i)%ny
      nvec = 5 * nyy
      if (pts(i)%itype == 2 .or. pts(i)%itype == 4) nvec = 2 * nvec
      allocate (y(nyy), eta(nyy), deta(nyy), d2eta(nyy))
      call stab_check(stab_grid(stab_ctx, pts(i), c_loc(y), c_loc(eta), c_loc(deta), c_loc(d2eta), c_null_ptr, &
                                c_null_ptr, c_null_ptr))
      if (npts == 1) then
        do off2 = 1, pts(i)%nev
          write (*, '(i3,2x,1p,2(e22.14,1x),"  resid ",e9.2)') off2, evs(off2), ress(off2)
        end do
      end if
      if (infs(i) /= STAB_INFO_OK) write (*, *) 'stab: point ', i, ' info = ', infs(i)
      ! fd_temporal / fd_spatial write their files with itype 1 / 2 (stab.f90:59,66)
      call write_eigensystem(trim(names(i)), inds(i), nyy, merge(int(pts(i)%itype) - 2, int(pts(i)%itype), &
                             pts(i)%itype >= 3), &
                             nvec, int(pts(i)%nev), &
                             evs((i - 1) * STAB_MAX_NEV + 1:(i - 1) * STAB_MAX_NEV + pts(i)%nev), &
                             vecs(off + 1:off + nvec * STAB_MAX_NEV), pts(i)%omega, pts(i)%alpha, pts(i)%beta, &
                             xs(i), y, eta, deta, d2eta, pts(i)%yi, pts(i)%ymax)
      off = off + nvec * STAB_MAX_NEV
      deallocate (y, eta, deta, d2eta)
    end do
  end subroutine solve_and_write

end module stab_path_util

!> makename (mtemporal.f90:53-76): base.<iver>
subroutine makename(base, iver, fname)
  implicit none
  character(*), intent(in) :: base
  integer, intent(in) :: iver
  character(len=80), intent(out) :: fname
  character(len=16) :: num
  write (num, '(i0)') iver
  fname = trim(base)//'.'//trim(num)
end subroutine makename

!> input.f90:14-128: the run parameters from stdin into module stuff.
subroutine input
  use stuff
  implicit none
  integer :: itmp
  do
    read (*, *) mattyp
    if (mattyp == 0 .or. mattyp == 1) exit
  end do
  if (mattyp == 1) then
    read (*, *) T0
    Te = T0 / (one + pt5 * gamma1 * Ma**2)       ! Ma is not read yet: Te = T0 (input.f90:24-25)
  end if
  read (*, *) Ma, Re, Pr
  Navier = .not. (Re >= 1.0d98 .or. Re == 0.0d0)
  read (*, *) ny, yi, ymax
  read (*, *) ievec
  read (*, *) itmp
  ider = (itmp /= 0)
  read (*, *) top, wall, wallt, curve
  ymin = zero
  do
    read (*, *) itype
    select case (itype)
    case (1, 3)
      read (*, *) alphar, alphai
      read (*, *) betar, betai
      exit
    case (2, 4)
      read (*, *) omegar, omegai
      read (*, *) betar, betai
      exit
    case (7)
      alphar = zero; alphai = zero; betar = zero; betai = zero
      exit
    case (8)
      omegar = zero; omegai = zero; alphar = zero; alphai = zero; betar = zero; betai = zero
      exit
    case (5, 6, 9)
      write (*, *) 'itype ', itype, ' (Stokes / bump receptivity) is outside the Chebyshev path'
      call exit(2)
    end select
  end do
  alpha = cmplx(alphar, alphai, kind=8)
  beta = cmplx(betar, betai, kind=8)
  omega = cmplx(omegar, omegai, kind=8)
end subroutine input

!> temporal(name, ind) (temporal.f90:2): eigenpairs of the temporal pencil
!> at module stuff's (alpha, beta) on profile.<ind>, written to `name`.
subroutine temporal(name, ind)
  use stab_path_util
  implicit none
  character(*), intent(in) :: name
  integer, intent(in) :: ind
  call solve_temporal_point(name, ind, 1)
end subroutine temporal

!> fd_temporal(name, ind) (fd_temporal.f90:2): the same with the
!> finite-difference discretisation (ogengrid grid, explicit stencils).
subroutine fd_temporal(name, ind)
  use stab_path_util
  implicit none
  character(*), intent(in) :: name
  integer, intent(in) :: ind
  call solve_temporal_point(name, ind, 3)
end subroutine fd_temporal

!> spatial(name, ind) (spatial.f90:2): eigenpairs of L(alpha) at module
!> stuff's (omega, beta) on profile.<ind>, written to `name`.
subroutine spatial(name, ind)
  use stab_path_util
  implicit none
  character(*), intent(in) :: name
  integer, intent(in) :: ind
  call solve_spatial_point(name, ind, 2)
end subroutine spatial

!> fd_spatial(name, ind) (fd_spatial.f90:2): the same with the
!> finite-difference discretisation (ogengrid grid, explicit stencils).
subroutine fd_spatial(name, ind)
  use stab_path_util
  implicit none
  character(*), intent(in) :: name
  integer, intent(in) :: ind
  call solve_spatial_point(name, ind, 4)
end subroutine fd_spatial

!> mtemporal(ind) (mtemporal.f90:2-50): the (alpha, beta) sweep, every point
!> in one batched GPU solve, eig.<iver> per point; shift alpha * stab_phase_guess.
subroutine mtemporal(ind)
  use stuff
  use stab_path_util
  implicit none
  integer, intent(in) :: ind
  real(8) :: amin, amax, ainc, bmin, bmax, binc
  integer :: ia, ib, na, nb, iver, npts, nym
  real(8), allocatable :: ym(:), vt(:,:)
  type(stab_point), allocatable :: pts(:)
  character(len=80), allocatable :: names(:)
  integer, allocatable :: inds(:)
  real(8), allocatable :: xs(:)
  read (*, *) amin, amax, ainc
  read (*, *) bmin, bmax, binc
  call read_guess_line()
  na = max(nint((amax - amin) / ainc), 1)
  nb = max(nint((bmax - bmin) / binc), 1)
  npts = na * nb
  allocate (pts(npts), names(npts), inds(npts), xs(npts))
  call fetch_profile(ind, nym, ym, vt)
  call stab_context()
  call send_profile(0, nym, ym, vt)
  iver = 0
  do ia = 1, na
    do ib = 1, nb
      iver = iver + 1
      alpha = amin + dble(ia - 1) * ainc
      beta = bmin + dble(ib - 1) * binc
      call makename('eig', iver, names(iver))
      write (*, '(2(1x,i3,1x),4(1pe13.6,1x))') ia, ib, real(alpha), aimag(alpha), real(beta), aimag(beta)
      pts(iver) = point_from_stuff(1, 0, alpha * stab_phase_guess, max(1, min(stab_nev, STAB_MAX_NEV)))
      inds(iver) = ind
      xs(iver) = x
    end do
  end do
  call solve_and_write(npts, pts, names, inds, xs)
end subroutine mtemporal

!> mspatial(ind) (mspatial.f90:2-109): stations ind1..ind2 of delta.dat
!> (x, delta; '#' lines skipped), Yi = 2 delta, profile.<station> each, an
!> omega x beta sweep per station; all points in one batched GPU solve,
!> eig.<iver> per point; shift omega / stab_phase_guess.
subroutine mspatial(ind)
  use stuff
  use stab_path_util
  implicit none
  integer, intent(inout) :: ind      ! the reference loops on it; it ends past the last station
  integer :: nym
  real(8), allocatable :: ym(:), vt(:,:)
  real(8) :: omin, omax, oinc, bmin, bmax, binc
  integer :: ind1, ind2, ind_inc, dtype, nbody, io, ib, no, nb, iver, npts, nst, k, u, ios, slot, itp
  real(8), allocatable :: xb(:), delta(:), xs(:)
  character(len=256) :: line
  type(stab_point), allocatable :: pts(:)
  character(len=80), allocatable :: names(:)
  integer, allocatable :: inds(:)
  read (*, *) omin, omax, oinc
  read (*, *) bmin, bmax, binc
  read (*, *) ind1, ind2, ind_inc
  read (*, *) dtype
  call read_guess_line()
  itp = merge(2, 4, dtype == 0)           ! mspatial.f90:85-89: spatial, or fd_spatial for dtype /= 0
  open (newunit=u, file='delta.dat', status='old', iostat=ios)
  if (ios /= 0) then
    write (*, *) 'ERROR: cannot open delta.dat'
    call exit(1)
  end if
  nbody = 0
  do
    read (u, '(a)', iostat=ios) line
    if (ios /= 0) exit
    if (line(1:1) /= '#') nbody = nbody + 1
  end do
  if (ind1 < 1 .or. ind2 > nbody) then
    write (*, *) ind1, ind2, 1, nbody
    write (*, "('ERROR: illegal index...check delta.dat')")
    call exit(1)
  end if
  allocate (xb(nbody), delta(nbody))
  rewind (u)
  k = 0
  do
    read (u, '(a)', iostat=ios) line
    if (ios /= 0) exit
    if (line(1:1) == '#') cycle
    k = k + 1
    read (line, *) xb(k), delta(k)
  end do
  close (u)
  if (oinc == zero) oinc = one
  if (binc == zero) binc = one
  if (ind_inc == 0) ind_inc = 1
  no = nint((omax - omin) / oinc) + 1
  nb = nint((bmax - bmin) / binc) + 1
  nst = 0
  do ind = ind1, ind2, ind_inc
    nst = nst + 1
  end do
  npts = nst * no * nb
  allocate (pts(npts), names(npts), inds(npts), xs(npts))
  do ind = ind1, ind2, ind_inc            ! every profile file exists before the GPU is touched
    call fetch_profile(ind, nym, ym, vt)
  end do
  call stab_context()
  iver = 0
  slot = 0
  write (*, "(/,95('='),/,'Index',' iver','      s       ','     Y_i      ','    omega_r   ','    omega_i   '," // &
            "'    beta_r    ','    beta_i    ',/,95('='))")
  do ind = ind1, ind2, ind_inc
    call fetch_profile(ind, nym, ym, vt)
    call send_profile(slot, nym, ym, vt)
    x = xb(ind)
    yi = two * delta(ind)
    do io = 1, no
      do ib = 1, nb
        iver = iver + 1
        omega = omin + dble(io - 1) * oinc
        beta = bmin + dble(ib - 1) * binc
        call makename('eig', iver, names(iver))
        write (*, '(2(1x,i3,1x),6(1pe13.6,1x))') ind, iver, x, yi, real(omega), aimag(omega), real(beta), &
          aimag(beta)
        pts(iver) = point_from_stuff(itp, slot, omega / stab_phase_guess, max(1, min(stab_nev, STAB_MAX_NEV)))
        inds(iver) = ind
        xs(iver) = x
      end do
    end do
    slot = slot + 1
  end do
  call solve_and_write(npts, pts, names, inds, xs)
end subroutine mspatial

!> Optional line after a sweep's inputs: "c_r c_i [nev]", the phase-speed
!> guess and pair count of this library (no reference counterpart).
subroutine read_guess_line()
  use stab_gpu_opts
  implicit none
  real(8) :: sr, si, rnev
  integer :: ios
  character(len=256) :: line
  read (*, '(a)', iostat=ios) line
  if (ios /= 0) return
  read (line, *, iostat=ios) sr, si, rnev
  if (ios == 0) then
    stab_phase_guess = cmplx(sr, si, kind=8)
    stab_nev = int(rnev)
    return
  end if
  read (line, *, iostat=ios) sr, si
  if (ios == 0) stab_phase_guess = cmplx(sr, si, kind=8)
end subroutine read_guess_line
