!==============================================================================
! stab_gpu: the main program of stab (stab.f90:42-101) for the Chebyshev path,
! with the arithmetic on the GPU (stab_path.f90, stab_hip_iface.f90).
!
! Reads the same list-directed stdin sequence as the reference (input.f90:14-128,
! then stab.f90:46-55: index, and x for spatial) and dispatches on itype:
!   1  temporal('evec.dat', ind)      2  spatial('evec.dat', ind)
!   3  fd_temporal('evec.dat', ind)   4  fd_spatial('evec.dat', ind)
!   7  mtemporal(ind)                 8  mspatial(ind)
! Stokes waves (5) and the bump receptivity drivers (6, 9) are not part of
! this build and exit with status 2.
!
! One extension: after the reference inputs of itype 1 to 4 an optional line
!     shift_r  shift_i  [nev]
! gives the shift (the value the reference's getevec -v reads, thesis/TStest/
! run.sh:13-17) and the number of pairs (default 8); for itype 7 and 8 the
! optional line after the sweep inputs is a phase-speed guess "c_r c_i [nev]"
! (shift = alpha c, or omega / c).  Set STAB_MODE=mixed for the FP32-LU path.
!==============================================================================
program stab_gpu
  use stuff
  use stab_gpu_opts
  implicit none
  character(len=80) :: name = 'evec.dat'
  integer :: ind = 0

  call input
  select case (itype)
  case (1)
    read (*, *) ind
    call read_shift_line()
    call temporal(name, ind)
  case (2)
    read (*, *) ind
    read (*, *) x
    call read_shift_line()
    call spatial(name, ind)
  case (3)
    read (*, *) ind
    itype = 1
    call read_shift_line()
    call fd_temporal(name, ind)
  case (4)
    read (*, *) ind
    read (*, *) x
    itype = 2
    call read_shift_line()
    call fd_spatial(name, ind)
  case (7)
    itype = 1
    call mtemporal(ind)
  case (8)
    itype = 2
    call mspatial(ind)
  case default
    write (*, *) 'itype ', itype, ' is outside the Chebyshev path of stab_gpu'
    call exit(2)
  end select

contains

  subroutine read_shift_line()
    real(8) :: sr, si, rnev
    integer :: ios
    character(len=256) :: line
    read (*, '(a)', iostat=ios) line
    if (ios /= 0) return
    read (line, *, iostat=ios) sr, si, rnev
    if (ios == 0) then
      stab_shift = cmplx(sr, si, kind=8)
      stab_have_shift = .true.
      stab_nev = int(rnev)
      return
    end if
    read (line, *, iostat=ios) sr, si
    if (ios == 0) then
      stab_shift = cmplx(sr, si, kind=8)
      stab_have_shift = .true.
    end if
  end subroutine read_shift_line

end program stab_gpu
