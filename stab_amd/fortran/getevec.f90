!==============================================================================
! getevec: eigenfunction extraction from an eigensystem file, the command of
! getevec.f90:2-239 written for this build (host only; no GPU).
!
! Usage: getevec [-i|-v|-h] [file]      (default file: prompt, as the reference)
! Reads the records of temporal.f90:883-890 / spatial.f90:1120-1126 (nmax = 5 ny
! or 10 ny values), prints the eigenvalue table, then repeatedly reads a
! selection from stdin -- an index (-i, default; 0 ends, -1 reprints) or an
! approximate eigenvalue "re im" (-v; the first value of least distance,
! getevec.f90:160-173; 0 0 ends) -- normalises that eigenvector by its first
! entry of largest modulus over the 5 ny flow components (getevec.f90:179-189)
! and writes time.<k> or space.<k>: four '#' lines, then rows from the top
! node down as y and the real/imaginary parts of rho, u, v, w, T in
! 1p,11(1pe21.13E3,1x) (getevec.f90:193-227).  Selection ends at end of input.
!==============================================================================
program getevec
  implicit none
  integer :: nyy, ndof, itp, iev, icurve, itop, iwall, iwallt, ider
  real(8), allocatable :: y(:), eta(:), deta(:), d2eta(:)
  complex(8), allocatable :: ev(:), vec(:,:)
  complex(8) :: om, al, be, k, val, sc
  real(8) :: xx, Re, Ma, Pr, yii, ymx, vr, vi, d
  character(len=256) :: fname, arg, base, outname
  integer :: narg, ia, nmax, j, i, i0, u, ios, iver, ind
  logical :: by_index

  by_index = .true.
  fname = ''
  narg = command_argument_count()
  do ia = 1, narg
    call get_command_argument(ia, arg)
    if (arg(1:1) /= '-') then
      if (len_trim(fname) > 0) then
        write (*, *) 'Only one filename argument allowed'
        call exit(1)
      end if
      fname = arg
    else if (arg(1:2) == '-i') then
      by_index = .true.
    else if (arg(1:2) == '-v') then
      by_index = .false.
    else if (arg(1:2) == '-h') then
      write (*, '(a)') 'Usage:  getevec [-i|-v] [evec file]'
      write (*, '(a)') '   -i:  locate based on index (default)'
      write (*, '(a)') '   -v:  locate based on eigenvalue'
      call exit(0)
    else
      write (*, "('Argument ',i2,' ignored.')") ia
    end if
  end do
  if (len_trim(fname) == 0) then
    write (*, "('Enter the Eigensystem filename ==> ',$)")
    read (*, '(a)') fname
  end if

  open (newunit=u, file=trim(fname), form='unformatted', status='old', iostat=ios)
  if (ios /= 0) then
    write (*, "('Error opening file...')")
    call exit(1)
  end if
  read (u) ind, nyy, ndof, itp, iev, icurve, itop, iwall, iwallt
  allocate (y(nyy), eta(nyy), deta(nyy), d2eta(nyy))
  nmax = ndof * nyy
  if (itp /= 1) nmax = 2 * nmax
  allocate (ev(nmax), vec(nmax, nmax))
  vec = (0.0d0, 0.0d0)
  read (u) om, al, be, Re, Ma, Pr
  read (u) xx, y, eta, deta, d2eta, yii, ymx
  read (u) ev
  if (iev == 1 .or. itp == 1) read (u) vec
  close (u)
  k = sqrt(al**2 + be**2)

  call table()
  iver = 0
  do
    if (by_index) then
      write (*, "('Which eigenfunction ==> ',$)")
      read (*, *, iostat=ios) j
      if (ios /= 0 .or. j == 0) exit
      if (j == -1) then
        call table()
        cycle
      end if
      if (j < -1 .or. j > nmax) cycle
    else
      write (*, "('Approximate eigenvalue: real, imag  ==> ',$)")
      read (*, *, iostat=ios) vr, vi
      if (ios /= 0) exit
      val = cmplx(vr, vi, kind=8)
      if (val == (0.0d0, 0.0d0)) exit
      j = 1
      d = abs(val - ev(1))
      do i = 2, nmax
        if (abs(val - ev(i)) < d) then
          d = abs(val - ev(i))
          j = i
        end if
      end do
    end if
    ! first entry of largest modulus over the flow components
    sc = (0.0d0, 0.0d0)
    do i = 1, nyy * ndof
      if (abs(vec(i, j)) > abs(sc)) sc = vec(i, j)
    end do
    if (sc /= (0.0d0, 0.0d0)) vec(1:nyy * ndof, j) = vec(1:nyy * ndof, j) / sc
    iver = iver + 1
    base = 'time'
    if (itp /= 1) base = 'space'
    write (outname, '(a,".",i0)') trim(base), iver
    open (newunit=u, file=trim(outname), form='formatted', status='unknown')
    write (u, "('# Re = ',1pe13.6,', Ma = ',1pe13.6,', Pr = ',1pe13.6)") Re, Ma, Pr
    if (itp == 1) then
      write (u, "('# Omega = (',1pe21.13E3,',',1pe21.13E3,')')") ev(j)
      write (u, "('# Alpha = (',1pe21.13E3,',',1pe21.13E3,')')") al
    else
      write (u, "('# Omega = (',1pe21.13E3,',',1pe21.13E3,')')") om
      write (u, "('# Alpha = (',1pe21.13E3,',',1pe21.13E3,')')") ev(j)
    end if
    write (u, "('# Beta  = (',1pe21.13E3,',',1pe21.13E3,')')") be
    do i = nyy, 1, -1
      i0 = (i - 1) * ndof
      write (u, '(1p,11(1pe21.13E3,1x))') y(i), (real(vec(i0 + ia, j)), aimag(vec(i0 + ia, j)), ia = 1, ndof)
    end do
    close (u)
    write (*, "('  Eigenfunction saved in: ',a)") trim(outname)
  end do

contains

  subroutine table()
    integer :: jj
    if (itp == 1) then
      write (*, "(/,'Temporal Eigensystem:',/)")
      write (*, "('s = ',1pe13.6,' Yi = ',1pe20.13,', Ymax = ',1pe13.6)") xx, yii, ymx
      write (*, "('Alpha = (',1pe13.6,',',1pe13.6,')')") al
    else
      write (*, "(/,'Spatial Eigensystem:',/)")
      write (*, "('s = ',1pe13.6,' Yi = ',1pe20.13,', Ymax = ',1pe13.6)") xx, yii, ymx
      write (*, "('Omega = (',1pe13.6,',',1pe13.6,')')") om
    end if
    write (*, "('Beta  = (',1pe13.6,',',1pe13.6,')')") be
    write (*, "('Re = ',1pe13.6,', Ma = ',1pe13.6,', Pr = ',1pe13.6)") Re, Ma, Pr
    write (*, "(100('='))")
    do jj = 1, nmax
      if (abs(real(ev(jj))) >= 1.0d299) cycle          ! slots this build leaves empty
      if (itp == 1) then
        write (*, "(i5,4x,4(1pe21.13E3,2x))") jj, real(ev(jj)), aimag(ev(jj)), real(ev(jj) / k), aimag(ev(jj) / k)
      else if (ev(jj) /= (0.0d0, 0.0d0)) then
        write (*, "(i5,4x,4(1pe21.13E3,2x))") jj, real(ev(jj)), aimag(ev(jj)), real(om / ev(jj)), &
          aimag(om / ev(jj))
      else
        write (*, "(i5,4x,4(1pe21.13E3,2x))") jj, real(ev(jj)), aimag(ev(jj)), 0.0d0, 0.0d0
      end if
    end do
    write (*, "(100('='),/)")
  end subroutine table

end program getevec
