!==============================================================================
! alpha_march: a caller written against the reference's interface only --
! module `stuff`, input, temporal(name, ind) and makename, as the reference's
! own mtemporal (mtemporal.f90:25-38) uses them -- compiled unchanged against
! this build's Fortran layer.  It marches alpha over a range at fixed beta and
! writes one eigensystem file eig.<k> per step.
!
! stdin: the reference's input sequence with itype 1 (its alpha/beta lines give
! the start), then the index of the profile, then "alpha_end nsteps".
!==============================================================================
program alpha_march
  use stuff
  implicit none
  integer :: ind, nsteps, k
  real(8) :: a0, a1
  character(len=80) :: name

  call input
  read (*, *) ind
  read (*, *) a1, nsteps
  a0 = real(alpha)
  do k = 1, nsteps
    alpha = a0 + (a1 - a0) * dble(k - 1) / dble(max(nsteps - 1, 1))
    call makename('eig', k, name)
    write (*, '(a,i3,a,1pe13.6)') ' step ', k, ' alpha ', real(alpha)
    call temporal(name, ind)
  end do
end program alpha_march
