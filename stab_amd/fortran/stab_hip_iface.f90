!==============================================================================
! ISO_C_BINDING interface of libstab_hip (include/stab_hip.h).
!
! Fortran callers of the reference's temporal/spatial path (temporal.f90:2,
! spatial.f90:2) reach the GPU through these bind(C) declarations.  The
! derived type stab_point mirrors the C struct field by field; complex values
! are complex(c_double_complex), which has the layout of {double re, im}.
!==============================================================================
module stab_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: STAB_OK = 0
  integer(c_int), parameter :: STAB_MAX_NEV = 16
  integer(c_int), parameter :: STAB_MODE_PLAIN_F64 = 0, STAB_MODE_MIXED = 1
  integer(c_int), parameter :: STAB_INFO_OK = 0, STAB_INFO_SINGULAR = 1, STAB_INFO_NOCONV = 2
  integer(c_int), parameter :: STAB_INFO_FALLBACK = 3
  integer(c_int), parameter :: STAB_LEAD_KRYLOV = 0, STAB_LEAD_INVIT = 1, STAB_TRACK = 2
  integer(c_int), parameter :: STAB_MAT_A0 = 0, STAB_MAT_B0 = 1, STAB_MAT_C0 = 2, STAB_MAT_C1 = 3, &
                               STAB_MAT_C2 = 4, STAB_MAT_SHIFTED = 5

  type, bind(C) :: stab_point
    integer(c_int32_t) :: itype, ny, mattyp, wallt, top, profile, nev, ider
    real(c_double) :: Ma, Re, Pr, Te, yi, ymax
    complex(c_double_complex) :: alpha, beta, omega, shift
  end type stab_point

  interface
    integer(c_int) function stab_ctx_create(ctx, device) bind(C, name="stab_ctx_create")
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
    end function

    integer(c_int) function stab_ctx_destroy(ctx) bind(C, name="stab_ctx_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function stab_last_error(ctx) bind(C, name="stab_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function stab_profile_set(ctx, slot, nym, y, vt) bind(C, name="stab_profile_set")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, nym
      real(c_double) :: y(*), vt(*)
    end function

    integer(c_int) function stab_grid(ctx, pt, y, eta, deta, d2eta, vm, d1, d2) bind(C, name="stab_grid")
      import :: c_ptr, c_int, stab_point
      type(c_ptr), value :: ctx
      type(stab_point) :: pt
      type(c_ptr), value :: y, eta, deta, d2eta, vm, d1, d2
    end function

    integer(c_int) function stab_solve(ctx, npts, pts, mode, eval, evec, resid, info, stats) &
        bind(C, name="stab_solve")
      import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex, stab_point
      type(c_ptr), value :: ctx
      integer(c_int), value :: npts, mode
      type(stab_point) :: pts(*)
      complex(c_double_complex) :: eval(*), evec(*)
      real(c_double) :: resid(*)
      integer(c_int32_t) :: info(*)
      type(c_ptr), value :: stats
    end function

    integer(c_int) function stab_pencil(ctx, pt, which, out) bind(C, name="stab_pencil")
      import :: c_ptr, c_int, c_double_complex, stab_point
      type(c_ptr), value :: ctx
      type(stab_point) :: pt
      integer(c_int), value :: which
      complex(c_double_complex) :: out(*)
    end function

    integer(c_int) function stab_shifted_solve(ctx, pt, mode, ir_steps, b, x, info) &
        bind(C, name="stab_shifted_solve")
      import :: c_ptr, c_int, c_int32_t, c_double_complex, stab_point
      type(c_ptr), value :: ctx
      type(stab_point) :: pt
      integer(c_int), value :: mode, ir_steps
      complex(c_double_complex) :: b(*), x(*)
      integer(c_int32_t) :: info
    end function

    integer(c_int) function stab_lu_factors(ctx, pt, mode, lu, perm, scales, info) &
        bind(C, name="stab_lu_factors")
      import :: c_ptr, c_int, c_int32_t, c_double_complex, stab_point
      type(c_ptr), value :: ctx
      type(stab_point) :: pt
      integer(c_int), value :: mode
      complex(c_double_complex) :: lu(*)
      integer(c_int32_t) :: perm(*)
      type(c_ptr), value :: scales
      integer(c_int32_t) :: info
    end function

    integer(c_size_t) function stab_evec_bytes(npts, pts) bind(C, name="stab_evec_bytes")
      import :: c_int, c_size_t, stab_point
      integer(c_int), value :: npts
      type(stab_point) :: pts(*)
    end function

    integer(c_int) function stab_solve_leading(ctx, npts, pts, mode, method, maxit, fallback_shift, eval, evec, &
                                               resid, farfield, iters, info, stats) &
        bind(C, name="stab_solve_leading")
      import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex, stab_point
      type(c_ptr), value :: ctx
      integer(c_int), value :: npts, mode, method, maxit
      type(stab_point) :: pts(*)
      type(c_ptr), value :: fallback_shift
      complex(c_double_complex) :: eval(*)
      type(c_ptr), value :: evec
      real(c_double) :: resid(*), farfield(*)
      integer(c_int32_t) :: iters(*), info(*)
      type(c_ptr), value :: stats
    end function

    integer(c_int) function stab_mlp_create(ctx, nin, width, depth, nout, seed, mlp) bind(C, name="stab_mlp_create")
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nin, width, depth, nout
      integer(c_int64_t), value :: seed
      type(c_ptr) :: mlp
    end function

    integer(c_int) function stab_mlp_destroy(mlp) bind(C, name="stab_mlp_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: mlp
    end function

    integer(c_int) function stab_mlp_nparams(mlp) bind(C, name="stab_mlp_nparams")
      import :: c_ptr, c_int
      type(c_ptr), value :: mlp
    end function

    integer(c_int) function stab_mlp_set_params(mlp, params) bind(C, name="stab_mlp_set_params")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: mlp
      real(c_float) :: params(*)
    end function

    integer(c_int) function stab_mlp_set_scaling(mlp, in_shift, in_scale, out_shift, out_scale) &
        bind(C, name="stab_mlp_set_scaling")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: mlp
      real(c_float) :: in_shift(*), in_scale(*), out_shift(*), out_scale(*)
    end function

    integer(c_int) function stab_mlp_predict(mlp, n, x, y) bind(C, name="stab_mlp_predict")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: mlp
      integer(c_int), value :: n
      real(c_float) :: x(*), y(*)
    end function
  end interface

contains

  !> Text of stab_last_error as a Fortran string.
  function stab_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(stab_last_error(ctx), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function stab_error_text

end module stab_hip
