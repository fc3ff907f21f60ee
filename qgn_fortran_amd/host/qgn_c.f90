! iso_c_binding interface to the HIP library declared in include/qgn.h.
! Arrays are passed in the reference's own layouts: real(dp)(nx,ny,nz) and
! complex(dp)(nx/2+1,ny,nz) (QGN_Module.f90:30-41).
module qgn_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: QGN_AB3 = 0, QGN_RK4 = 1, QGN_BSRK32 = 2, QGN_ARK43 = 3, &
                               QGN_ARK43_QGLEITH = 4
  integer(c_int), parameter :: QGN_FP64 = 0, QGN_MIXED = 1

  ! mirrors struct qgn_params
  type, bind(C) :: qgn_params
    integer(c_int) :: nx, ny, nz
    real(c_double) :: Lx, Ly, Htot, beta, A2, A8, r0, C_d, TOL, dt_max
    integer(c_int) :: is_explicit, scheme, precision, max_rejects
  end type qgn_params

  type, bind(C) :: qgn_stats
    integer(c_long_long) :: steps, attempts, rejects, rhs_evals
    real(c_double) :: last_err, err0
  end type qgn_stats

  interface
    integer(c_int) function qgn_create(ctx, p, H, Modes, EVals, uBar, qyBar) bind(C, name="qgn_create")
      import :: c_int, c_ptr, c_double, qgn_params
      type(c_ptr), intent(out) :: ctx
      type(qgn_params), intent(in) :: p
      real(c_double), intent(in) :: H(*), Modes(*), EVals(*), uBar(*), qyBar(*)
    end function

    integer(c_int) function qgn_destroy(ctx) bind(C, name="qgn_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function qgn_initialize(ctx, qp, q_hat) bind(C, name="qgn_initialize")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: qp(*)
      complex(c_double_complex), intent(out) :: q_hat(*)
    end function

    integer(c_int) function qgn_set_state(ctx, q_hat) bind(C, name="qgn_set_state")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: q_hat(*)
    end function

    integer(c_int) function qgn_get_state(ctx, q_hat) bind(C, name="qgn_get_state")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: q_hat(*)
    end function

    integer(c_int) function qgn_time_step(ctx, t, dt) bind(C, name="qgn_time_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: t, dt
    end function

    integer(c_int) function qgn_rhs(ctx, q_hat, rhs) bind(C, name="qgn_rhs")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: q_hat(*)
      complex(c_double_complex), intent(out) :: rhs(*)
    end function

    integer(c_int) function qgn_fields(ctx, q_phys, psi_phys) bind(C, name="qgn_fields")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: q_phys(*), psi_phys(*)
    end function

    integer(c_int) function qgn_modal_fields(ctx, q_mode, psi_mode, j_mode) bind(C, name="qgn_modal_fields")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: q_mode(*), psi_mode(*), j_mode(*)
    end function

    integer(c_int) function qgn_spectra(ctx, nbins, ke, ens) bind(C, name="qgn_spectra")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbins
      real(c_double), intent(out) :: ke(*), ens(*)
    end function

    integer(c_int) function qgn_get_stats(ctx, s) bind(C, name="qgn_get_stats")
      import :: c_int, c_ptr, qgn_stats
      type(c_ptr), value :: ctx
      type(qgn_stats), intent(out) :: s
    end function

    integer(c_int) function qgn_vertical_modes(nz, Htot, H_in, S, uBar, H_out, L, Modes, EVals, qyBar) &
        bind(C, name="qgn_vertical_modes")
      import :: c_int, c_double
      integer(c_int), value :: nz
      real(c_double), value :: Htot
      real(c_double), intent(in) :: H_in(*), S(*), uBar(*)
      real(c_double), intent(out) :: H_out(*), L(*), Modes(*), EVals(*), qyBar(*)
    end function

    type(c_ptr) function qgn_last_error() bind(C, name="qgn_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! error stop with the library's message on a non-zero status
  subroutine qgn_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: i
    if (status == 0) return
    call c_f_pointer(qgn_last_error(), msg, [256])
    text = ' '
    do i = 1, 256
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    print *, 'qgn error in ', where, ': status ', status, ' ', trim(text)
    error stop 1
  end subroutine qgn_check

end module qgn_c
