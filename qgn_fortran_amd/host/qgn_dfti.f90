! DFTI descriptors over the HIP library (include/qgn.h, qgn_dfti_*): the
! calls that MKL_DFTI_Setup.f90:11-50 and QGN_Module.f90 make (create, set
! values, commit, compute forward / backward, free), with the same names,
! argument order and integer status results, so a host written against
! MKL's Fortran DFTI interface needs only its `use` line changed.  The
! transforms are those the reference configures: batched 2-D real
! transforms, CCE half spectrum along x, out of place, unnormalised, on
! real(dp) and complex(dp) host arrays.  A nonzero status comes with a
! message from DftiErrorMessage.
module qgn_dfti
  use iso_c_binding
  implicit none
  private
  public :: DFTI_DESCRIPTOR, DftiCreateDescriptor, DftiSetValue, DftiCommitDescriptor, DftiComputeForward, &
            DftiComputeBackward, DftiFreeDescriptor, DftiErrorMessage
  public :: DFTI_DOUBLE, DFTI_REAL, DFTI_COMPLEX_COMPLEX, DFTI_CCE_FORMAT, DFTI_NOT_INPLACE, DFTI_INPLACE
  public :: DFTI_NUMBER_OF_TRANSFORMS, DFTI_INPUT_DISTANCE, DFTI_OUTPUT_DISTANCE, DFTI_CONJUGATE_EVEN_STORAGE, &
            DFTI_PACKED_FORMAT, DFTI_PLACEMENT, DFTI_INPUT_STRIDES, DFTI_OUTPUT_STRIDES, DFTI_FORWARD_SCALE, &
            DFTI_BACKWARD_SCALE

  ! configuration values and parameters (the enums of include/qgn.h)
  integer, parameter :: DFTI_DOUBLE = 1, DFTI_REAL = 2, DFTI_COMPLEX_COMPLEX = 3, DFTI_CCE_FORMAT = 4, &
                        DFTI_NOT_INPLACE = 5, DFTI_INPLACE = 6
  integer, parameter :: DFTI_NUMBER_OF_TRANSFORMS = 1, DFTI_INPUT_DISTANCE = 2, DFTI_OUTPUT_DISTANCE = 3, &
                        DFTI_CONJUGATE_EVEN_STORAGE = 4, DFTI_PACKED_FORMAT = 5, DFTI_PLACEMENT = 6, &
                        DFTI_INPUT_STRIDES = 7, DFTI_OUTPUT_STRIDES = 8, DFTI_FORWARD_SCALE = 9, &
                        DFTI_BACKWARD_SCALE = 10

  type :: DFTI_DESCRIPTOR
    type(c_ptr) :: h = c_null_ptr
  end type DFTI_DESCRIPTOR

  interface DftiSetValue
    module procedure set_int, set_ints, set_real
  end interface DftiSetValue

  interface
    integer(c_int) function c_create(h, prec, dom, ndim, lens) bind(C, name="qgn_dfti_create_descriptor")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: prec, dom, ndim
      integer(c_long_long), intent(in) :: lens(*)
    end function
    integer(c_int) function c_set_int(h, param, v) bind(C, name="qgn_dfti_set_value_int")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: h
      integer(c_int), value :: param
      integer(c_long_long), value :: v
    end function
    integer(c_int) function c_set_ints(h, param, v, n) bind(C, name="qgn_dfti_set_value_ints")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: h
      integer(c_int), value :: param, n
      integer(c_long_long), intent(in) :: v(*)
    end function
    integer(c_int) function c_set_real(h, param, v) bind(C, name="qgn_dfti_set_value_double")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: param
      real(c_double), value :: v
    end function
    integer(c_int) function c_commit(h) bind(C, name="qgn_dfti_commit_descriptor")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function c_forward(h, grid, spec) bind(C, name="qgn_dfti_compute_forward")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      real(c_double), intent(in) :: grid(*)
      complex(c_double_complex), intent(out) :: spec(*)
    end function
    integer(c_int) function c_backward(h, spec, grid) bind(C, name="qgn_dfti_compute_backward")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: spec(*)
      real(c_double), intent(out) :: grid(*)
    end function
    integer(c_int) function c_free(h) bind(C, name="qgn_dfti_free_descriptor")
      import :: c_int, c_ptr
      type(c_ptr), intent(inout) :: h
    end function
    type(c_ptr) function c_last_error() bind(C, name="qgn_last_error")
      import :: c_ptr
    end function
  end interface

contains

  integer function DftiCreateDescriptor(desc, precision, domain, dimension, lengths)
    type(DFTI_DESCRIPTOR), pointer :: desc
    integer, intent(in) :: precision, domain, dimension
    integer, intent(in) :: lengths(*)
    integer(c_long_long) :: lens(2)
    allocate(desc)
    lens = 0
    lens(1:min(dimension, 2)) = lengths(1:min(dimension, 2))
    DftiCreateDescriptor = c_create(desc%h, precision, domain, dimension, lens)
  end function DftiCreateDescriptor

  integer function set_int(desc, param, value)
    type(DFTI_DESCRIPTOR), pointer :: desc
    integer, intent(in) :: param, value
    set_int = c_set_int(desc%h, param, int(value, c_long_long))
  end function set_int

  integer function set_ints(desc, param, values)
    type(DFTI_DESCRIPTOR), pointer :: desc
    integer, intent(in) :: param
    integer, intent(in) :: values(:)
    integer(c_long_long) :: v(size(values))
    v = values
    set_ints = c_set_ints(desc%h, param, v, size(values))
  end function set_ints

  integer function set_real(desc, param, value)
    type(DFTI_DESCRIPTOR), pointer :: desc
    integer, intent(in) :: param
    real(c_double), intent(in) :: value
    set_real = c_set_real(desc%h, param, value)
  end function set_real

  integer function DftiCommitDescriptor(desc)
    type(DFTI_DESCRIPTOR), pointer :: desc
    DftiCommitDescriptor = c_commit(desc%h)
  end function DftiCommitDescriptor

  ! DftiComputeForward(g2s_q, gridE, specE): real(dp) in, complex(dp) out
  integer function DftiComputeForward(desc, grid, spec)
    type(DFTI_DESCRIPTOR), pointer :: desc
    real(c_double), intent(in) :: grid(*)
    complex(c_double_complex), intent(out) :: spec(*)
    DftiComputeForward = c_forward(desc%h, grid, spec)
  end function DftiComputeForward

  ! DftiComputeBackward(s2g_q, specE, gridE): complex(dp) in, real(dp) out
  integer function DftiComputeBackward(desc, spec, grid)
    type(DFTI_DESCRIPTOR), pointer :: desc
    complex(c_double_complex), intent(in) :: spec(*)
    real(c_double), intent(out) :: grid(*)
    DftiComputeBackward = c_backward(desc%h, spec, grid)
  end function DftiComputeBackward

  integer function DftiFreeDescriptor(desc)
    type(DFTI_DESCRIPTOR), pointer :: desc
    DftiFreeDescriptor = c_free(desc%h)
    deallocate(desc)
    nullify(desc)
  end function DftiFreeDescriptor

  ! the library's message for the last nonzero status of this thread
  function DftiErrorMessage(status) result(msg)
    integer, intent(in) :: status
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    if (status == 0) return
    call c_f_pointer(c_last_error(), s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function DftiErrorMessage

end module qgn_dfti
