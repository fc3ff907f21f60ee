! Drives the qgn_dfti module the way a DFTI host does: two descriptors
! configured call by call (create, number of transforms, distances,
! conjugate-even storage, packed format, placement, strides, commit; the
! sequence of MKL_DFTI_Setup.f90:11-50), one forward and one backward
! transform of host arrays, then a descriptor the library must refuse
! (in place) and a forward compute on a backward-configured descriptor.
! Usage: dfti_check nx ny nz dir  (dir holds grid.bin and spec.bin; the
! program writes fwd.bin and bwd.bin there and prints one line per check).
program dfti_check
  use qgn_dfti
  implicit none
  integer, parameter :: dp = kind(1.d0)
  integer :: nx, ny, nz, Status, Dims(2), real_strides(3), cplx_strides(3), u
  character(len=512) :: arg, dir
  real(dp), allocatable :: grid(:), gback(:)
  complex(dp), allocatable :: spec(:), specin(:)
  type(DFTI_DESCRIPTOR), pointer :: g2s, s2g, bad

  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) nz
  call get_command_argument(4, dir)
  allocate(grid(nx*ny*nz), gback(nx*ny*nz), spec((nx/2+1)*ny*nz), specin((nx/2+1)*ny*nz))
  open(newunit=u, file=trim(dir)//'/grid.bin', access='stream', status='old'); read(u) grid; close(u)
  open(newunit=u, file=trim(dir)//'/spec.bin', access='stream', status='old'); read(u) specin; close(u)

  Dims = [nx, ny]
  real_strides = [0, 1, nx]
  cplx_strides = [0, 1, nx/2 + 1]
  Status = DftiCreateDescriptor(g2s, DFTI_DOUBLE, DFTI_REAL, 2, Dims);                call chk(Status)
  Status = DftiSetValue(g2s, DFTI_NUMBER_OF_TRANSFORMS, nz);                         call chk(Status)
  Status = DftiSetValue(g2s, DFTI_INPUT_DISTANCE, nx*ny);                            call chk(Status)
  Status = DftiSetValue(g2s, DFTI_OUTPUT_DISTANCE, (nx/2+1)*ny);                     call chk(Status)
  Status = DftiSetValue(g2s, DFTI_CONJUGATE_EVEN_STORAGE, DFTI_COMPLEX_COMPLEX);     call chk(Status)
  Status = DftiSetValue(g2s, DFTI_PACKED_FORMAT, DFTI_CCE_FORMAT);                   call chk(Status)
  Status = DftiSetValue(g2s, DFTI_PLACEMENT, DFTI_NOT_INPLACE);                      call chk(Status)
  Status = DftiSetValue(g2s, DFTI_INPUT_STRIDES, real_strides);                      call chk(Status)
  Status = DftiSetValue(g2s, DFTI_OUTPUT_STRIDES, cplx_strides);                     call chk(Status)
  Status = DftiCommitDescriptor(g2s);                                                call chk(Status)
  Status = DftiCreateDescriptor(s2g, DFTI_DOUBLE, DFTI_REAL, 2, Dims);                call chk(Status)
  Status = DftiSetValue(s2g, DFTI_NUMBER_OF_TRANSFORMS, nz);                         call chk(Status)
  Status = DftiSetValue(s2g, DFTI_INPUT_DISTANCE, (nx/2+1)*ny);                      call chk(Status)
  Status = DftiSetValue(s2g, DFTI_OUTPUT_DISTANCE, nx*ny);                           call chk(Status)
  Status = DftiSetValue(s2g, DFTI_CONJUGATE_EVEN_STORAGE, DFTI_COMPLEX_COMPLEX);     call chk(Status)
  Status = DftiSetValue(s2g, DFTI_PLACEMENT, DFTI_NOT_INPLACE);                      call chk(Status)
  Status = DftiSetValue(s2g, DFTI_PACKED_FORMAT, DFTI_CCE_FORMAT);                   call chk(Status)
  Status = DftiSetValue(s2g, DFTI_INPUT_STRIDES, cplx_strides);                      call chk(Status)
  Status = DftiSetValue(s2g, DFTI_OUTPUT_STRIDES, real_strides);                     call chk(Status)
  Status = DftiCommitDescriptor(s2g);                                                call chk(Status)

  Status = DftiComputeForward(g2s, grid, spec);                                      call chk(Status)
  Status = DftiComputeBackward(s2g, specin, gback);                                  call chk(Status)
  open(newunit=u, file=trim(dir)//'/fwd.bin', access='stream', status='replace'); write(u) spec; close(u)
  open(newunit=u, file=trim(dir)//'/bwd.bin', access='stream', status='replace'); write(u) gback; close(u)
  print '(a)', 'transforms done'

  ! in place is not provided: commit refuses it with a message
  Status = DftiCreateDescriptor(bad, DFTI_DOUBLE, DFTI_REAL, 2, Dims);                call chk(Status)
  Status = DftiSetValue(bad, DFTI_PLACEMENT, DFTI_INPLACE);                          call chk(Status)
  Status = DftiCommitDescriptor(bad)
  if (Status /= 0) print '(a,a)', 'refused in place: ', trim(DftiErrorMessage(Status))
  Status = DftiFreeDescriptor(bad)
  ! a backward descriptor does not compute forward
  Status = DftiComputeForward(s2g, grid, spec)
  if (Status /= 0) print '(a,a)', 'refused direction: ', trim(DftiErrorMessage(Status))

  Status = DftiFreeDescriptor(g2s);                                                  call chk(Status)
  Status = DftiFreeDescriptor(s2g);                                                  call chk(Status)

contains
  subroutine chk(st)
    integer, intent(in) :: st
    if (st /= 0) then
      print '(a,i0,a,a)', 'DFTI status ', st, ': ', trim(DftiErrorMessage(st))
      error stop 1
    end if
  end subroutine chk
end program dfti_check
