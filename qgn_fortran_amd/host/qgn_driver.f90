!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!
! Driver of the QGN host (behaviour of QGN_Driver.f90): the last command-line
! argument is the restart index N0; N0 > 1 reads q.N0.dat (physical PV,
! raw float64 (nx,ny,nz)); otherwise the reference's analytic initial
! condition is used.  Cadence and step count come from the &driver group of
! qgn.nml (defaults as in QGN_Driver.f90:6-7).
!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!
program main
use QGN_Module
implicit none
real(dp), allocatable :: q0(:,:,:)
complex(dp), allocatable :: q_hat(:,:,:)
real(dp) :: t = 0._dp, dt = 750._dp
integer :: i, j, N0 = 1, Nt = 10, diag_freq = 10000, chkpt_freq = 1, out_freq = 1, u
character(len=32) :: arg
character(len=16) :: file_name
logical :: there
namelist /driver/ dt, Nt, diag_freq, chkpt_freq, out_freq

call ReadParameters()
inquire(file='qgn.nml', exist=there)
if (there) then
    open(newunit=u, file='qgn.nml', status='OLD', action='READ')
    read(u, nml=driver)
    close(u)
end if
do i = 1, command_argument_count()
    call get_command_argument(i, arg)
    if (len_trim(arg) == 0) exit
    read (arg,*) N0
end do
print *,'---------------------------------------------------------------'
print *,' Starting iteration  '
print *, N0
print *,'---------------------------------------------------------------'
allocate(q0(nx,ny,nz), q_hat(nx/2+1,ny,nz))
if (N0 > 1) then
    write (file_name,'(A2,I0.9,A4)') "q.",N0,'.dat'
    print *,'---------------------------------------------------------------'
    print *,' Opening q0 file  '
    print *,  file_name
    print *,'---------------------------------------------------------------'
    open(newunit=u, file=file_name, access='STREAM', status='OLD')
    read(u) q0
    close(u)
else
    do i = 1, nx
        do j = 1, ny
            q0(i,j,:) = 1.E-5_dp*sin(48._dp*pi*real(i-1,dp)/real(nx,dp))
            q0(i,j,:) = q0(i,j,:) + 2.E-5_dp*cos(32._dp*pi*real(j-1,dp)/real(ny,dp))
        end do
    end do
end if

call Initialize(q0, q_hat)
call WriteQP(q_hat, 1, 0)
do i = N0, N0 + Nt
    call TimeStep(q_hat, t, dt)
    if (mod(i, out_freq) == 0) print *, 'Time since inception = ', real(t/86400._dp), &
        'days; time step size = ', real(dt), 's'
    if (mod(i, diag_freq) == 0) then
        print *, ' Diagnostics, RecNo = ', i/diag_freq
        call WriteSpectra(i)
    end if
    if (mod(i, chkpt_freq) == 0) then
        if (.not. sync_host) call GetState(q_hat)
        call WriteQP(q_hat, 1, i)
        call WriteQPmodal(q_hat, 1, i)
    end if
end do
call Cleanup
end program main
