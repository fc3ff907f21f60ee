!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!
! QGN_Module over the MI355X library: the public procedures of the
! reference module (QGN_Module.f90:7) with the arithmetic done by the HIP
! library through qgn_c.  Set-up reads the same binary profile files and
! writes the same diagnostics files; TimeStep's scheme is chosen by the
! include below, as in the reference (QGN_Module.f90:306).
!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!
module QGN_Module
use iso_c_binding
use qgn_c
implicit none
private
public dp, nx, ny, nz, pi, Initialize, TimeStep, WriteQP, WriteQPmodal, Cleanup, GetRHS, ReadParameters, &
       GetState, WriteSpectra, sync_host

include "parameters.f90"
include "ARK43_QGLeith.f90"

type(c_ptr) :: ctx = c_null_ptr
real(dp), allocatable :: H(:), S(:), uBar(:), qyBar(:), EVals(:)
real(dp), allocatable :: L(:,:), Modes(:,:)

contains

! Optional namelist override of the parameters (qgn.nml, group &qgn).
subroutine ReadParameters()
    logical :: there
    integer :: u
    inquire(file='qgn.nml', exist=there)
    if (there) then
        open(newunit=u, file='qgn.nml', status='OLD', action='READ')
        read(u, nml=qgn)
        close(u)
    end if
    if (dt_max <= 0._dp) dt_max = pi / (beta * Lx)
end subroutine ReadParameters

! Initialize (QGN_Module.f90:53-188): profiles, vertical modes, plans,
! transform of the initial condition.
subroutine Initialize(qp, q_hat)
    real(dp), intent(in)     :: qp(nx,ny,nz)
    complex(dp), intent(out) :: q_hat(nx/2+1,ny,nz)
    real(dp) :: Hraw(nz)
    type(qgn_params) :: p
    integer :: u, k

    print *,'---------------------------------------------------------------'
    print *,' dx = ', Lx/real(nx,dp),' dy = ', Ly/real(ny,dp)
    allocate(H(nz), S(0:nz), uBar(nz), qyBar(nz), EVals(nz), L(nz,nz), Modes(nz,nz))
    open(newunit=u, file='H.dat', access='STREAM', status='OLD')
    read(u) Hraw
    close(u)
    open(newunit=u, file='S.dat', access='STREAM', status='OLD')
    read(u) S
    close(u)
    open(newunit=u, file='UBar.dat', access='STREAM', status='OLD')
    read(u) uBar
    close(u)
    call qgn_check(qgn_vertical_modes(int(nz, c_int), Htot, Hraw, S, uBar, H, L, Modes, EVals, qyBar), &
                   'qgn_vertical_modes')
    print *,'---------------------------------------------------------------'
    print *,' Layer depths  '
    print *, H(:)
    print *,'---------------------------------------------------------------'
    print *,' f^2/N^2  '
    print *, S(:)
    print *,'---------------------------------------------------------------'
    open(newunit=u, file='L.dat', access='STREAM', status='REPLACE')
    write(u) L
    close(u)
    open(newunit=u, file='modes.dat', access='STREAM', status='REPLACE')
    write(u) Modes
    close(u)
    print *,'---------------------------------------------------------------'
    print *,' The eigenvalues are '
    print *, EVals(:)
    print *,'---------------------------------------------------------------'
    print *,' The deformation radii are '
    print *, (1._dp/sqrt(-EVals(k)), k = 1, nz)
    print *,'---------------------------------------------------------------'
    print *,' ubar  '
    print *, uBar(:)
    print *,'---------------------------------------------------------------'
    print *,' qybar  '
    print *, qyBar(:)
    print *,'---------------------------------------------------------------'
    open(newunit=u, file='Qy.dat', access='STREAM', status='REPLACE')
    write(u) qyBar
    close(u)

    p%nx = nx; p%ny = ny; p%nz = nz
    p%Lx = Lx; p%Ly = Ly; p%Htot = Htot; p%beta = beta
    p%A2 = A2; p%A8 = A8; p%r0 = r0; p%C_d = C_d; p%TOL = TOL; p%dt_max = dt_max
    p%is_explicit = merge(1, 0, is_explicit)
    p%scheme = time_scheme
    p%precision = precision
    p%max_rejects = max_rejects
    call qgn_check(qgn_create(ctx, p, H, Modes, EVals, uBar, qyBar), 'qgn_create')
    print *,' Transforming the initial condition '
    call qgn_check(qgn_initialize(ctx, qp, q_hat), 'qgn_initialize')
    print *,' Initialization complete '
    print *,'---------------------------------------------------------------'
end subroutine Initialize

! TimeStep (QGN_Module.f90:302-309).  The state lives on the GPU; q_hat is
! refreshed from it when sync_host is set.
subroutine TimeStep(q_hat, t, dt)
    complex(dp), intent(inout) :: q_hat(nx/2+1,ny,nz)
    real(dp),    intent(inout) :: t, dt
    type(qgn_stats) :: s0, s1
    integer(c_long_long) :: k
    call qgn_check(qgn_get_stats(ctx, s0), 'qgn_get_stats')
    call qgn_check(qgn_time_step(ctx, t, dt), 'qgn_time_step')
    if (sync_host) call qgn_check(qgn_get_state(ctx, q_hat), 'qgn_get_state')
    ! ARK43_QGLeith reports each rejected attempt (ARK43_QGLeith.f90:126)
    if (time_scheme == QGN_ARK43_QGLEITH) then
        call qgn_check(qgn_get_stats(ctx, s1), 'qgn_get_stats')
        do k = 1, s1%rejects - s0%rejects
            print *, 'Rejected step'
        end do
    end if
end subroutine TimeStep

subroutine GetState(q_hat)
    complex(dp), intent(out) :: q_hat(nx/2+1,ny,nz)
    call qgn_check(qgn_get_state(ctx, q_hat), 'qgn_get_state')
end subroutine GetState

! GetRHS (QGN_Module.f90:258-296)
subroutine GetRHS(q_hat, RHS)
    complex(dp), intent(in)  :: q_hat(nx/2+1,ny,nz)
    complex(dp), intent(out) :: RHS(nx/2+1,ny,nz)
    call qgn_check(qgn_rhs(ctx, q_hat, RHS), 'qgn_rhs')
end subroutine GetRHS

! WriteQP (QGN_Module.f90:348-379): q and psi of q_hat as q.%09d.dat and
! p.%09d.dat, raw float64 in (nx,ny,nz) order.
subroutine WriteQP(q_hat, RecNo, nt)
    complex(dp), intent(in) :: q_hat(nx/2+1,ny,nz)
    integer,     intent(in) :: RecNo, nt
    real(dp), allocatable :: q_phys(:,:,:), psi_phys(:,:,:)
    character(len=50) :: file_name
    integer :: u
    allocate(q_phys(nx,ny,nz), psi_phys(nx,ny,nz))
    call qgn_check(qgn_set_state(ctx, q_hat), 'qgn_set_state')
    call qgn_check(qgn_fields(ctx, q_phys, psi_phys), 'qgn_fields')
    write (file_name,'(A2,I0.9,A4)') "q.",nt,'.dat'
    open(newunit=u, file=file_name, access='STREAM', status='UNKNOWN')
    if (RecNo == 1) rewind u
    write(u) q_phys
    close(u)
    write (file_name,'(A2,I0.9,A4)') "p.",nt,'.dat'
    open(newunit=u, file=file_name, access='STREAM', status='UNKNOWN')
    if (RecNo == 1) rewind u
    write(u) psi_phys
    close(u)
end subroutine WriteQP

! WriteQPmodal (QGN_Module.f90:381-446): modal q, psi and Jacobian of q_hat
! as q_mode.%09d.dat, p_mode.%09d.dat and J_mode.%09d.dat.
subroutine WriteQPmodal(q_hat, RecNo, nt)
    complex(dp), intent(in) :: q_hat(nx/2+1,ny,nz)
    integer,     intent(in) :: RecNo, nt
    real(dp), allocatable :: q_mode(:,:,:), psi_mode(:,:,:), j_mode(:,:,:)
    allocate(q_mode(nx,ny,nz), psi_mode(nx,ny,nz), j_mode(nx,ny,nz))
    call qgn_check(qgn_set_state(ctx, q_hat), 'qgn_set_state')
    call qgn_check(qgn_modal_fields(ctx, q_mode, psi_mode, j_mode), 'qgn_modal_fields')
    call write_field("q_mode.", nt, RecNo, q_mode)
    call write_field("p_mode.", nt, RecNo, psi_mode)
    call write_field("J_mode.", nt, RecNo, j_mode)
end subroutine WriteQPmodal

! Diagnostics for the hook of QGN_Driver.f90:52-55: isotropic KE and
! enstrophy spectra of the device state, per layer, as spec.%09d.dat holding
! ke(nbins,nz) then ens(nbins,nz); bin b covers |k| within dk/2 of b*dk.
subroutine WriteSpectra(nt)
    integer, intent(in) :: nt
    real(dp), allocatable :: ke(:,:), ens(:,:)
    character(len=50) :: file_name
    integer :: u, nbins
    real(dp) :: dkx, dky
    dkx = 2._dp*pi/Lx
    dky = 2._dp*pi/Ly
    nbins = floor(hypot(0.5_dp*nx*dkx, 0.5_dp*ny*dky)/min(dkx, dky) + 0.5_dp) + 1
    allocate(ke(nbins,nz), ens(nbins,nz))
    call qgn_check(qgn_spectra(ctx, int(nbins, c_int), ke, ens), 'qgn_spectra')
    write (file_name,'(A5,I0.9,A4)') "spec.", nt, '.dat'
    open(newunit=u, file=file_name, access='STREAM', status='REPLACE')
    write(u) ke, ens
    close(u)
end subroutine WriteSpectra

subroutine write_field(prefix, nt, RecNo, f)
    character(len=*), intent(in) :: prefix
    integer,          intent(in) :: nt, RecNo
    real(dp),         intent(in) :: f(nx,ny,nz)
    character(len=50) :: file_name
    integer :: u
    write (file_name,'(A,I0.9,A4)') prefix, nt, '.dat'
    open(newunit=u, file=file_name, access='STREAM', status='UNKNOWN')
    if (RecNo == 1) rewind u
    write(u) f
    close(u)
end subroutine write_field

! Cleanup (QGN_Module.f90:448-458)
subroutine Cleanup
    print *, '------------------------'
    print *, ' Cleaning up '
    if (c_associated(ctx)) call qgn_check(qgn_destroy(ctx), 'qgn_destroy')
    ctx = c_null_ptr
    print *, '------------------------'
end subroutine Cleanup

end module QGN_Module
