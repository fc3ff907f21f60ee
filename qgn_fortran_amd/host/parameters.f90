!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!
! Run-time parameters of the QGN host (the names and default values of the
! reference's compile-time parameters.f90).  A namelist file qgn.nml in the
! working directory may override any of them (group &qgn).
!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!
    integer, parameter :: dp = C_DOUBLE
    real(dp), parameter :: pi = 4._dp*atan(1.0_dp)

! Physical & grid parameters:
    real(dp) :: Lx = 1.3824E6_dp, Ly = 1.3824E6_dp, Htot = 4800._dp
    ! Grid: the 2/3 rule is exact when nx, ny are multiples of 3
    integer :: nx = (3*256)/2, ny = (3*256)/2, nz = 32
    real(dp) :: f0 = 1.E-4_dp, beta = 2.E-11_dp
    real(dp) :: g = 9.81_dp, rho0 = 1028.8_dp
    real(dp) :: A2 = 0._dp, A8 = 1.E23_dp, r0 = 1.E-7_dp
    real(dp) :: C_d = 0.00001_dp
    real(dp) :: TOL = 1.E-10_dp
    logical :: is_explicit = .FALSE.
    ! CFL for the fastest Rossby wave; 0 means pi/(beta*Lx)
    real(dp) :: dt_max = 0._dp
    ! GPU arithmetic: 0 = FP64, 1 = FP32 FFT/Jacobian with FP64 state and stages
    integer :: precision = 0
    ! bound on rejected attempts per step
    integer :: max_rejects = 1000
    ! copy the state back to the caller's q_hat after every TimeStep
    logical :: sync_host = .TRUE.

    namelist /qgn/ Lx, Ly, Htot, nx, ny, nz, f0, beta, g, rho0, A2, A8, r0, C_d, TOL, &
                   is_explicit, dt_max, precision, max_rejects, sync_host
