! solid_hip.f90 -- a drop-in for the reference's module SolidSolver
! (SolidSolver.f90) on the library's host solid solver (fsibe_solid_*):
! the same module name and the public procedures the time loop calls,
!     call InitializeSolidSolver(beams)             time_step.f90:72
!     call GravityImpulse(time)                     time_step.f90:140, 209
!     call SolidSolverWrapper(time, dt1, beams)     time_step.f90:141, 210
! so that a build without MKL (which the reference's IBES_solver and
! Solve_LGS need) compiles this file in place of SolidSolver.f90.  The beam
! (type Solid, share_vars.f90:50-63) stays the caller's: every call copies
! its arrays into the solver and back.  One beam (iBeam = 1).
module SolidSolver
  use iso_c_binding
  use share_vars
  use fsibe_hip_api
  implicit none
  type(c_ptr), save :: fsibe_beam = c_null_ptr

contains

  ! the library's parameter record from the share_vars values the solid
  ! solver reads (geometry, kinematics, material, time scheme)
  subroutine solid_params(p)
    type(fsibe_params_t), intent(out) :: p
    p%nx = nx; p%ny = ny; p%ns = ns
    p%iMeanVelocity = iMeanVelocity; p%iSponge = iSponge; p%iBeam = iBeam
    p%iWalls = iWalls; p%iCylinder = iCylinder; p%iSharpTrailing = iSharpTrailing
    p%iFLUSI = iFLUSI; p%iMotion = iMotion; p%iViscous = iViscous
    p%iSaveBeam = iSaveBeam; p%idealis = idealis
    p%xl = xl; p%yl = yl; p%nu = nu; p%eps = eps; p%eps_sponge = eps_sponge
    p%cfl = cfl; p%dt_fixed = dt_fixed; p%Time_end = Time_end
    p%N_smooth = N_smooth; p%t_beam = t_beam; p%length = length
    p%SpongeSize = SpongeSize; p%R_cylinder = R_cylinder; p%h_channel = h_channel
    p%x0 = x0; p%y0 = y0; p%AngleBeam = AngleBeam; p%theta_inf = theta_inf
    p%T_fullspeed = T_fullspeed; p%T_release = T_release; p%tau = tau
    p%U_mean_true = U_mean_true
    p%heave_amp = 0.0_c_double; p%heave_freq = 0.0_c_double; p%motion_coef = 0.0_c_double
    p%mue = mue; p%eta = eta; p%sigma = sigma; p%grav = grav
    p%TimeMethodSolid = TimeMethodSolid; p%iImpulse = iImpulse
    p%iSmoothCutoff = 0
  end subroutine solid_params

  ! InitializeSolidSolver (SolidSolver.f90:54-67): the next step is the first
  subroutine InitializeSolidSolver(beams)
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    type(fsibe_params_t) :: p
    integer :: i
    if (iBeam > 1) then
      write (*, '(A)') "solid_hip: one beam only (iBeam = 1)"
      stop
    end if
    if (.not. c_associated(fsibe_beam)) then
      call solid_params(p)
      call fsibe_check(fsibe_solid_create(p, fsibe_beam), "InitializeSolidSolver")
    end if
    do i = 1, iBeam
      beams(i)%StartupStep = .true.
    end do
  end subroutine InitializeSolidSolver

  ! GravityImpulse (SolidSolver.f90:874-911): sets the share_vars grav
  subroutine GravityImpulse(time)
    real(kind=pr), intent(in) :: time
    type(fsibe_params_t) :: p
    call solid_params(p)
    call fsibe_check(fsibe_gravity_impulse(p, time, grav), "GravityImpulse")
  end subroutine GravityImpulse

  ! SolidSolverWrapper (SolidSolver.f90:13-44): IBES step or the held
  ! beam's leading-edge motion, then SolidEnergies
  subroutine SolidSolverWrapper(time, dt, beams)
    real(kind=pr), intent(in) :: time, dt
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    type(fsibe_solid_info_t) :: info
    if (.not. c_associated(fsibe_beam)) call InitializeSolidSolver(beams)
    call put(FSIBE_SOLID_X, beams(1)%x)
    call put(FSIBE_SOLID_Y, beams(1)%y)
    call put(FSIBE_SOLID_VX, beams(1)%vx)
    call put(FSIBE_SOLID_VY, beams(1)%vy)
    call put(FSIBE_SOLID_THETA, beams(1)%theta)
    call put(FSIBE_SOLID_THETA_DOT, beams(1)%theta_dot)
    call put(FSIBE_SOLID_AX, beams(1)%ax)
    call put(FSIBE_SOLID_AY, beams(1)%ay)
    call put(FSIBE_SOLID_PRESSURE_OLD, beams(1)%pressure_old)
    call put(FSIBE_SOLID_PRESSURE_NEW, beams(1)%pressure_new)
    call put(FSIBE_SOLID_TAU_OLD, beams(1)%tau_old)
    call put(FSIBE_SOLID_TAU_NEW, beams(1)%tau_new)
    call fsibe_check(fsibe_solid_set(fsibe_beam, FSIBE_SOLID_OLDOLD, beams(1)%beam_oldold), "SolidSolverWrapper")
    info%dt_old = beams(1)%dt_old
    info%grav = grav
    info%StartupStep = merge(1_c_int32_t, 0_c_int32_t, beams(1)%StartupStep)
    call fsibe_check(fsibe_solid_set_info(fsibe_beam, info), "SolidSolverWrapper")
    call fsibe_check(fsibe_solid_wrapper(fsibe_beam, time, dt), "SolidSolverWrapper")
    call get(FSIBE_SOLID_X, beams(1)%x)
    call get(FSIBE_SOLID_Y, beams(1)%y)
    call get(FSIBE_SOLID_VX, beams(1)%vx)
    call get(FSIBE_SOLID_VY, beams(1)%vy)
    call get(FSIBE_SOLID_THETA, beams(1)%theta)
    call get(FSIBE_SOLID_THETA_DOT, beams(1)%theta_dot)
    call get(FSIBE_SOLID_AX, beams(1)%ax)
    call get(FSIBE_SOLID_AY, beams(1)%ay)
    call fsibe_check(fsibe_solid_get(fsibe_beam, FSIBE_SOLID_OLDOLD, beams(1)%beam_oldold), "SolidSolverWrapper")
    call fsibe_check(fsibe_solid_get_info(fsibe_beam, info), "SolidSolverWrapper")
    beams(1)%dt_old = info%dt_old
    beams(1)%StartupStep = info%StartupStep /= 0
    beams(1)%E_kinetic = info%E_kinetic
    beams(1)%E_pot = info%E_pot
    beams(1)%E_elastic = info%E_elastic
    beams(1)%Inertial_Force = info%Inertial_Force
    ! the emergency brake ends the loop after this step (SolidSolver.f90:370-380)
    if (info%brake /= 0) continue_timestep = .false.
  contains
    subroutine put(which, a)
      integer(c_int), intent(in) :: which
      real(kind=pr), dimension(0:ns-1), intent(in) :: a
      call fsibe_check(fsibe_solid_set(fsibe_beam, which, a), "SolidSolverWrapper")
    end subroutine put
    subroutine get(which, a)
      integer(c_int), intent(in) :: which
      real(kind=pr), dimension(0:ns-1), intent(out) :: a
      call fsibe_check(fsibe_solid_get(fsibe_beam, which, a), "SolidSolverWrapper")
    end subroutine get
  end subroutine SolidSolverWrapper

end module SolidSolver
