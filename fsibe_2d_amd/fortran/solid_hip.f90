! solid_hip.f90 -- a drop-in for the reference's module SolidSolver
! (SolidSolver.f90) on the library's host solid solver (fsibe_solid_*):
! the same module name and the public procedures the time loop calls,
!     call InitializeSolidSolver(beams)             time_step.f90:72
!     call GravityImpulse(time)                     time_step.f90:140, 209
!     call SolidSolverWrapper(time, dt1, beams)     time_step.f90:141, 210
! so that a build without MKL (which the reference's IBES_solver and
! Solve_LGS need) compiles this file in place of SolidSolver.f90.  The beam
! (type Solid, share_vars.f90:50-63) stays the caller's: every call copies
! its arrays into the solver and back.  One solver per beam, up to two beams
! (init_beam.f90:44-76 places the second).
module SolidSolver
  use iso_c_binding
  use share_vars
  use fsibe_hip_api
  implicit none
  type(c_ptr), save :: fsibe_beams(2) = c_null_ptr

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
    if (iBeam > 2) then
      write (*, '(A)') "solid_hip: at most two beams"
      stop
    end if
    call solid_params(p)
    do i = 1, iBeam
      if (.not. c_associated(fsibe_beams(i))) &
        call fsibe_check(fsibe_solid_create_beam(p, int(i - 1, c_int), fsibe_beams(i)), "InitializeSolidSolver")
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

  ! SolidSolverWrapper (SolidSolver.f90:13-44): per beam (:25-33) an IBES
  ! step or the beam's leading-edge motion, then SolidEnergies
  subroutine SolidSolverWrapper(time, dt, beams)
    real(kind=pr), intent(in) :: time, dt
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    type(fsibe_solid_info_t) :: info
    type(c_ptr) :: h
    integer :: i
    if (.not. c_associated(fsibe_beams(1))) call InitializeSolidSolver(beams)
    do i = 1, iBeam
      h = fsibe_beams(i)
      call put(FSIBE_SOLID_X, beams(i)%x)
      call put(FSIBE_SOLID_Y, beams(i)%y)
      call put(FSIBE_SOLID_VX, beams(i)%vx)
      call put(FSIBE_SOLID_VY, beams(i)%vy)
      call put(FSIBE_SOLID_THETA, beams(i)%theta)
      call put(FSIBE_SOLID_THETA_DOT, beams(i)%theta_dot)
      call put(FSIBE_SOLID_AX, beams(i)%ax)
      call put(FSIBE_SOLID_AY, beams(i)%ay)
      call put(FSIBE_SOLID_PRESSURE_OLD, beams(i)%pressure_old)
      call put(FSIBE_SOLID_PRESSURE_NEW, beams(i)%pressure_new)
      call put(FSIBE_SOLID_TAU_OLD, beams(i)%tau_old)
      call put(FSIBE_SOLID_TAU_NEW, beams(i)%tau_new)
      call fsibe_check(fsibe_solid_set(h, FSIBE_SOLID_OLDOLD, beams(i)%beam_oldold), "SolidSolverWrapper")
      info%dt_old = beams(i)%dt_old
      info%grav = grav
      info%StartupStep = merge(1_c_int32_t, 0_c_int32_t, beams(i)%StartupStep)
      call fsibe_check(fsibe_solid_set_info(h, info), "SolidSolverWrapper")
      call fsibe_check(fsibe_solid_wrapper(h, time, dt), "SolidSolverWrapper")
      call get(FSIBE_SOLID_X, beams(i)%x)
      call get(FSIBE_SOLID_Y, beams(i)%y)
      call get(FSIBE_SOLID_VX, beams(i)%vx)
      call get(FSIBE_SOLID_VY, beams(i)%vy)
      call get(FSIBE_SOLID_THETA, beams(i)%theta)
      call get(FSIBE_SOLID_THETA_DOT, beams(i)%theta_dot)
      call get(FSIBE_SOLID_AX, beams(i)%ax)
      call get(FSIBE_SOLID_AY, beams(i)%ay)
      call fsibe_check(fsibe_solid_get(h, FSIBE_SOLID_OLDOLD, beams(i)%beam_oldold), "SolidSolverWrapper")
      call fsibe_check(fsibe_solid_get_info(h, info), "SolidSolverWrapper")
      beams(i)%dt_old = info%dt_old
      beams(i)%StartupStep = info%StartupStep /= 0
      beams(i)%E_kinetic = info%E_kinetic
      beams(i)%E_pot = info%E_pot
      beams(i)%E_elastic = info%E_elastic
      beams(i)%Inertial_Force = info%Inertial_Force
      ! the emergency brake ends the loop after this step (SolidSolver.f90:370-380)
      if (info%brake /= 0) continue_timestep = .false.
    end do
  contains
    subroutine put(which, a)
      integer(c_int), intent(in) :: which
      real(kind=pr), dimension(0:ns-1), intent(in) :: a
      call fsibe_check(fsibe_solid_set(h, which, a), "SolidSolverWrapper")
    end subroutine put
    subroutine get(which, a)
      integer(c_int), intent(in) :: which
      real(kind=pr), dimension(0:ns-1), intent(out) :: a
      call fsibe_check(fsibe_solid_get(h, which, a), "SolidSolverWrapper")
    end subroutine get
  end subroutine SolidSolverWrapper

end module SolidSolver
