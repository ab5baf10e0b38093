! fluid_hip.f90 -- Tier-B binding for the reference's time loop
! (time_step.f90:178-230): the per-step calls
!     create_mask(time, beams)                                  :187
!     EvolveFluidAB2(time, dt0, dt1, n0, n1, beams, vortk, ...) :193
!     GetForces(time, beams, press, u, "new")                   :199
! become hip_create_mask / hip_evolve_ab2 / hip_get_forces below, with the
! fluid state (vortk, nlk, u, masks, vor_mean, press) resident on the GPU.
! The beam (type Solid, share_vars.f90:50-63) stays on the host and only its
! node arrays and the per-node loads cross the bus.  One beam (iBeam = 1).
module fluid_hip
  use iso_c_binding
  use share_vars
  use fsibe_hip_api
  implicit none
  type(c_ptr), save :: fsibe_sim = c_null_ptr

contains

  ! after params (params.f90) and the initial condition: allocate the device
  ! state and upload vortk(:,:,n0) (packed layout)
  subroutine hip_fluid_create(vortk_n0)
    real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: vortk_n0
    type(fsibe_params_t) :: p
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
    call fsibe_check(fsibe_sim_create(p, fsibe_sim), "fsibe_sim_create")
    call fsibe_check(fsibe_set_vortk(fsibe_sim, vortk_n0), "fsibe_set_vortk")
  end subroutine hip_fluid_create

  subroutine hip_fluid_destroy()
    call fsibe_check(fsibe_sim_destroy(fsibe_sim), "fsibe_sim_destroy")
    fsibe_sim = c_null_ptr
  end subroutine hip_fluid_destroy

  ! create_mask(time, beams) (create_mask.f90:2-218): the leading-edge
  ! kinematics come from the host's own mouvement (mouvement.f90:5-241)
  subroutine hip_create_mask(time, beams)
    use motion
    real(kind=pr), intent(in) :: time
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    real(kind=pr) :: alpha, alpha_t, alpha_tt, LeadingEdge(1:6)
    call mouvement(time, alpha, alpha_t, alpha_tt, LeadingEdge, beams(1))
    call fsibe_check(fsibe_create_mask(fsibe_sim, time, beams(1)%x, beams(1)%y, beams(1)%vx, &
                                       beams(1)%vy, beams(1)%theta, alpha, LeadingEdge), "create_mask")
  end subroutine hip_create_mask

  ! EvolveFluidExplicit(time, dt1, ...) (NavierStokes.f90:422-459)
  subroutine hip_evolve_explicit(time, dt1, beams)
    real(kind=pr), intent(in) :: time
    real(kind=pr), intent(out) :: dt1
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    call fsibe_check(fsibe_evolve_explicit(fsibe_sim, time, dt1), "EvolveFluidExplicit")
    call fetch_drag(beams)
  end subroutine hip_evolve_explicit

  ! EvolveFluidAB2(time, dt0, dt1, ...) (NavierStokes.f90:463-508)
  subroutine hip_evolve_ab2(time, dt0, dt1, beams)
    real(kind=pr), intent(in) :: time, dt0
    real(kind=pr), intent(out) :: dt1
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    call fsibe_check(fsibe_evolve_ab2(fsibe_sim, time, dt0, dt1), "EvolveFluidAB2")
    call fetch_drag(beams)
  end subroutine hip_evolve_ab2

  ! cal_drag's results, stored in type Solid (NavierStokes.f90:567-589)
  subroutine fetch_drag(beams)
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    type(fsibe_step_out_t) :: o
    call fsibe_check(fsibe_get_step_out(fsibe_sim, o), "fsibe_get_step_out")
    if (iBeam > 0) then
      beams(1)%Force = o%Force
      beams(1)%Force_unst = o%Force_unst
    end if
  end subroutine fetch_drag

  ! GetForces(time, beams, press, u, "new") (BeamForces.f90:6-51)
  subroutine hip_get_forces(time, beams)
    real(kind=pr), intent(in) :: time
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    call fsibe_check(fsibe_get_forces(fsibe_sim, time, beams(1)%pressure_new, beams(1)%tau_new, &
                                      beams(1)%Force_press), "GetForces")
  end subroutine hip_get_forces

  ! inter = n1; n1 = n0; n0 = inter (time_step.f90:222)
  subroutine hip_swap_levels()
    call fsibe_check(fsibe_swap_levels(fsibe_sim), "swap")
  end subroutine hip_swap_levels

  ! download for save_fields / diagnostics: which = FSIBE_FIELD_*
  subroutine hip_download(which, field)
    integer(c_int), intent(in) :: which
    real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: field
    call fsibe_check(fsibe_get_field(fsibe_sim, which, field), "fsibe_get_field")
  end subroutine hip_download

end module fluid_hip
