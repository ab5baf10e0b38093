! fluid_hip.f90 -- Tier-B binding for the reference's time loop
! (time_step.f90:178-230): the per-step calls
!     create_mask(time, beams)                                  :187
!     EvolveFluidAB2(time, dt0, dt1, n0, n1, beams, vortk, ...) :193
!     GetForces(time, beams, press, u, "new")                   :199
! become hip_create_mask / hip_evolve_ab2 / hip_get_forces below, with the
! fluid state (vortk, nlk, u, masks, vor_mean, press) resident on the GPU.
! The beam (type Solid, share_vars.f90:50-63) stays on the host and only its
! node arrays and the per-node loads cross the bus.  Up to two beams (the
! reference's loops over beams(1:iBeam)).
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
  ! kinematics come from the host's own mouvement (mouvement.f90:5-241) of
  ! each beam (create_mask.f90:77); the beams' node arrays go side by side
  subroutine hip_create_mask(time, beams)
    use motion
    real(kind=pr), intent(in) :: time
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    real(kind=pr) :: alpha(1:max(iBeam, 1)), alpha_t, alpha_tt, LeadingEdge(1:6, 1:max(iBeam, 1))
    real(kind=pr), dimension(0:ns * max(iBeam, 1) - 1) :: bx, by, bvx, bvy, bth
    integer :: i, o
    if (iBeam < 1) then
      alpha = 0.0_pr
      LeadingEdge = 0.0_pr
      bx = 0.0_pr; by = 0.0_pr; bvx = 0.0_pr; bvy = 0.0_pr; bth = 0.0_pr
    end if
    do i = 1, iBeam
      call mouvement(time, alpha(i), alpha_t, alpha_tt, LeadingEdge(:, i), beams(i))
      o = (i - 1) * ns
      bx(o:o + ns - 1) = beams(i)%x
      by(o:o + ns - 1) = beams(i)%y
      bvx(o:o + ns - 1) = beams(i)%vx
      bvy(o:o + ns - 1) = beams(i)%vy
      bth(o:o + ns - 1) = beams(i)%theta
    end do
    call fsibe_check(fsibe_create_mask_beams(fsibe_sim, time, bx, by, bvx, bvy, bth, alpha, LeadingEdge), &
                     "create_mask")
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
    integer :: i
    do i = 1, iBeam
      call fsibe_check(fsibe_get_step_out_beam(fsibe_sim, int(i - 1, c_int), o), "fsibe_get_step_out_beam")
      beams(i)%Force = o%Force
      beams(i)%Force_unst = o%Force_unst
    end do
  end subroutine fetch_drag

  ! GetForces(time, beams, press, u, "new") (BeamForces.f90:6-51), every beam
  subroutine hip_get_forces(time, beams)
    real(kind=pr), intent(in) :: time
    type(solid), dimension(1:iBeam), intent(inout) :: beams
    real(kind=pr), dimension(0:ns * max(iBeam, 1) - 1) :: pb, tb
    real(kind=pr) :: fp(1:2, 1:max(iBeam, 1))
    integer :: i, o
    call fsibe_check(fsibe_get_forces(fsibe_sim, time, pb, tb, fp), "GetForces")
    do i = 1, iBeam
      o = (i - 1) * ns
      beams(i)%pressure_new = pb(o:o + ns - 1)
      beams(i)%tau_new = tb(o:o + ns - 1)
      beams(i)%Force_press = fp(:, i)
    end do
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
