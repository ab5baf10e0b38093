! fsibe_hip_api.f90 -- iso_c_binding interfaces of include/fsibe_hip.h for a
! Fortran host (the reference is Fortran 90 compiled -r8: REAL(8) = c_double).
!
! Arrays are passed by address as the reference's implicit interfaces do
! (call coftxy(penal(:,:,2), work2), NavierStokes.f90:127): explicit-shape
! dummies f(0:nx-1,0:ny-1) associate with the assumed-size f(*) below.
! fsibe_check turns a nonzero return code into the reference's fatal-error
! behaviour, a message and STOP (e.g. Interpolation.f90:19-24).
module fsibe_hip_api
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FSIBE_OK = 0, FSIBE_EARG = 1, FSIBE_ESIZE = 2, FSIBE_EHIP = 3, &
                               FSIBE_EUNSUPPORTED = 4, FSIBE_EDOMAIN = 5, FSIBE_ESTATE = 6
  integer(c_int), parameter :: FSIBE_FIELD_VORTK = 0, FSIBE_FIELD_NLK = 1, FSIBE_FIELD_U1 = 3, &
                               FSIBE_FIELD_U2 = 4, FSIBE_FIELD_PRESS = 5, FSIBE_FIELD_MASK = 6, &
                               FSIBE_FIELD_MASKVX = 7, FSIBE_FIELD_MASKVY = 8, &
                               FSIBE_FIELD_MASK_SPONGE = 9, FSIBE_FIELD_VOR_MEAN = 10, &
                               FSIBE_FIELD_VORT = 14

  ! struct fsibe_params (share_vars.f90:8-25 names)
  type, bind(C) :: fsibe_params_t
    integer(c_int32_t) :: nx, ny, ns
    integer(c_int32_t) :: iMeanVelocity, iSponge, iBeam, iWalls, iCylinder, iSharpTrailing
    integer(c_int32_t) :: iFLUSI, iMotion, iViscous, iSaveBeam, idealis
    real(c_double) :: xl, yl, nu, eps, eps_sponge, cfl, dt_fixed, Time_end
    real(c_double) :: N_smooth, t_beam, length, SpongeSize, R_cylinder, h_channel
    real(c_double) :: x0, y0, AngleBeam, theta_inf, T_fullspeed, T_release, tau
    real(c_double) :: U_mean_true
    real(c_double) :: heave_amp, heave_freq
    real(c_double) :: motion_coef(8)
    real(c_double) :: mue, eta, sigma, grav
    integer(c_int32_t) :: TimeMethodSolid, iImpulse
    integer(c_int32_t) :: iSmoothCutoff
  end type

  ! struct fsibe_step_out
  type, bind(C) :: fsibe_step_out_t
    real(c_double) :: dt1, u_max
    real(c_double) :: Force(2), Force_unst(2), Force_press(2)
    real(c_double) :: energies(6)
    real(c_double) :: bsi(10)
    real(c_double) :: fx_int, fy_int
    integer(c_int32_t) :: status, ibes_iter
  end type

  ! struct fsibe_solid_info
  type, bind(C) :: fsibe_solid_info_t
    real(c_double) :: dt_old, grav, E_kinetic, E_pot, E_elastic, Inertial_Force(2)
    integer(c_int32_t) :: StartupStep, iter, brake, method
  end type

  integer(c_int), parameter :: FSIBE_SOLID_X = 0, FSIBE_SOLID_Y = 1, FSIBE_SOLID_VX = 2, FSIBE_SOLID_VY = 3, &
                               FSIBE_SOLID_THETA = 4, FSIBE_SOLID_THETA_DOT = 5, FSIBE_SOLID_AX = 6, &
                               FSIBE_SOLID_AY = 7, FSIBE_SOLID_PRESSURE_OLD = 8, FSIBE_SOLID_PRESSURE_NEW = 9, &
                               FSIBE_SOLID_TAU_OLD = 10, FSIBE_SOLID_TAU_NEW = 11, FSIBE_SOLID_OLDOLD = 12

  interface
    function fsibe_last_error() bind(C, name="fsibe_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    ! ---- Tier A: FFT / spectral-operator plugin ----------------------------
    function fsibe_spectral_create(nx, ny, xl, yl, nu, ctx) bind(C, name="fsibe_spectral_create") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny
      real(c_double), value :: xl, yl, nu
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function
    function fsibe_spectral_destroy(ctx) bind(C, name="fsibe_spectral_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function fsibe_coftxy(ctx, f, fk) bind(C, name="fsibe_coftxy") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fk(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofitxy(ctx, fk, f) bind(C, name="fsibe_cofitxy") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: f(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofdx(ctx, fk, out) bind(C, name="fsibe_cofdx") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofdy(ctx, fk, out) bind(C, name="fsibe_cofdy") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofdxdx(ctx, fk, out) bind(C, name="fsibe_cofdxdx") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofdydy(ctx, fk, out) bind(C, name="fsibe_cofdydy") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofdxdy(ctx, fk, out) bind(C, name="fsibe_cofdxdy") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_poisson(ctx, f, ans) bind(C, name="fsibe_poisson") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: ans(*)
      integer(c_int) :: rc
    end function
    function fsibe_vis(ctx, dt, work) bind(C, name="fsibe_vis") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      real(c_double), intent(out) :: work(*)
      integer(c_int) :: rc
    end function
    function fsibe_dealiase_mask(ctx, dealiase) bind(C, name="fsibe_dealiase_mask") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dealiase(*)
      integer(c_int) :: rc
    end function
    function fsibe_cofts(f, fk, L, n) bind(C, name="fsibe_cofts") result(rc)
      import :: c_int, c_double
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fk(*)
      integer(c_int), value :: L, n
      integer(c_int) :: rc
    end function
    function fsibe_cofits(fk, f, L, n) bind(C, name="fsibe_cofits") result(rc)
      import :: c_int, c_double
      real(c_double), intent(in) :: fk(*)
      real(c_double), intent(out) :: f(*)
      integer(c_int), value :: L, n
      integer(c_int) :: rc
    end function

    ! ---- Tier B: device-resident fluid step -------------------------------
    function fsibe_sim_create(p, sim) bind(C, name="fsibe_sim_create") result(rc)
      import :: c_int, c_ptr, fsibe_params_t
      type(fsibe_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: sim
      integer(c_int) :: rc
    end function
    function fsibe_sim_destroy(sim) bind(C, name="fsibe_sim_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function
    function fsibe_set_vortk(sim, vortk) bind(C, name="fsibe_set_vortk") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), intent(in) :: vortk(*)
      integer(c_int) :: rc
    end function
    function fsibe_create_mask(sim, time, x, y, vx, vy, theta, alpha, LeadingEdge) &
        bind(C, name="fsibe_create_mask") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: time, alpha
      real(c_double), intent(in) :: x(*), y(*), vx(*), vy(*), theta(*), LeadingEdge(6)
      integer(c_int) :: rc
    end function
    function fsibe_evolve_explicit(sim, time, dt1) bind(C, name="fsibe_evolve_explicit") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: time
      real(c_double), intent(out) :: dt1
      integer(c_int) :: rc
    end function
    function fsibe_evolve_ab2(sim, time, dt0, dt1) bind(C, name="fsibe_evolve_ab2") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: time, dt0
      real(c_double), intent(out) :: dt1
      integer(c_int) :: rc
    end function
    function fsibe_get_forces(sim, time, pressure_beam, tau_beam, Force_press) &
        bind(C, name="fsibe_get_forces") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: time
      real(c_double), intent(out) :: pressure_beam(*), tau_beam(*), Force_press(*)
      integer(c_int) :: rc
    end function
    function fsibe_get_step_out(sim, out) bind(C, name="fsibe_get_step_out") result(rc)
      import :: c_int, c_ptr, fsibe_step_out_t
      type(c_ptr), value :: sim
      type(fsibe_step_out_t), intent(out) :: out
      integer(c_int) :: rc
    end function
    function fsibe_get_step_out_beam(sim, ib, out) bind(C, name="fsibe_get_step_out_beam") result(rc)
      import :: c_int, c_ptr, fsibe_step_out_t
      type(c_ptr), value :: sim
      integer(c_int), value :: ib
      type(fsibe_step_out_t), intent(out) :: out
      integer(c_int) :: rc
    end function
    function fsibe_create_mask_beams(sim, time, x, y, vx, vy, theta, alpha, LeadingEdge) &
        bind(C, name="fsibe_create_mask_beams") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: time
      real(c_double), intent(in) :: x(*), y(*), vx(*), vy(*), theta(*), alpha(*), LeadingEdge(*)
      integer(c_int) :: rc
    end function
    function fsibe_swap_levels(sim) bind(C, name="fsibe_swap_levels") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: rc
    end function
    function fsibe_get_field(sim, which, out) bind(C, name="fsibe_get_field") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: which
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_run(sim, nsteps, out) bind(C, name="fsibe_run") result(rc)
      import :: c_int, c_ptr, fsibe_step_out_t
      type(c_ptr), value :: sim
      integer(c_int), value :: nsteps
      type(fsibe_step_out_t), intent(out) :: out
      integer(c_int) :: rc
    end function

    ! ---- host solid solver ----------------------------------------------------
    function fsibe_solid_create(p, b) bind(C, name="fsibe_solid_create") result(rc)
      import :: c_int, c_ptr, fsibe_params_t
      type(fsibe_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: b
      integer(c_int) :: rc
    end function
    function fsibe_solid_create_beam(p, ib, b) bind(C, name="fsibe_solid_create_beam") result(rc)
      import :: c_int, c_ptr, fsibe_params_t
      type(fsibe_params_t), intent(in) :: p
      integer(c_int), value :: ib
      type(c_ptr), intent(out) :: b
      integer(c_int) :: rc
    end function
    function fsibe_solid_destroy(b) bind(C, name="fsibe_solid_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function
    function fsibe_solid_get(b, which, out) bind(C, name="fsibe_solid_get") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: which
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function fsibe_solid_set(b, which, in) bind(C, name="fsibe_solid_set") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: which
      real(c_double), intent(in) :: in(*)
      integer(c_int) :: rc
    end function
    function fsibe_solid_get_info(b, info) bind(C, name="fsibe_solid_get_info") result(rc)
      import :: c_int, c_ptr, fsibe_solid_info_t
      type(c_ptr), value :: b
      type(fsibe_solid_info_t), intent(out) :: info
      integer(c_int) :: rc
    end function
    function fsibe_solid_set_info(b, info) bind(C, name="fsibe_solid_set_info") result(rc)
      import :: c_int, c_ptr, fsibe_solid_info_t
      type(c_ptr), value :: b
      type(fsibe_solid_info_t), intent(in) :: info
      integer(c_int) :: rc
    end function
    function fsibe_solid_step(b, time, dt) bind(C, name="fsibe_solid_step") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), value :: time, dt
      integer(c_int) :: rc
    end function
    function fsibe_solid_wrapper(b, time, dt) bind(C, name="fsibe_solid_wrapper") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), value :: time, dt
      integer(c_int) :: rc
    end function
    function fsibe_gravity_impulse(p, time, grav) bind(C, name="fsibe_gravity_impulse") result(rc)
      import :: c_int, c_double, fsibe_params_t
      type(fsibe_params_t), intent(in) :: p
      real(c_double), value :: time
      real(c_double), intent(inout) :: grav
      integer(c_int) :: rc
    end function
    ! discrete adjoint (no reference counterpart): J = (1/T) sum(-Force(1) dt)
    ! over nsteps steps of motion 601 and dJ/dcoef(8), checkpoints every chk
    function fsibe_adjoint_create(p, vortk, nsteps, chk, adj) bind(C, name="fsibe_adjoint_create") result(rc)
      import :: c_int, c_double, c_ptr, fsibe_params_t
      type(fsibe_params_t), intent(in) :: p
      real(c_double), intent(in) :: vortk(*)
      integer(c_int), value :: nsteps, chk
      type(c_ptr), intent(out) :: adj
      integer(c_int) :: rc
    end function
    ! the same with at most nsnap checkpoint slots (binomial schedule;
    ! nsnap < 0: one per unit of chk steps); identical results
    function fsibe_adjoint_create_binomial(p, vortk, nsteps, chk, nsnap, adj) &
        bind(C, name="fsibe_adjoint_create_binomial") result(rc)
      import :: c_int, c_double, c_ptr, fsibe_params_t
      type(fsibe_params_t), intent(in) :: p
      real(c_double), intent(in) :: vortk(*)
      integer(c_int), value :: nsteps, chk, nsnap
      type(c_ptr), intent(out) :: adj
      integer(c_int) :: rc
    end function
    function fsibe_adjoint_destroy(adj) bind(C, name="fsibe_adjoint_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: adj
      integer(c_int) :: rc
    end function
    function fsibe_adjoint_objective(adj, coef, J) bind(C, name="fsibe_adjoint_objective") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: adj
      real(c_double), intent(in) :: coef(8)
      real(c_double), intent(out) :: J
      integer(c_int) :: rc
    end function
    ! restart files (MakeRuntimeBackup, init_fields inicond = 2 and 111)
    function fsibe_backup_write(sim, path, ivideo, colorscale, colorscale_done) &
        bind(C, name="fsibe_backup_write") result(rc)
      import :: c_int, c_double, c_ptr, c_char
      type(c_ptr), value :: sim
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: ivideo, colorscale_done
      real(c_double), value :: colorscale
      integer(c_int) :: rc
    end function
    function fsibe_backup_read(sim, path, ivideo, colorscale, colorscale_done) &
        bind(C, name="fsibe_backup_read") result(rc)
      import :: c_int, c_double, c_ptr, c_char
      type(c_ptr), value :: sim
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: ivideo, colorscale_done
      real(c_double), intent(out) :: colorscale
      integer(c_int) :: rc
    end function
    function fsibe_inicond_write(sim, path) bind(C, name="fsibe_inicond_write") result(rc)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: sim
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function
    function fsibe_inicond_read(sim, path) bind(C, name="fsibe_inicond_read") result(rc)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: sim
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function
    ! ComputePressureSnapshot and SaveField
    function fsibe_pressure_snapshot(sim, time, level, press) bind(C, name="fsibe_pressure_snapshot") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: sim
      real(c_double), value :: time
      integer(c_int), value :: level
      real(c_double), intent(out) :: press(*)
      integer(c_int) :: rc
    end function
    function fsibe_save_field(path, field, nx, ny, nsave, xl, yl, fieldtype, header) &
        bind(C, name="fsibe_save_field") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*), fieldtype(*), header(*)
      real(c_double), intent(in) :: field(*)
      integer(c_int), value :: nx, ny, nsave
      real(c_double), value :: xl, yl
      integer(c_int) :: rc
    end function
    function fsibe_save_gif(path, field, nx, ny, palette, have_range, color_min, color_max) &
        bind(C, name="fsibe_save_gif") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(in) :: field(*)
      integer(c_int), value :: nx, ny, palette, have_range
      real(c_double), value :: color_min, color_max
      integer(c_int) :: rc
    end function
    function fsibe_adjoint_gradient(adj, coef, J, grad) bind(C, name="fsibe_adjoint_gradient") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: adj
      real(c_double), intent(in) :: coef(8)
      real(c_double), intent(out) :: J, grad(8)
      integer(c_int) :: rc
    end function
  end interface

contains

  ! message of the last failed call
  function fsibe_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    msg = ' '
    call c_f_pointer(fsibe_last_error(), cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function

  ! the reference's reaction to a fatal error: message, then STOP
  subroutine fsibe_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= FSIBE_OK) then
      write (*, '("fsibe_hip: ",A,": error ",I0,": ",A)') where, rc, trim(fsibe_message())
      stop
    end if
  end subroutine

end module fsibe_hip_api
