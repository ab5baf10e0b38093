! drive_step_hip.f90 -- the reference's time loop (time_step.f90:120-230) for
! prescribed kinematics, written against the reference's own modules
! (share_vars, motion, integrate_position, compiled from the reference by
! oracle/ref.mk) with the fluid step on the GPU through
! fsibe_2d_amd/fortran/fluid_hip.f90.  The initial vorticity is the packed
! coftxy (through cof_hip.f90) of a fixed field.  Writes the state after the
! startup step and 9 AB2 steps to the stream file named on the command line;
! tests/test_fortran_binding.py runs the oracle on the same inputs.  A second
! argument "2" runs two beams in tandem (iMotion 303, the second beam placed
! as init_beam.f90:50-76 places it).
program drive_step_hip
  use share_vars
  use motion
  use fluid_hip
  implicit none
  type(solid), dimension(:), allocatable :: beams
  real(kind=pr), allocatable :: f(:,:), vortk(:,:), out(:,:)
  real(kind=pr) :: time, dt0, dt1, alpha, alpha_t, alpha_tt, LE(1:6)
  character(len=512) :: path, arg
  integer :: ix, iy, it, u, k, i, nb

  call get_command_argument(1, path)
  nb = 1
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read (arg, *) nb
  end if
  nx = 64; ny = 64; ns = 24
  xl = 4.d0; yl = 4.d0; nu = 5.d-3; eps = 1.d-3; eps_sponge = 1.d-3
  cfl = 1.d0; dt_fixed = 2.d-4; Time_end = 1.d6
  N_smooth = 2.d0; t_beam = 0.05d0; length = 1.d0; SpongeSize = 0.4d0
  R_cylinder = 0.d0; h_channel = 0.d0; x0 = 1.d0; y0 = 2.d0; AngleBeam = 5.d0
  theta_inf = 0.d0; T_fullspeed = 0.d0; T_release = 0.d0; tau = 0.d0; U_mean_true = 1.d0
  iMeanVelocity = 3; iSponge = 3; iBeam = nb; iWalls = 0; iCylinder = 0; iSharpTrailing = 0
  iFLUSI = 0; iMotion = 2; iViscous = 1; iSaveBeam = 1; idealis = 1
  if (nb == 2) then
    iMotion = 303; x0 = 0.8d0
  end if
  pi = 4.d0 * atan(1.d0)
  scalex = (2.d0*pi/xl)**2
  scaley = (2.d0*pi/yl)**2
  ds = length / real(ns-1, pr)
  allocate(dealiase(0:nx-1, 0:ny-1))
  allocate(f(0:nx-1, 0:ny-1), vortk(0:nx-1, 0:ny-1), out(0:nx-1, 0:ny-1))

  ! initial vorticity: packed transform of a fixed smooth field
  do iy = 0, ny-1
    do ix = 0, nx-1
      f(ix, iy) = 0.3d0*sin(2.d0*pi*ix/nx) * cos(4.d0*pi*iy/ny) + 0.1d0*cos(2.d0*pi*(3*ix+iy)/nx)
    end do
  end do
  call fft_initialize
  call coftxy(f, vortk)

  ! the beams as init_beam.f90:48-76 places them and :130-167 builds them
  allocate(beams(1:nb))
  do i = 1, nb
    allocate(beams(i)%x(0:ns-1), beams(i)%y(0:ns-1), beams(i)%vx(0:ns-1), beams(i)%vy(0:ns-1))
    allocate(beams(i)%theta(0:ns-1), beams(i)%theta_dot(0:ns-1))
    allocate(beams(i)%pressure_new(0:ns-1), beams(i)%tau_new(0:ns-1))
    beams(i)%x0 = x0; beams(i)%y0 = y0; beams(i)%AngleBeam = AngleBeam; beams(i)%phase = 0.d0
    beams(i)%iMouvement = iMotion
    if (i == 2) then
      beams(i)%x0 = x0 + 1.25
      beams(i)%phase = 0.75*pi
    end if
    call mouvement(0.d0, alpha, alpha_t, alpha_tt, LE, beams(i))
    do k = 0, ns-1
      beams(i)%x(k) = LE(1) + real(k, pr)*ds
      beams(i)%y(k) = LE(2)
    end do
    beams(i)%vx = 0.d0; beams(i)%vy = 0.d0; beams(i)%theta = 0.d0; beams(i)%theta_dot = 0.d0
    call integrate_position(0.d0, beams(i))
  end do

  call hip_fluid_create(vortk)
  time = 0.d0
  ! startup, time_step.f90:120-162
  call hip_create_mask(time, beams)
  call hip_evolve_explicit(time, dt1, beams)
  call hip_get_forces(time, beams)
  do i = 1, nb
    call integrate_position(time, beams(i))
  end do
  call hip_swap_levels()
  time = time + dt1
  ! time loop, time_step.f90:178-230
  do it = 1, 9
    dt0 = dt1
    call hip_create_mask(time, beams)
    call hip_evolve_ab2(time, dt0, dt1, beams)
    call hip_get_forces(time, beams)
    do i = 1, nb
      call integrate_position(time, beams(i))
    end do
    call hip_swap_levels()
    time = time + dt1
  end do

  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
  write(u) nx, ny, ns
  write(u) f
  write(u) time, dt1
  call hip_download(FSIBE_FIELD_VORTK, out); write(u) out
  do i = 1, nb
    write(u) beams(i)%Force, beams(i)%Force_unst, beams(i)%Force_press
    write(u) beams(i)%pressure_new, beams(i)%tau_new, beams(i)%x, beams(i)%y
  end do
  close(u)
  call hip_fluid_destroy()
  call fft_free
end program drive_step_hip
