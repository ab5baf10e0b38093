! drive_step_hip.f90 -- the reference's time loop (time_step.f90:120-230) for
! prescribed kinematics, written against the reference's own modules
! (share_vars, motion, integrate_position, compiled from the reference by
! oracle/ref.mk) with the fluid step on the GPU through
! fsibe_2d_amd/fortran/fluid_hip.f90.  The initial vorticity is the packed
! coftxy (through cof_hip.f90) of a fixed field.  Writes the state after the
! startup step and 9 AB2 steps to the stream file named on the command line;
! tests/test_fortran_binding.py runs the oracle on the same inputs.
program drive_step_hip
  use share_vars
  use motion
  use fluid_hip
  implicit none
  type(solid), dimension(1:1) :: beams
  real(kind=pr), allocatable :: f(:,:), vortk(:,:), out(:,:)
  real(kind=pr) :: time, dt0, dt1, alpha, alpha_t, alpha_tt, LE(1:6)
  character(len=512) :: path
  integer :: ix, iy, it, u, k

  call get_command_argument(1, path)
  nx = 64; ny = 64; ns = 24
  xl = 4.d0; yl = 4.d0; nu = 5.d-3; eps = 1.d-3; eps_sponge = 1.d-3
  cfl = 1.d0; dt_fixed = 2.d-4; Time_end = 1.d6
  N_smooth = 2.d0; t_beam = 0.05d0; length = 1.d0; SpongeSize = 0.4d0
  R_cylinder = 0.d0; h_channel = 0.d0; x0 = 1.d0; y0 = 2.d0; AngleBeam = 5.d0
  theta_inf = 0.d0; T_fullspeed = 0.d0; T_release = 0.d0; tau = 0.d0; U_mean_true = 1.d0
  iMeanVelocity = 3; iSponge = 3; iBeam = 1; iWalls = 0; iCylinder = 0; iSharpTrailing = 0
  iFLUSI = 0; iMotion = 2; iViscous = 1; iSaveBeam = 1; idealis = 1
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

  ! one beam as init_beam.f90:130-167 builds it
  allocate(beams(1)%x(0:ns-1), beams(1)%y(0:ns-1), beams(1)%vx(0:ns-1), beams(1)%vy(0:ns-1))
  allocate(beams(1)%theta(0:ns-1), beams(1)%theta_dot(0:ns-1))
  allocate(beams(1)%pressure_new(0:ns-1), beams(1)%tau_new(0:ns-1))
  beams(1)%x0 = x0; beams(1)%y0 = y0; beams(1)%AngleBeam = AngleBeam; beams(1)%phase = 0.d0
  beams(1)%iMouvement = iMotion
  call mouvement(0.d0, alpha, alpha_t, alpha_tt, LE, beams(1))
  do k = 0, ns-1
    beams(1)%x(k) = LE(1) + real(k, pr)*ds
    beams(1)%y(k) = LE(2)
  end do
  beams(1)%vx = 0.d0; beams(1)%vy = 0.d0; beams(1)%theta = 0.d0; beams(1)%theta_dot = 0.d0
  call integrate_position(0.d0, beams(1))

  call hip_fluid_create(vortk)
  time = 0.d0
  ! startup, time_step.f90:120-162
  call hip_create_mask(time, beams)
  call hip_evolve_explicit(time, dt1, beams)
  call hip_get_forces(time, beams)
  call integrate_position(time, beams(1))
  call hip_swap_levels()
  time = time + dt1
  ! time loop, time_step.f90:178-230
  do it = 1, 9
    dt0 = dt1
    call hip_create_mask(time, beams)
    call hip_evolve_ab2(time, dt0, dt1, beams)
    call hip_get_forces(time, beams)
    call integrate_position(time, beams(1))
    call hip_swap_levels()
    time = time + dt1
  end do

  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
  write(u) nx, ny, ns
  write(u) f
  write(u) time, dt1
  call hip_download(FSIBE_FIELD_VORTK, out); write(u) out
  write(u) beams(1)%Force, beams(1)%Force_unst, beams(1)%Force_press
  write(u) beams(1)%pressure_new, beams(1)%tau_new, beams(1)%x, beams(1)%y
  close(u)
  call hip_fluid_destroy()
  call fft_free
end program drive_step_hip
