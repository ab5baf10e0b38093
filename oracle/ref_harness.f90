! ref_harness.f90 -- bind(C) entry points into the reference's own Fortran
! routines, compiled unmodified from /root/reference by oracle/ref.mk into
! oracle/_ref/libfsibe_ref.so.
!
! TEST INFRASTRUCTURE ONLY.  This file is a driver: it sets the reference's
! module globals (share_vars) the way params.f90:139-141,190 and dns.f90:183
! do, then calls the reference's external subroutines.  It restates nothing.
!
! Routines reached: cofdx, cofdy, cofdxdx, cofdydy, cofdxdy (cofd*.f90),
! poisson (poisson.f90), vis (cal_vis.f90), dealiase_mask (dealiase_mask.f90),
! Mean_ux / Mean_uy (mean_velocity.f90), mouvement (mouvement.f90) and
! integrate_position (integrate_position.f90).
module ref_harness
  use iso_c_binding
  use share_vars
  use motion
  implicit none
  type(solid), save :: hb
contains

  subroutine ref_setup(nx_, ny_, ns_, xl_, yl_, nu_, length_) bind(C, name="ref_setup")
    integer(c_int), value :: nx_, ny_, ns_
    real(c_double), value :: xl_, yl_, nu_, length_
    nx = nx_; ny = ny_; ns = ns_
    xl = xl_; yl = yl_; nu = nu_; length = length_
    pi = 4.0 * atan(1.0)
    scalex = (2.0*pi / xl)**2
    scaley = (2.0*pi / yl)**2
    ds = length/(ns-1)
    if (allocated(dealiase)) deallocate(dealiase)
    allocate(dealiase(0:nx-1,0:ny-1))
    if (allocated(hb%x)) then
      deallocate(hb%x, hb%y, hb%vx, hb%vy, hb%theta, hb%theta_dot)
    endif
    allocate(hb%x(0:ns-1), hb%y(0:ns-1), hb%vx(0:ns-1), hb%vy(0:ns-1))
    allocate(hb%theta(0:ns-1), hb%theta_dot(0:ns-1))
  end subroutine

  subroutine ref_set_flow(iMean, theta_inf_, T_fullspeed_) bind(C, name="ref_set_flow")
    integer(c_int), value :: iMean
    real(c_double), value :: theta_inf_, T_fullspeed_
    iMeanVelocity = iMean
    theta_inf = theta_inf_
    T_fullspeed = T_fullspeed_
    u_mean = 0.0
  end subroutine

  subroutine ref_op(which, fk, out) bind(C, name="ref_op")
    integer(c_int), value :: which
    real(c_double), intent(in) :: fk(nx*ny)
    real(c_double), intent(out) :: out(nx*ny)
    select case (which)
    case (0); call cofdx(fk, out)
    case (1); call cofdy(fk, out)
    case (2); call cofdxdx(fk, out)
    case (3); call cofdydy(fk, out)
    case (4); call cofdxdy(fk, out)
    case (5); call poisson(fk, out)
    end select
  end subroutine

  subroutine ref_vis(dt, out) bind(C, name="ref_vis")
    real(c_double), value :: dt
    real(c_double), intent(out) :: out(nx*ny)
    call vis(dt, out)
  end subroutine

  subroutine ref_dealiase(out) bind(C, name="ref_dealiase")
    real(c_double), intent(out) :: out(nx*ny)
    dealiase = 1.0d0
    call dealiase_mask
    out = reshape(dealiase, (/ nx*ny /))
  end subroutine

  function ref_mean_ux(t) bind(C, name="ref_mean_ux") result(v)
    real(c_double), value :: t
    real(c_double) :: v
    real(kind=pr) :: Mean_ux
    v = Mean_ux(t)
  end function

  function ref_mean_uy(t) bind(C, name="ref_mean_uy") result(v)
    real(c_double), value :: t
    real(c_double) :: v
    real(kind=pr) :: Mean_uy
    v = Mean_uy(t)
  end function

  subroutine ref_mouvement(imouv, x0_, y0_, angle_, phase_, t, alpha, alpha_t, alpha_tt, le) &
      bind(C, name="ref_mouvement")
    integer(c_int), value :: imouv
    real(c_double), value :: x0_, y0_, angle_, phase_, t
    real(c_double), intent(out) :: alpha, alpha_t, alpha_tt, le(6)
    hb%iMouvement = imouv
    hb%x0 = x0_; hb%y0 = y0_; hb%AngleBeam = angle_; hb%phase = phase_
    call mouvement(t, alpha, alpha_t, alpha_tt, le, hb)
  end subroutine

  subroutine ref_integrate_position(imouv, x0_, y0_, angle_, phase_, t, theta, theta_dot, &
      x, y, vx, vy) bind(C, name="ref_integrate_position")
    integer(c_int), value :: imouv
    real(c_double), value :: x0_, y0_, angle_, phase_, t
    real(c_double), intent(in) :: theta(ns), theta_dot(ns)
    real(c_double), intent(out) :: x(ns), y(ns), vx(ns), vy(ns)
    hb%iMouvement = imouv
    hb%x0 = x0_; hb%y0 = y0_; hb%AngleBeam = angle_; hb%phase = phase_
    hb%theta = theta
    hb%theta_dot = theta_dot
    call integrate_position(t, hb)
    x = hb%x; y = hb%y; vx = hb%vx; vy = hb%vy
  end subroutine

end module ref_harness
