! ops_hip.f90 -- replacements of the reference's spectral operator files
! (cofdx.f90, cofdy.f90, cofdxdx.f90, cofdydy.f90, cofdxdy.f90, poisson.f90,
! cal_vis.f90, dealiase_mask.f90), same external names and dummies, computed by
! the HIP library on the context that cof_hip.f90's fft_initialize created.
subroutine cofdx(fk, fk_dx)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: fk
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: fk_dx
  call fsibe_check(fsibe_cofdx(fsibe_ctx, fk, fk_dx), "cofdx")
end subroutine cofdx

subroutine cofdy(fk, fk_dy)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: fk
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: fk_dy
  call fsibe_check(fsibe_cofdy(fsibe_ctx, fk, fk_dy), "cofdy")
end subroutine cofdy

subroutine cofdxdx(fk, fk_dxdx)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: fk
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: fk_dxdx
  call fsibe_check(fsibe_cofdxdx(fsibe_ctx, fk, fk_dxdx), "cofdxdx")
end subroutine cofdxdx

subroutine cofdydy(fk, fk_dydy)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: fk
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: fk_dydy
  call fsibe_check(fsibe_cofdydy(fsibe_ctx, fk, fk_dydy), "cofdydy")
end subroutine cofdydy

subroutine cofdxdy(fk, fk_dxdy)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: fk
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: fk_dxdy
  call fsibe_check(fsibe_cofdxdy(fsibe_ctx, fk, fk_dxdy), "cofdxdy")
end subroutine cofdxdy

subroutine poisson(f, ans)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: f
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: ans
  call fsibe_check(fsibe_poisson(fsibe_ctx, f, ans), "poisson")
end subroutine poisson

subroutine vis(dt, work)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), intent(in) :: dt
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: work
  call fsibe_check(fsibe_vis(fsibe_ctx, real(dt, c_double), work), "vis")
end subroutine vis

! dealiase_mask fills the share_vars array dealiase (dealiase_mask.f90:1-33)
subroutine dealiase_mask
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  call fsibe_check(fsibe_dealiase_mask(fsibe_ctx, dealiase), "dealiase_mask")
end subroutine dealiase_mask
