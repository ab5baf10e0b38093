! cof_hip.f90 -- drop-in replacement of the reference's FFT plugin file
! (cof_fftw33.f90, chosen by COF_FILE in makefile:40,75): the same external
! subroutines with the same dummies, forwarding to the HIP library through
! fsibe_hip_api.  Sizes come from share_vars (nx, ny, xl, yl, nu) exactly as in
! the file it replaces; the context handle takes the place of the FFTW plans
! (module fftw33_descriptors).
module fsibe_cof_state
  use iso_c_binding
  implicit none
  type(c_ptr), save :: fsibe_ctx = c_null_ptr
end module fsibe_cof_state

! fft_initialize, cof_fftw33.f90:18-66
subroutine fft_initialize
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  call fsibe_check(fsibe_spectral_create(int(nx, c_int), int(ny, c_int), real(xl, c_double), &
                                         real(yl, c_double), real(nu, c_double), fsibe_ctx), &
                   "fft_initialize")
end subroutine fft_initialize

! fft_free, cof_fftw33.f90:119-137
subroutine fft_free
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  call fsibe_check(fsibe_spectral_destroy(fsibe_ctx), "fft_free")
  fsibe_ctx = c_null_ptr
end subroutine fft_free

! coftxy(f, fk), cof_fftw33.f90:142-193
subroutine coftxy(f, fk)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: f
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: fk
  call fsibe_check(fsibe_coftxy(fsibe_ctx, f, fk), "coftxy")
end subroutine coftxy

! cofitxy(fk, f), cof_fftw33.f90:197-243
subroutine cofitxy(fk, f)
  use share_vars
  use fsibe_hip_api
  use fsibe_cof_state
  implicit none
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(in) :: fk
  real(kind=pr), dimension(0:nx-1, 0:ny-1), intent(out) :: f
  call fsibe_check(fsibe_cofitxy(fsibe_ctx, fk, f), "cofitxy")
end subroutine cofitxy

! cofts(f, fk, L, n), cof_fftw33.f90:248-296 (host-side in the library)
subroutine cofts(f, fk, L, n)
  use share_vars
  use fsibe_hip_api
  implicit none
  integer, intent(in) :: L, n
  real(kind=pr), dimension(0:L-1, 0:n-1), intent(in) :: f
  real(kind=pr), dimension(0:L-1, 0:n-1), intent(out) :: fk
  call fsibe_check(fsibe_cofts(f, fk, int(L, c_int), int(n, c_int)), "cofts")
end subroutine cofts

! cofits(fk, f, L, n), cof_fftw33.f90:300-340
subroutine cofits(fk, f, L, n)
  use share_vars
  use fsibe_hip_api
  implicit none
  integer, intent(in) :: L, n
  real(kind=pr), dimension(0:L-1, 0:n-1), intent(in) :: fk
  real(kind=pr), dimension(0:L-1, 0:n-1), intent(out) :: f
  call fsibe_check(fsibe_cofits(fk, f, int(L, c_int), int(n, c_int)), "cofits")
end subroutine cofits
