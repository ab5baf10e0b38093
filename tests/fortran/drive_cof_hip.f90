! drive_cof_hip.f90 -- a Fortran host that uses the reference's own module
! share_vars (compiled from the reference by oracle/ref.mk) and calls the FFT
! plugin and operators through fsibe_2d_amd/fortran/cof_hip.f90 and
! ops_hip.f90 with the reference's call signatures.  Writes every input and
! output to the stream file named on the command line; tests/test_fortran_binding.py
! compares them with the oracle.
program drive_cof_hip
  use share_vars
  implicit none
  real(kind=pr), allocatable :: f(:,:), fk(:,:), g(:,:), w(:,:), t(:,:), tk(:,:), ti(:,:)
  character(len=512) :: path
  integer :: ix, iy, u, L, n

  call get_command_argument(1, path)
  nx = 96; ny = 64; xl = 4.0d0; yl = 3.0d0; nu = 1.3d-2
  pi = 4.d0 * atan(1.d0)
  scalex = (2.d0*pi/xl)**2
  scaley = (2.d0*pi/yl)**2
  allocate(dealiase(0:nx-1, 0:ny-1))
  allocate(f(0:nx-1, 0:ny-1), fk(0:nx-1, 0:ny-1), g(0:nx-1, 0:ny-1), w(0:nx-1, 0:ny-1))
  do iy = 0, ny-1
    do ix = 0, nx-1
      f(ix, iy) = sin(0.37d0*ix + 1.3d0*iy) + 0.5d0*cos(0.011d0*ix*iy) + 0.25d0*real(mod(7*ix+3*iy, 11), pr)
    end do
  end do

  call fft_initialize
  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
  write(u) nx, ny
  write(u) f
  call coftxy(f, fk);   write(u) fk
  call cofitxy(fk, g);  write(u) g
  call cofdx(fk, w);    write(u) w
  call cofdy(fk, w);    write(u) w
  call cofdxdx(fk, w);  write(u) w
  call cofdydy(fk, w);  write(u) w
  call cofdxdy(fk, w);  write(u) w
  call poisson(fk, w);  write(u) w
  call vis(0.37d0, w);  write(u) w
  dealiase = -1.d0
  call dealiase_mask;   write(u) dealiase
  L = 37; n = 2
  allocate(t(0:L-1, 0:n-1), tk(0:L-1, 0:n-1), ti(0:L-1, 0:n-1))
  do iy = 0, n-1
    do ix = 0, L-1
      t(ix, iy) = cos(0.2d0*ix*(iy+1)) + 0.1d0*ix
    end do
  end do
  call cofts(t, tk, L, n)
  call cofits(tk, ti, L, n)
  write(u) L, n
  write(u) t, tk, ti
  close(u)
  call fft_free
end program drive_cof_hip
