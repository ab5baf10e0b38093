! Reads a runtime_backup file with the record list of init_fields
! (inicond = 2, init_fields.f90:55-94) and prints what it read, so that a
! test can hold the library's writer to a Fortran reader.
!   read_backup FILE NX NY NS
program read_backup
  implicit none
  integer, parameter :: pr = kind(0.d0)
  character(len=512) :: path, arg
  integer :: nx, ny, ns, n1, ivideo, ns_file, imv
  real(kind=pr) :: time, checksum, dt1, colorscale, E_kin, E_pot, E_ela, x0, y0, angle, dun, duo, lun, luo, dt_old
  logical :: colorscale_done, ucr
  real(kind=pr), allocatable :: vortk(:,:,:), nlk(:,:,:), workvis(:,:,:), u(:,:,:)
  real(kind=pr), allocatable :: mask(:,:), maskvx(:,:), maskvy(:,:), vor_mean(:,:)
  real(kind=pr), allocatable :: x(:), y(:), vx(:), vy(:), theta(:), theta_dot(:)
  real(kind=pr), allocatable :: p_new(:), p_old(:), t_new(:), t_old(:), oldold(:,:), ax(:), ay(:)
  real(kind=pr) :: Force(2), Force_unst(2), Force_press(2), Inertial_Force(2)

  call get_command_argument(1, path)
  call get_command_argument(2, arg); read (arg, *) nx
  call get_command_argument(3, arg); read (arg, *) ny
  call get_command_argument(4, arg); read (arg, *) ns
  allocate (vortk(0:nx-1,0:ny-1,0:1), nlk(0:nx-1,0:ny-1,0:1), workvis(0:nx-1,0:ny-1,0:1), u(0:nx-1,0:ny-1,0:1))
  allocate (mask(0:nx-1,0:ny-1), maskvx(0:nx-1,0:ny-1), maskvy(0:nx-1,0:ny-1), vor_mean(0:nx-1,0:ny-1))
  allocate (x(0:ns-1), y(0:ns-1), vx(0:ns-1), vy(0:ns-1), theta(0:ns-1), theta_dot(0:ns-1))
  allocate (p_new(0:ns-1), p_old(0:ns-1), t_new(0:ns-1), t_old(0:ns-1), oldold(0:ns-1,1:6), ax(0:ns-1), ay(0:ns-1))

  open (15, file=trim(path), form='unformatted', status='old')
  read (15) time
  read (15) checksum
  read (15) n1, dt1, vortk, nlk, workvis, mask, maskvx, maskvy, ivideo, u
  read (15) ns_file
  read (15) colorscale, colorscale_done
  read (15) x, y, vx, vy, theta, theta_dot
  read (15) p_new, p_old, t_new, t_old
  read (15) Force, Force_unst, Force_press, E_kin, E_pot, E_ela, x0, y0
  read (15) angle, imv, dun, duo, lun, luo
  read (15) ucr, dt_old, oldold
  read (15) ax, ay, Inertial_Force
  read (15) vor_mean
  close (15)

  write (*, '(A,ES25.17)') 'time ', time
  write (*, '(A,ES25.17)') 'checksum ', checksum
  write (*, '(A,I0)') 'n1 ', n1
  write (*, '(A,ES25.17)') 'dt1 ', dt1
  write (*, '(A,I0)') 'ivideo ', ivideo
  write (*, '(A,I0)') 'ns ', ns_file
  write (*, '(A,ES25.17)') 'colorscale ', colorscale
  write (*, '(A,L1)') 'colorscale_done ', colorscale_done
  write (*, '(A,2ES25.17)') 'vortk ', sum(vortk(:,:,0)), sum(vortk(:,:,1))
  write (*, '(A,2ES25.17)') 'nlk ', sum(nlk(:,:,0)), sum(nlk(:,:,1))
  write (*, '(A,2ES25.17)') 'workvis ', sum(workvis(:,:,0)), sum(workvis(:,:,1))
  write (*, '(A,3ES25.17)') 'mask ', sum(mask), sum(maskvx), sum(maskvy)
  write (*, '(A,2ES25.17)') 'u ', sum(u(:,:,0)), sum(u(:,:,1))
  write (*, '(A,2ES25.17)') 'beam_xy ', sum(x), sum(y)
  write (*, '(A,2ES25.17)') 'beam_theta ', sum(theta), sum(theta_dot)
  write (*, '(A,2ES25.17)') 'loads ', sum(p_new), sum(t_old)
  write (*, '(A,6ES25.17)') 'forces ', Force, Force_unst, Force_press
  write (*, '(A,2ES25.17)') 'origin ', x0, y0
  write (*, '(A,I0)') 'iMouvement ', imv
  write (*, '(A,2ES25.17)') 'unst ', duo, luo
  write (*, '(A,L1)') 'UnsteadyCorrectionsReady ', ucr
  write (*, '(A,ES25.17)') 'oldold ', sum(oldold)
  write (*, '(A,ES25.17)') 'vor_mean ', sum(vor_mean)
end program read_backup
