! Writes a field file with SaveField's records (FieldExport.f90:626-644)
! and WriteHeaderToFile's CHARACTER(256) header lines, for values read
! from a raw little-endian REAL(8) file, so that a test can compare the
! library's writer with the Fortran runtime's formatting byte for byte.
!   write_field VALUES NX NY NSAVE FIELDTYPE OUT
program write_field
  implicit none
  integer, parameter :: pr = kind(0.d0)
  character(len=512) :: vpath, opath, arg, ftype
  character(len=256) :: hdr(2)
  integer :: nx1, ny1, nsave, ix, iy
  real(kind=pr), allocatable :: field(:,:)
  real(kind=pr) :: lx, ly

  call get_command_argument(1, vpath)
  call get_command_argument(2, arg); read (arg, *) nx1
  call get_command_argument(3, arg); read (arg, *) ny1
  call get_command_argument(4, arg); read (arg, *) nsave
  call get_command_argument(5, ftype)
  call get_command_argument(6, opath)
  allocate (field(nx1, ny1))
  open (11, file=trim(vpath), access='stream', form='unformatted', status='old')
  read (11) field
  close (11)
  lx = 4.0_pr
  ly = 2.5_pr
  hdr(1) = '% nx=64 ny=64'
  hdr(2) = '% Re=1000'

  open (10, file=trim(opath), form='formatted', status='replace')
  do ix = 1, 2
    write (10, '(A)') hdr(ix)
  end do
  write (10,'(A)') "%format1"
  write (10,'(i5.5,3x,A)') nx1, "% nx"
  write (10,'(i5.5,3x,A)') ny1, "% ny"
  write (10,'(es11.4,3x,A)') lx, "% xl, domain size in x-direction"
  write (10,'(es11.4,3x,A)') ly, "% yl, domain size in y-direction"
  write (10,'(A)') "% The next entry can be used to identify the type of the field: "
  write (10,'(A)') "% "//trim(ftype)
  if (trim(ftype) == "precision") then
    do iy = 0, ny1-1, nsave
      write (10, '(es16.9)') (field(ix, ny1-iy), ix=1,nx1, nsave)
    end do
  else
    do iy = 0, ny1-1, nsave
      write (10, '(es11.4)') (field(ix, ny1-iy), ix=1,nx1, nsave)
    end do
  endif
  close (10)
end program write_field
