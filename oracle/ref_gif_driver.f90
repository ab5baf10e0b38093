! Test infrastructure only (parity checker): runs the reference's SaveGIF
! (FieldExport.f90:305-384, compiled where it lies together with
! gif_util.f90 and share_vars.f90 by oracle/ref.mk) on a field read from a
! raw file, so that tests can compare the GIF it writes byte for byte.
!
! usage: save_gif_ref <input> <stem>   (writes <stem>.gif)
!   input, native stream: int32 nx, ny, palette, have_range;
!   float64 color_min, color_max; float64 field(nx, ny) (x fastest)
!   palette = 0 and have_range = 0 call SaveGIF without the optional
!   arguments (palette 13, the field's own range)
program save_gif_ref
  use fieldexport
  implicit none
  character(len=1024) :: inp, stem
  integer(4) :: nx, ny, ipal, hr
  integer :: u
  real(8) :: cmin, cmax
  real(8), allocatable :: f(:, :)

  call get_command_argument(1, inp)
  call get_command_argument(2, stem)
  open(newunit=u, file=trim(inp), access='stream', form='unformatted', status='old')
  read(u) nx, ny, ipal, hr
  read(u) cmin, cmax
  allocate(f(nx, ny))
  read(u) f
  close(u)
  if (ipal == 0 .and. hr == 0) then
    call SaveGIF(f, trim(stem))
  else if (hr == 0) then
    call SaveGIF(f, trim(stem), ipal)
  else
    call SaveGIF(f, trim(stem), ipal, cmin, cmax)
  end if
end program save_gif_ref
