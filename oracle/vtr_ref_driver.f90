! Test infrastructure only: a host program that calls the reference's own
! VTR writer module (src/VTR_mod.f90, compiled by oracle/Makefile from the
! reference's sources) in the order PrintResultVTR3D uses
! (PrintResult.f90:131-146), on fields read from raw files.  Its output is the
! golden file the engine's VTR writer is compared with.
!
! usage: vtr_ref_driver <dir with meta.txt and *.bin> <output dir ending in />
program vtr_ref_driver
  use PrecisionVar
  use vtr
  implicit none
  character(len=512) :: din, dout
  integer :: ni, nj, nk
  integer(kind=it8b) :: itt
  real(kind=dp), allocatable :: x(:), y(:), z(:)
  real(kind=dp), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:)
  real(kind=dp), allocatable :: phi(:,:,:), phil(:,:,:), vof(:,:,:), vofl(:,:,:), mres(:,:,:)
  type(VTR_file_handle) :: fd
  call get_command_argument(1, din)
  call get_command_argument(2, dout)
  open(unit=20, file=trim(din)//'/meta.txt', status='old', action='read')
  read(20, *) ni, nj, nk, itt
  close(20)
  allocate(x(ni), y(nj), z(nk))
  allocate(u(ni,nj,nk), v(ni,nj,nk), w(ni,nj,nk), p(ni,nj,nk), phi(ni,nj,nk), phil(ni,nj,nk))
  allocate(vof(ni,nj,nk), vofl(ni,nj,nk), mres(ni,nj,nk))
  call rd1(x, 'x'); call rd1(y, 'y'); call rd1(z, 'z')
  call rd3(u, 'u'); call rd3(v, 'v'); call rd3(w, 'w'); call rd3(p, 'p')
  call rd3(phi, 'phi'); call rd3(phil, 'phi_l'); call rd3(vof, 'vof'); call rd3(vofl, 'vof_l')
  call rd3(mres, 'mres')
  call VTR_open_file(Prefix='case', dir=trim(dout), itera=itt, FD=fd)
  call VTR_write_mesh(fd, x, y, z)
  call VTR_write_var(fd, "Velocity", u, v, w)
  call VTR_write_var(fd, "Pressure", p)
  call VTR_write_var(fd, "FluidLvs", phi)
  call VTR_write_var(fd, "LiquidLvs", phil)
  call VTR_write_var(fd, "FluidVof", vof)
  call VTR_write_var(fd, "LiquidVof", vofl)
  call VTR_write_var(fd, "Mass error", mres)
  call VTR_close_file(fd)
contains
  subroutine rd1(a, name)
    real(kind=dp), intent(out) :: a(:)
    character(len=*), intent(in) :: name
    open(unit=21, file=trim(din)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(21) a
    close(21)
  end subroutine rd1
  subroutine rd3(a, name)
    real(kind=dp), intent(out) :: a(:,:,:)
    character(len=*), intent(in) :: name
    open(unit=21, file=trim(din)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(21) a
    close(21)
  end subroutine rd3
end program vtr_ref_driver
