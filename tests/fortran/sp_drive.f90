! Fortran host for the C ABI: the calling sequence SIMCOFlow's UpdatePUVW
! would use (ComputePUVW.f90:127-129 replaced by one sp_project call), on
! arrays with the host's own bounds.
!
!   sp_drive <dir>
! reads <dir>/meta.txt and raw little-endian arrays written by the test,
! runs one projection and writes p, u, v, w, pocoef and stats back to <dir>.
program sp_drive
  use iso_c_binding
  use simco_proj
  implicit none
  character(len=512) :: dir
  integer :: Imax, Jmax, Kmax, pin(3), prec, bcf(6)
  real(c_double) :: dt, rho_l, rho_g
  real(c_double), allocatable, target :: ustar(:,:,:), vstar(:,:,:), wstar(:,:,:)
  real(c_double), allocatable, target :: dpu(:,:,:), dpv(:,:,:), dpw(:,:,:)
  real(c_double), allocatable, target :: eea(:,:,:), nea(:,:,:), tea(:,:,:)
  real(c_double), allocatable, target :: vof(:,:,:), vofo(:,:,:), vofl(:,:,:), phil(:,:,:)
  real(c_double), allocatable, target :: p(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:)
  real(c_double), allocatable, target :: pocoef(:,:,:,:)
  integer(c_int32_t), allocatable, target :: ctp(:,:,:), ctu(:,:,:), ctv(:,:,:), ctw(:,:,:)
  integer(c_int32_t), allocatable, target :: mu(:,:,:), mv(:,:,:), mw(:,:,:)
  real(c_double), allocatable :: dxp(:), dyp(:), dzp(:), dxu(:), dyv(:), dzw(:)
  type(sp_params) :: prm
  type(sp_inputs) :: inp
  type(sp_outputs) :: outp
  type(sp_stats) :: st
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  integer :: f

  call get_command_argument(1, dir)
  open(newunit=f, file=trim(dir)//'/meta.txt', status='old', action='read')
  read(f, *) Imax, Jmax, Kmax
  read(f, *) dt, rho_l, rho_g
  read(f, *) pin, prec
  read(f, *) bcf
  close(f)

  allocate(ustar(0:Imax+1, 0:Jmax+1, 0:Kmax+1))
  allocate(vstar, wstar, dpu, dpv, dpw, eea, nea, tea, vof, vofo, vofl, phil, p, u, v, w, &
           mold=ustar)
  allocate(ctp(Imax, Jmax, Kmax))
  allocate(ctu, ctv, ctw, mu, mv, mw, mold=ctp)
  allocate(pocoef(Imax, Jmax, Kmax, 6))
  allocate(dxp(Imax), dyp(Jmax), dzp(Kmax), dxu(Imax), dyv(Jmax), dzw(Kmax))

  call rd(ustar, 'u_star'); call rd(vstar, 'v_star'); call rd(wstar, 'w_star')
  call rd(dpu, 'dp_u'); call rd(dpv, 'dp_v'); call rd(dpw, 'dp_w')
  call rd(eea, 'ee_area'); call rd(nea, 'ne_area'); call rd(tea, 'te_area')
  call rd(vof, 'vof'); call rd(vofo, 'vof_old'); call rd(vofl, 'vof_l'); call rd(phil, 'phi_l')
  call rdi(ctp, 'cell_type_p'); call rdi(ctu, 'cell_type_u'); call rdi(ctv, 'cell_type_v')
  call rdi(ctw, 'cell_type_w'); call rdi(mu, 'moex_u'); call rdi(mv, 'moex_v')
  call rdi(mw, 'moex_w')
  open(newunit=f, file=trim(dir)//'/grid.bin', access='stream', form='unformatted', &
       status='old', action='read')
  read(f) dxp, dyp, dzp, dxu, dyv, dzw
  close(f)
  u = 0.0_c_double; v = 0.0_c_double; w = 0.0_c_double; p = 0.0_c_double

  call sp_params_default(prm)
  prm%imax = Imax; prm%jmax = Jmax; prm%kmax = Kmax
  prm%precision = prec
  prm%rho_l = rho_l; prm%rho_g = rho_g; prm%rho_ref = rho_l
  prm%pin_i = pin(1); prm%pin_j = pin(2); prm%pin_k = pin(3)
  rc = sp_create(ctx, prm, 0_c_int)
  call check(rc, 'sp_create')
  rc = sp_set_grid_1d(ctx, dxp, dyp, dzp, dxu, dyv, dzw)
  call check(rc, 'sp_set_grid_1d')

  inp%u_star = c_loc(ustar); inp%v_star = c_loc(vstar); inp%w_star = c_loc(wstar)
  inp%dp_u = c_loc(dpu); inp%dp_v = c_loc(dpv); inp%dp_w = c_loc(dpw)
  inp%ee_area = c_loc(eea); inp%ne_area = c_loc(nea); inp%te_area = c_loc(tea)
  inp%vof = c_loc(vof); inp%vof_old = c_loc(vofo)
  inp%vof_l = c_loc(vofl); inp%phi_l = c_loc(phil)
  inp%cell_type_p = c_loc(ctp)
  inp%cell_type_u = c_loc(ctu); inp%cell_type_v = c_loc(ctv); inp%cell_type_w = c_loc(ctw)
  inp%moex_u = c_loc(mu); inp%moex_v = c_loc(mv); inp%moex_w = c_loc(mw)
  inp%bc_flag = bcf
  inp%dt = dt
  outp%p = c_loc(p); outp%u = c_loc(u); outp%v = c_loc(v); outp%w = c_loc(w)
  outp%po_coef = c_loc(pocoef)

  rc = sp_project(ctx, inp, outp, st)
  call check(rc, 'sp_project')

  call wr(p, 'out_p'); call wr(u, 'out_u'); call wr(v, 'out_v'); call wr(w, 'out_w')
  open(newunit=f, file=trim(dir)//'/out_pocoef.bin', access='stream', form='unformatted', &
       status='replace', action='write')
  write(f) pocoef
  close(f)
  open(newunit=f, file=trim(dir)//'/stats.txt', status='replace', action='write')
  write(f, '(i0,1x,i0,1x,i0,1x,es24.16)') st%iterations, st%converged, st%levels, st%final_div
  close(f)
  rc = sp_destroy(ctx)
  call check(rc, 'sp_destroy')
  print '(a,i0,a,es10.3)', 'sp_drive: iterations ', st%iterations, ' final_div ', st%final_div

contains

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= SP_OK) then
      print '(a,a,a,a,a,a)', what, ': ', sp_strerror_f(status), ' (', sp_last_error_f(ctx), ')'
      stop 1
    end if
  end subroutine check

  subroutine rd(a, name)
    real(c_double), intent(out) :: a(:,:,:)
    character(len=*), intent(in) :: name
    integer :: g
    open(newunit=g, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(g) a
    close(g)
  end subroutine rd

  subroutine rdi(a, name)
    integer(c_int32_t), intent(out) :: a(:,:,:)
    character(len=*), intent(in) :: name
    integer :: g
    open(newunit=g, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(g) a
    close(g)
  end subroutine rdi

  subroutine wr(a, name)
    real(c_double), intent(in) :: a(:,:,:)
    character(len=*), intent(in) :: name
    integer :: g
    open(newunit=g, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    write(g) a
    close(g)
  end subroutine wr

end program sp_drive
