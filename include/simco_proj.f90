! simco_proj.f90 -- Fortran 2003 (ISO_C_BINDING) interface to the C ABI in
! simco_proj.h, for SIMCOFlow's Fortran host.
!
! The derived types are bind(C) twins of the C structs (same members, same
! order; the companion C compiler's layout rules apply to both).  Array
! arguments are passed by address with the host's own bounds:
!   halo arrays  real(dp)   (0:Imax+1, 0:Jmax+1, 0:Kmax+1)
!   cell arrays  integer(it4b) (1:Imax, 1:Jmax, 1:Kmax)
!   PoCoef       real(dp)   (1:Imax, 1:Jmax, 1:Kmax, 6)
! and stored in the type(c_ptr) members with c_loc (the actual arrays need
! the TARGET attribute).  real(dp) = real(c_double), integer(it4b) =
! integer(c_int32_t) (Precision.f90:4,7).
module simco_proj
  use iso_c_binding
  implicit none
  private
  public :: sp_params, sp_inputs, sp_outputs, sp_stats, sp_adj_inputs, sp_adj_outputs, &
            sp_wall_bc, sp_vtr_fields, sp_vtk_fields, sp_solid_body, sp_lvs_grid, sp_clsvof_inputs, sp_clsvof_state, sp_cell_geom, sp_cell_interp, sp_liquid_field, &
            sp_pred_inputs, sp_pred_outputs
  public :: sp_params_default, sp_create, sp_destroy, sp_set_params, sp_set_grid, &
            sp_set_grid_1d, sp_project, sp_project_dev, sp_adjoint, sp_adjoint_dev, &
            sp_wall_step, sp_wall_step_dev, sp_write_vtr, sp_write_vtk, sp_interpolate_lvs, sp_clsvof_scheme, sp_grid_preprocess, &
            sp_grid_preprocess_dev, sp_new_cell_face, sp_define_momentum_exchange_cell, &
            sp_compute_uvw_liquid_field, sp_predict_uvw, sp_number_external_cell_noalloc, &
            sp_number_external_cell, sp_strerror_f, sp_last_error_f
  public :: SP_OK, SP_ERR_INVALID_ARG, SP_ERR_HIP, SP_ERR_NOT_CONVERGED, SP_ERR_NONFINITE, &
            SP_ERR_OOM, SP_ERR_GRID, SP_ERR_STATE, SP_ERR_NO_DEVICE, SP_ERR_CAPACITY, &
            SP_PREC_FP64, &
            SP_PREC_MIXED, SP_PRECOND_MG, SP_PRECOND_JACOBI

  integer(c_int), parameter :: SP_OK = 0, SP_ERR_INVALID_ARG = -1, SP_ERR_HIP = -2, &
                               SP_ERR_NOT_CONVERGED = -3, SP_ERR_NONFINITE = -4, &
                               SP_ERR_OOM = -5, SP_ERR_GRID = -6, SP_ERR_STATE = -7, &
                               SP_ERR_NO_DEVICE = -8, SP_ERR_CAPACITY = -9
  integer(c_int32_t), parameter :: SP_PREC_FP64 = 0, SP_PREC_MIXED = 1
  integer(c_int32_t), parameter :: SP_PRECOND_MG = 0, SP_PRECOND_JACOBI = 1

  type, bind(C) :: sp_params
    integer(c_int32_t) :: imax, jmax, kmax
    integer(c_int32_t) :: precision, precond
    integer(c_int32_t) :: mg_levels, mg_pre, mg_post, mg_coarse_sweeps
    real(c_double) :: mg_coarse_scale
    integer(c_int32_t) :: max_iter
    real(c_double) :: div_tol
    real(c_double) :: rho_l, rho_g, rho_ref
    real(c_double) :: epsi
    integer(c_int32_t) :: pin_i, pin_j, pin_k
    integer(c_int32_t) :: flags
    real(c_double) :: mg_omega, mg_omega_coarse
  end type sp_params

  ! BCu/BCv/BCw%flag and BCu/BCv/BCw/BCp%const for the wall step after the
  ! correction (BoundaryConditionVarNew, the Dirichlet-pressure mass fix and
  ! VariablesInternalCellCondition); the pressure flags are sp_inputs%bc_flag.
  type, bind(C) :: sp_wall_bc
    integer(c_int32_t) :: flag_u(6) = 1, flag_v(6) = 1, flag_w(6) = 1
    real(c_double) :: const_u(6) = 0.0_c_double, const_v(6) = 0.0_c_double
    real(c_double) :: const_w(6) = 0.0_c_double, const_p(6) = 0.0_c_double
  end type sp_wall_bc

  type, bind(C) :: sp_inputs
    type(c_ptr) :: u_star = c_null_ptr, v_star = c_null_ptr, w_star = c_null_ptr
    type(c_ptr) :: dp_u = c_null_ptr, dp_v = c_null_ptr, dp_w = c_null_ptr
    type(c_ptr) :: ee_area = c_null_ptr, ne_area = c_null_ptr, te_area = c_null_ptr
    type(c_ptr) :: vof = c_null_ptr, vof_old = c_null_ptr
    type(c_ptr) :: vof_l = c_null_ptr, phi_l = c_null_ptr
    type(c_ptr) :: cell_type_p = c_null_ptr
    type(c_ptr) :: cell_type_u = c_null_ptr, cell_type_v = c_null_ptr, cell_type_w = c_null_ptr
    type(c_ptr) :: moex_u = c_null_ptr, moex_v = c_null_ptr, moex_w = c_null_ptr
    integer(c_int32_t) :: bc_flag(6) = 1
    type(c_ptr) :: bc_val(6) = c_null_ptr
    real(c_double) :: dt = 0.0_c_double
    type(c_ptr) :: beta_e = c_null_ptr, beta_n = c_null_ptr, beta_t = c_null_ptr
    type(c_ptr) :: wall = c_null_ptr      ! c_loc of a type(sp_wall_bc), or null
    ! 1: geometry arrays unchanged since the previous sp_project call (no re-copy)
    integer(c_int32_t) :: static_geometry = 0
  end type sp_inputs

  type, bind(C) :: sp_outputs
    type(c_ptr) :: p = c_null_ptr
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr
    type(c_ptr) :: po_coef = c_null_ptr
    ! BCu%VarW..VarT, BCv, BCw, BCp after the wall step (null entries skipped)
    type(c_ptr) :: bc_var(24) = c_null_ptr
  end type sp_outputs

  type, bind(C) :: sp_stats
    integer(c_int32_t) :: iterations, converged, levels, singular
    integer(c_int64_t) :: active_cells, stale_faces
    real(c_double) :: initial_div, final_div, residual_div
    real(c_double) :: t_assemble_ms, t_solve_ms, t_correct_ms, t_total_ms
    real(c_double) :: wall_div
  end type sp_stats

  ! PrintResultVTR3D's fields (c_loc of the host arrays); bit q of halo_mask
  ! marks field q (u, v, w, p, phi, phi_l, vof, vof_l, mres) as a
  ! (0:Imax+1, ...) array, otherwise (1:Imax, ...)
  type, bind(C) :: sp_vtr_fields
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr, p = c_null_ptr
    type(c_ptr) :: phi = c_null_ptr, phi_l = c_null_ptr
    type(c_ptr) :: vof = c_null_ptr, vof_l = c_null_ptr
    type(c_ptr) :: mres = c_null_ptr
    integer(c_int32_t) :: halo_mask = 0
  end type sp_vtr_fields

  ! PrintResultVTK's fields (c_loc of the host arrays): cell centres x, y, z
  ! and widths dx, dy, dz (TGrid%x(:,1,1), TGrid%dx(:,1,1), ...); bit q of
  ! halo_mask marks field q (p, u, v, w, phi_l, phi, vof_l, vof) as a
  ! (0:Imax+1, ...) array, otherwise (1:Imax, ...)
  type, bind(C) :: sp_vtk_fields
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: dx = c_null_ptr, dy = c_null_ptr, dz = c_null_ptr
    type(c_ptr) :: p = c_null_ptr, u = c_null_ptr, v = c_null_ptr, w = c_null_ptr
    type(c_ptr) :: phi_l = c_null_ptr, phi = c_null_ptr
    type(c_ptr) :: vof_l = c_null_ptr, vof = c_null_ptr
    integer(c_int32_t) :: halo_mask = 0
  end type sp_vtk_fields

  ! A solid body of SolidBody.f90:25-54: local mesh size, uniform spacings,
  ! c_loc of this%x, y, z and this%phi, this%CMpoint, this%rotMatLtoG
  type, bind(C) :: sp_solid_body
    integer(c_int32_t) :: imax = 0, jmax = 0, kmax = 0
    real(c_double) :: dx = 0, dy = 0, dz = 0
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: phi = c_null_ptr
    real(c_double) :: cm(3) = 0
    real(c_double) :: rot(3,3) = 0
  end type sp_solid_body

  ! One grid of interpolateLvS: c_loc of TGrid%x, y, z, dx, dy, dz and of
  ! TCell%phi, vof, nx, ny, nz
  type, bind(C) :: sp_lvs_grid
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: dx = c_null_ptr, dy = c_null_ptr, dz = c_null_ptr
    type(c_ptr) :: phi = c_null_ptr, vof = c_null_ptr
    type(c_ptr) :: nx = c_null_ptr, ny = c_null_ptr, nz = c_null_ptr
  end type sp_lvs_grid

  ! Clsvof_Scheme inputs: c_loc of TVar%u, v, w, PGrid%x..dz, PCell%nx, ny,
  ! nz, BCu%VarW, BCv%VarS, BCw%VarB; BCLvs and BCVof %flag and %const; dt
  type, bind(C) :: sp_clsvof_inputs
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: dx = c_null_ptr, dy = c_null_ptr, dz = c_null_ptr
    type(c_ptr) :: nx = c_null_ptr, ny = c_null_ptr, nz = c_null_ptr
    type(c_ptr) :: bcu_w = c_null_ptr, bcv_s = c_null_ptr, bcw_b = c_null_ptr
    integer(c_int32_t) :: lvs_flag(6) = 1
    real(c_double) :: lvs_const(6) = 0
    integer(c_int32_t) :: vof_flag(6) = 1
    real(c_double) :: vof_const(6) = 0
    real(c_double) :: dt = 0
  end type sp_clsvof_inputs

  ! Clsvof_Scheme state: c_loc of PCell%vofL, phiL, nxL, nyL, nzL and of
  ! BCLvs%VarW..VarT, BCVof%VarW..VarT
  type, bind(C) :: sp_clsvof_state
    type(c_ptr) :: vof_l = c_null_ptr, phi_l = c_null_ptr
    type(c_ptr) :: nx_l = c_null_ptr, ny_l = c_null_ptr, nz_l = c_null_ptr
    type(c_ptr) :: lvs_var(6) = c_null_ptr
    type(c_ptr) :: vof_var(6) = c_null_ptr
  end type sp_clsvof_state

  ! One staggered grid of GridPreprocess (CutCell.f90:62-325): c_loc of
  ! TGrid%dx, dy, dz and of the TCell members (allocated as in AllocateVar,
  ! Main.f90:229-399).  fce/fcn/fct are TCell%FCE/FCN/FCT(0:Imax+1,0:Jmax+1,
  ! 0:Kmax+1,3), cell_cent is TCell%Cell_Cent(Imax,Jmax,Kmax,3).
  type, bind(C) :: sp_cell_geom
    type(c_ptr) :: dx = c_null_ptr, dy = c_null_ptr, dz = c_null_ptr
    type(c_ptr) :: phi = c_null_ptr, nx = c_null_ptr, ny = c_null_ptr, nz = c_null_ptr
    type(c_ptr) :: vof = c_null_ptr, vof_l = c_null_ptr, cell_type = c_null_ptr
    type(c_ptr) :: ee_area = c_null_ptr, ne_area = c_null_ptr, te_area = c_null_ptr
    type(c_ptr) :: fce = c_null_ptr, fcn = c_null_ptr, fct = c_null_ptr
    type(c_ptr) :: cell_cent = c_null_ptr, wllh = c_null_ptr, delh = c_null_ptr
    ! 1 when phi, nx, ny, nz, vof, vofL carry the halo (0:Imax+1,..)
    integer(c_int32_t) :: halo_fields = 0
  end type sp_cell_geom

  ! NewCellFace members of one U/V/W cell besides sp_cell_geom: c_loc of
  ! TCell%MoExCell, MsCe(Imax,Jmax,Kmax,3) and SxE, EtaE, AlE, DAlE, SyN,
  ! EtaN, AlN, DAlN, SzT, EtaT, AlT, DAlT (halo arrays)
  type, bind(C) :: sp_cell_interp
    type(c_ptr) :: moex = c_null_ptr, msce = c_null_ptr
    type(c_ptr) :: sxe = c_null_ptr, etae = c_null_ptr, ale = c_null_ptr, dale = c_null_ptr
    type(c_ptr) :: syn = c_null_ptr, etan = c_null_ptr, aln = c_null_ptr, daln = c_null_ptr
    type(c_ptr) :: szt = c_null_ptr, etat = c_null_ptr, alt = c_null_ptr, dalt = c_null_ptr
  end type sp_cell_interp

  ! ComputeUVWLiquidField fields of one grid: c_loc of TGrid%dx, dy, dz and
  ! of TCell%phiL, nxL, nyL, nzL, vofL
  type, bind(C) :: sp_liquid_field
    type(c_ptr) :: dx = c_null_ptr, dy = c_null_ptr, dz = c_null_ptr
    type(c_ptr) :: phi_l = c_null_ptr, nx_l = c_null_ptr, ny_l = c_null_ptr
    type(c_ptr) :: nz_l = c_null_ptr, vof_l = c_null_ptr
    ! 1 when phiL, nxL, nyL, nzL, vofL carry the halo (0:Imax+1,..)
    integer(c_int32_t) :: halo_fields = 0
  end type sp_liquid_field

  ! PredictingUVW inputs: c_loc of PCellO%EEArea.., of type(sp_cell_geom) /
  ! type(sp_cell_interp) variables describing UCellO.. and UCell.., of TVar
  ! and TVar_n velocities, FluxDivOld, the BCu/BCv/BCw face arrays
  type, bind(C) :: sp_pred_inputs
    type(c_ptr) :: p_ee_area = c_null_ptr, p_ne_area = c_null_ptr, p_te_area = c_null_ptr
    type(c_ptr) :: u_old = c_null_ptr, v_old = c_null_ptr, w_old = c_null_ptr
    type(c_ptr) :: u_old_x = c_null_ptr, v_old_x = c_null_ptr, w_old_x = c_null_ptr
    type(c_ptr) :: u_cur = c_null_ptr, v_cur = c_null_ptr, w_cur = c_null_ptr
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr
    type(c_ptr) :: u_n = c_null_ptr, v_n = c_null_ptr, w_n = c_null_ptr
    type(c_ptr) :: flux_div_old = c_null_ptr
    integer(c_int32_t) :: bc_flag(18) = 1
    type(c_ptr) :: bc_var(18) = c_null_ptr
    real(c_double) :: dt = 0.0_c_double
    integer(c_int64_t) :: itt = 1
    integer(c_int32_t) :: time_order = 1, space_order = 1
    real(c_double) :: mu_w_ref = 1.0_c_double, mu_a_ref = 1.0_c_double
    real(c_double) :: rey = 1.0_c_double, fr = 1.0_c_double
    real(c_double) :: gx = 0.0_c_double, gy = 0.0_c_double, gz = 0.0_c_double, g = 1.0_c_double
  end type sp_pred_inputs

  ! c_loc of Pred%u, v, w and PU%Dp, PV%Dp, PW%Dp
  type, bind(C) :: sp_pred_outputs
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr
    type(c_ptr) :: dp_u = c_null_ptr, dp_v = c_null_ptr, dp_w = c_null_ptr
  end type sp_pred_outputs

  type, bind(C) :: sp_adj_inputs
    type(c_ptr) :: gu = c_null_ptr, gv = c_null_ptr, gw = c_null_ptr
    type(c_ptr) :: gp = c_null_ptr  ! dL/dPp, or null
  end type sp_adj_inputs

  type, bind(C) :: sp_adj_outputs
    type(c_ptr) :: gu_star = c_null_ptr, gv_star = c_null_ptr, gw_star = c_null_ptr
    type(c_ptr) :: gk_e = c_null_ptr, gk_n = c_null_ptr, gk_t = c_null_ptr
    type(c_ptr) :: gbeta_e = c_null_ptr, gbeta_n = c_null_ptr, gbeta_t = c_null_ptr
  end type sp_adj_outputs

  interface
    subroutine sp_params_default(p) bind(C, name="sp_params_default")
      import :: sp_params
      type(sp_params), intent(out) :: p
    end subroutine sp_params_default

    integer(c_int) function sp_create(ctx, params, device) bind(C, name="sp_create")
      import :: c_ptr, c_int, sp_params
      type(c_ptr), intent(out) :: ctx
      type(sp_params), intent(in) :: params
      integer(c_int), value :: device
    end function sp_create

    integer(c_int) function sp_destroy(ctx) bind(C, name="sp_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function sp_destroy

    integer(c_int) function sp_set_params(ctx, params) bind(C, name="sp_set_params")
      import :: c_ptr, c_int, sp_params
      type(c_ptr), value :: ctx
      type(sp_params), intent(in) :: params
    end function sp_set_params

    ! PGrid%dx,dy,dz, UGrid%dx, VGrid%dy, WGrid%dz (Imax x Jmax x Kmax each)
    integer(c_int) function sp_set_grid(ctx, dxp, dyp, dzp, dxu, dyv, dzw) &
        bind(C, name="sp_set_grid")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: dxp(*), dyp(*), dzp(*), dxu(*), dyv(*), dzw(*)
    end function sp_set_grid

    integer(c_int) function sp_set_grid_1d(ctx, dxp, dyp, dzp, dxu, dyv, dzw) &
        bind(C, name="sp_set_grid_1d")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: dxp(*), dyp(*), dzp(*), dxu(*), dyv(*), dzw(*)
    end function sp_set_grid_1d

    integer(c_int) function sp_project(ctx, inp, outp, st) bind(C, name="sp_project")
      import :: c_ptr, c_int, sp_inputs, sp_outputs, sp_stats
      type(c_ptr), value :: ctx
      type(sp_inputs), intent(in) :: inp
      type(sp_outputs), intent(in) :: outp
      type(sp_stats), intent(out) :: st
    end function sp_project

    integer(c_int) function sp_project_dev(ctx, inp, outp, st) bind(C, name="sp_project_dev")
      import :: c_ptr, c_int, sp_inputs, sp_outputs, sp_stats
      type(c_ptr), value :: ctx
      type(sp_inputs), intent(in) :: inp
      type(sp_outputs), intent(in) :: outp
      type(sp_stats), intent(out) :: st
    end function sp_project_dev

    integer(c_int) function sp_adjoint(ctx, inp, outp, st) bind(C, name="sp_adjoint")
      import :: c_ptr, c_int, sp_adj_inputs, sp_adj_outputs, sp_stats
      type(c_ptr), value :: ctx
      type(sp_adj_inputs), intent(in) :: inp
      type(sp_adj_outputs), intent(in) :: outp
      type(sp_stats), intent(out) :: st
    end function sp_adjoint

    integer(c_int) function sp_adjoint_dev(ctx, inp, outp, st) bind(C, name="sp_adjoint_dev")
      import :: c_ptr, c_int, sp_adj_inputs, sp_adj_outputs, sp_stats
      type(c_ptr), value :: ctx
      type(sp_adj_inputs), intent(in) :: inp
      type(sp_adj_outputs), intent(in) :: outp
      type(sp_stats), intent(out) :: st
    end function sp_adjoint_dev

    integer(c_int) function sp_wall_step(ctx, inp, outp, st) bind(C, name="sp_wall_step")
      import :: c_ptr, c_int, sp_inputs, sp_outputs, sp_stats
      type(c_ptr), value :: ctx
      type(sp_inputs), intent(in) :: inp
      type(sp_outputs), intent(in) :: outp
      type(sp_stats), intent(out) :: st
    end function sp_wall_step

    integer(c_int) function sp_wall_step_dev(ctx, inp, outp, st) bind(C, name="sp_wall_step_dev")
      import :: c_ptr, c_int, sp_inputs, sp_outputs, sp_stats
      type(c_ptr), value :: ctx
      type(sp_inputs), intent(in) :: inp
      type(sp_outputs), intent(in) :: outp
      type(sp_stats), intent(out) :: st
    end function sp_wall_step_dev

    ! dir and prefix: NUL-terminated (trim(dir)//c_null_char)
    integer(c_int) function sp_write_vtr(ctx, dir, prefix, itt, fields, device) &
        bind(C, name="sp_write_vtr")
      import :: c_ptr, c_int, c_char, c_int64_t, sp_vtr_fields
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: dir(*), prefix(*)
      integer(c_int64_t), value :: itt
      type(sp_vtr_fields), intent(in) :: fields
      integer(c_int), value :: device
    end function sp_write_vtr

    ! dir: NUL-terminated, "" for the working folder or ending with '/'
    integer(c_int) function sp_write_vtk(ctx, dir, itt, fields, device) &
        bind(C, name="sp_write_vtk")
      import :: c_ptr, c_int, c_char, c_int64_t, sp_vtk_fields
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int64_t), value :: itt
      type(sp_vtk_fields), intent(in) :: fields
      integer(c_int), value :: device
    end function sp_write_vtk

    integer(c_int) function sp_interpolate_lvs(ctx, body, grid, device) &
        bind(C, name="sp_interpolate_lvs")
      import :: c_ptr, c_int, sp_solid_body, sp_lvs_grid
      type(c_ptr), value :: ctx
      type(sp_solid_body), intent(in) :: body
      type(sp_lvs_grid), intent(in) :: grid
      integer(c_int), value :: device
    end function sp_interpolate_lvs

    integer(c_int) function sp_clsvof_scheme(ctx, inputs, state, device) &
        bind(C, name="sp_clsvof_scheme")
      import :: c_ptr, c_int, sp_clsvof_inputs, sp_clsvof_state
      type(c_ptr), value :: ctx
      type(sp_clsvof_inputs), intent(in) :: inputs
      type(sp_clsvof_state), intent(inout) :: state
      integer(c_int), value :: device
    end function sp_clsvof_scheme

    integer(c_int) function sp_grid_preprocess(ctx, p, u, v, w) &
        bind(C, name="sp_grid_preprocess")
      import :: c_ptr, c_int, sp_cell_geom
      type(c_ptr), value :: ctx
      type(sp_cell_geom), intent(in) :: p, u, v, w
    end function sp_grid_preprocess

    integer(c_int) function sp_grid_preprocess_dev(ctx, p, u, v, w) &
        bind(C, name="sp_grid_preprocess_dev")
      import :: c_ptr, c_int, sp_cell_geom
      type(c_ptr), value :: ctx
      type(sp_cell_geom), intent(in) :: p, u, v, w
    end function sp_grid_preprocess_dev

    integer(c_int) function sp_compute_uvw_liquid_field(ctx, p, u, v, w, device) &
        bind(C, name="sp_compute_uvw_liquid_field")
      import :: c_ptr, c_int, sp_liquid_field
      type(c_ptr), value :: ctx
      type(sp_liquid_field), intent(in) :: p, u, v, w
      integer(c_int), value :: device
    end function sp_compute_uvw_liquid_field

    ! PredictingUVW with SpaceOrder 1 or 2; device = 0 for host arrays
    integer(c_int) function sp_predict_uvw(ctx, inp, outp, device) bind(C, name="sp_predict_uvw")
      import :: c_ptr, c_int, sp_pred_inputs, sp_pred_outputs
      type(c_ptr), value :: ctx
      type(sp_pred_inputs), intent(in) :: inp
      type(sp_pred_outputs), intent(inout) :: outp
      integer(c_int), value :: device
    end function sp_predict_uvw

    ! NumberExternalCell without device allocation: cell_type, posnu and
    ! ext_cell are device memory (c_ptr); the call only enqueues work
    integer(c_int) function sp_number_external_cell_noalloc(ctx, iu, iv, iw, cell_type, posnu, &
        ext_cell) bind(C, name="sp_number_external_cell_noalloc")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: iu, iv, iw
      type(c_ptr), value :: cell_type, posnu, ext_cell
    end function sp_number_external_cell_noalloc

    integer(c_int) function sp_new_cell_face(ctx, p, u, v, w, iu, iv, iw, device) &
        bind(C, name="sp_new_cell_face")
      import :: c_ptr, c_int, sp_cell_geom, sp_cell_interp
      type(c_ptr), value :: ctx
      type(sp_cell_geom), intent(in) :: p, u, v, w
      type(sp_cell_interp), intent(in) :: iu, iv, iw
      integer(c_int), value :: device
    end function sp_new_cell_face

    ! c_loc of PCell%Cell_Type, UCell/VCell/WCell%Cell_Type and of
    ! UCell/VCell/WCell%MoExCell; device = 0 for host arrays
    integer(c_int) function sp_define_momentum_exchange_cell(ctx, ctp, ctu, ctv, ctw, mu, &
        mv, mw, device) bind(C, name="sp_define_momentum_exchange_cell")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, ctp, ctu, ctv, ctw, mu, mv, mw
      integer(c_int), value :: device
    end function sp_define_momentum_exchange_cell

    ! TCell%PosNu numbering; ext_cell receives TCell%ExtCell
    integer(c_int) function sp_number_external_cell(ctx, iu, iv, iw, cell_type, posnu, &
        ext_cell, device) bind(C, name="sp_number_external_cell")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: iu, iv, iw
      type(c_ptr), value :: cell_type, posnu
      integer(c_int32_t), intent(out) :: ext_cell
      integer(c_int), value :: device
    end function sp_number_external_cell

    type(c_ptr) function sp_strerror_c(s) bind(C, name="sp_strerror")
      import :: c_ptr, c_int
      integer(c_int), value :: s
    end function sp_strerror_c

    type(c_ptr) function sp_last_error_c(ctx) bind(C, name="sp_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function sp_last_error_c

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! C string -> Fortran character
  function c_to_f(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
      s(i:i) = chars(i)
    end do
  end function c_to_f

  function sp_strerror_f(status) result(s)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: s
    s = c_to_f(sp_strerror_c(status))
  end function sp_strerror_f

  function sp_last_error_f(ctx) result(s)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: s
    s = c_to_f(sp_last_error_c(ctx))
  end function sp_last_error_f

end module simco_proj
