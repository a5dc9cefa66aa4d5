! Prints the size and selected member offsets of the bind(C) types so the
! tests can compare them with the C structs of simco_proj.h.
program sp_layout
  use iso_c_binding
  use simco_proj
  implicit none
  type(sp_params), target :: prm
  type(sp_inputs), target :: inp
  type(sp_outputs), target :: outp
  type(sp_stats), target :: st
  type(sp_adj_inputs), target :: ai
  type(sp_adj_outputs), target :: ao
  type(sp_wall_bc), target :: wb
  type(sp_vtr_fields), target :: vf
  type(sp_vtk_fields), target :: kf
  type(sp_solid_body), target :: sb
  type(sp_lvs_grid), target :: lg
  type(sp_clsvof_inputs), target :: csi
  type(sp_clsvof_state), target :: css
  type(sp_cell_geom), target :: cg
  type(sp_cell_interp), target :: ci
  type(sp_liquid_field), target :: lf
  type(sp_pred_inputs), target :: pin
  type(sp_pred_outputs), target :: pout
  print '(a,1x,i0)', 'sp_params', c_sizeof(prm)
  print '(a,1x,i0)', 'sp_inputs', c_sizeof(inp)
  print '(a,1x,i0)', 'sp_outputs', c_sizeof(outp)
  print '(a,1x,i0)', 'sp_stats', c_sizeof(st)
  print '(a,1x,i0)', 'sp_adj_inputs', c_sizeof(ai)
  print '(a,1x,i0)', 'sp_adj_outputs', c_sizeof(ao)
  print '(a,1x,i0)', 'sp_wall_bc', c_sizeof(wb)
  print '(a,1x,i0)', 'sp_vtr_fields', c_sizeof(vf)
  print '(a,1x,i0)', 'sp_vtk_fields', c_sizeof(kf)
  print '(a,1x,i0)', 'sp_solid_body', c_sizeof(sb)
  print '(a,1x,i0)', 'sp_lvs_grid', c_sizeof(lg)
  print '(a,1x,i0)', 'sp_clsvof_inputs', c_sizeof(csi)
  print '(a,1x,i0)', 'sp_clsvof_state', c_sizeof(css)
  print '(a,1x,i0)', 'sp_cell_geom', c_sizeof(cg)
  print '(a,1x,i0)', 'sp_cell_interp', c_sizeof(ci)
  print '(a,1x,i0)', 'sp_liquid_field', c_sizeof(lf)
  print '(a,1x,i0)', 'sp_pred_inputs', c_sizeof(pin)
  print '(a,1x,i0)', 'sp_pred_outputs', c_sizeof(pout)
  print '(a,1x,i0)', 'sp_pred_inputs.bc_flag', off(c_loc(pin%bc_flag), c_loc(pin))
  print '(a,1x,i0)', 'sp_pred_inputs.bc_var', off(c_loc(pin%bc_var), c_loc(pin))
  print '(a,1x,i0)', 'sp_pred_inputs.itt', off(c_loc(pin%itt), c_loc(pin))
  print '(a,1x,i0)', 'sp_pred_inputs.space_order', off(c_loc(pin%space_order), c_loc(pin))
  print '(a,1x,i0)', 'sp_pred_inputs.g', off(c_loc(pin%g), c_loc(pin))
  print '(a,1x,i0)', 'sp_pred_outputs.dp_w', off(c_loc(pout%dp_w), c_loc(pout))
  print '(a,1x,i0)', 'sp_liquid_field.vof_l', off(c_loc(lf%vof_l), c_loc(lf))
  print '(a,1x,i0)', 'sp_cell_interp.sxe', off(c_loc(ci%sxe), c_loc(ci))
  print '(a,1x,i0)', 'sp_cell_interp.dalt', off(c_loc(ci%dalt), c_loc(ci))
  print '(a,1x,i0)', 'sp_cell_geom.cell_type', off(c_loc(cg%cell_type), c_loc(cg))
  print '(a,1x,i0)', 'sp_cell_geom.fct', off(c_loc(cg%fct), c_loc(cg))
  print '(a,1x,i0)', 'sp_cell_geom.delh', off(c_loc(cg%delh), c_loc(cg))
  print '(a,1x,i0)', 'sp_vtr_fields.mres', off(c_loc(vf%mres), c_loc(vf))
  print '(a,1x,i0)', 'sp_vtr_fields.halo_mask', off(c_loc(vf%halo_mask), c_loc(vf))
  print '(a,1x,i0)', 'sp_vtk_fields.dz', off(c_loc(kf%dz), c_loc(kf))
  print '(a,1x,i0)', 'sp_vtk_fields.vof', off(c_loc(kf%vof), c_loc(kf))
  print '(a,1x,i0)', 'sp_vtk_fields.halo_mask', off(c_loc(kf%halo_mask), c_loc(kf))
  print '(a,1x,i0)', 'sp_solid_body.dx', off(c_loc(sb%dx), c_loc(sb))
  print '(a,1x,i0)', 'sp_solid_body.phi', off(c_loc(sb%phi), c_loc(sb))
  print '(a,1x,i0)', 'sp_solid_body.cm', off(c_loc(sb%cm), c_loc(sb))
  print '(a,1x,i0)', 'sp_solid_body.rot', off(c_loc(sb%rot), c_loc(sb))
  print '(a,1x,i0)', 'sp_lvs_grid.nz', off(c_loc(lg%nz), c_loc(lg))
  print '(a,1x,i0)', 'sp_clsvof_inputs.lvs_flag', off(c_loc(csi%lvs_flag), c_loc(csi))
  print '(a,1x,i0)', 'sp_clsvof_inputs.vof_const', off(c_loc(csi%vof_const), c_loc(csi))
  print '(a,1x,i0)', 'sp_clsvof_inputs.dt', off(c_loc(csi%dt), c_loc(csi))
  print '(a,1x,i0)', 'sp_clsvof_state.lvs_var', off(c_loc(css%lvs_var), c_loc(css))
  print '(a,1x,i0)', 'sp_clsvof_state.vof_var', off(c_loc(css%vof_var), c_loc(css))
  print '(a,1x,i0)', 'sp_wall_bc.const_u', off(c_loc(wb%const_u), c_loc(wb))
  print '(a,1x,i0)', 'sp_wall_bc.const_p', off(c_loc(wb%const_p), c_loc(wb))
  print '(a,1x,i0)', 'sp_inputs.wall', off(c_loc(inp%wall), c_loc(inp))
  print '(a,1x,i0)', 'sp_inputs.static_geometry', off(c_loc(inp%static_geometry), c_loc(inp))
  print '(a,1x,i0)', 'sp_outputs.bc_var', off(c_loc(outp%bc_var), c_loc(outp))
  print '(a,1x,i0)', 'sp_stats.wall_div', off(c_loc(st%wall_div), c_loc(st))
  print '(a,1x,i0)', 'sp_params.mg_coarse_scale', off(c_loc(prm%mg_coarse_scale), c_loc(prm))
  print '(a,1x,i0)', 'sp_params.max_iter', off(c_loc(prm%max_iter), c_loc(prm))
  print '(a,1x,i0)', 'sp_params.div_tol', off(c_loc(prm%div_tol), c_loc(prm))
  print '(a,1x,i0)', 'sp_params.pin_i', off(c_loc(prm%pin_i), c_loc(prm))
  print '(a,1x,i0)', 'sp_params.flags', off(c_loc(prm%flags), c_loc(prm))
  print '(a,1x,i0)', 'sp_params.mg_omega', off(c_loc(prm%mg_omega), c_loc(prm))
  print '(a,1x,i0)', 'sp_params.mg_omega_coarse', off(c_loc(prm%mg_omega_coarse), c_loc(prm))
  print '(a,1x,i0)', 'sp_inputs.bc_flag', off(c_loc(inp%bc_flag), c_loc(inp))
  print '(a,1x,i0)', 'sp_inputs.bc_val', off(c_loc(inp%bc_val), c_loc(inp))
  print '(a,1x,i0)', 'sp_inputs.dt', off(c_loc(inp%dt), c_loc(inp))
  print '(a,1x,i0)', 'sp_inputs.beta_t', off(c_loc(inp%beta_t), c_loc(inp))
  print '(a,1x,i0)', 'sp_stats.active_cells', off(c_loc(st%active_cells), c_loc(st))
  print '(a,1x,i0)', 'sp_stats.t_total_ms', off(c_loc(st%t_total_ms), c_loc(st))
contains
  integer function off(a, b)
    type(c_ptr), intent(in) :: a, b
    integer(c_intptr_t) :: ia, ib
    ia = transfer(a, ia)
    ib = transfer(b, ib)
    off = int(ia - ib)
  end function off
end program sp_layout
