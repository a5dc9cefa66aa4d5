/*
 * simco_proj.h -- C ABI of the MI355X pressure-projection engine.
 *
 * One call replaces what SIMCOFlow's Fortran host does for the pressure
 * projection today:
 *   PoissonEquationSolver(PGrid,UGrid,VGrid,WGrid,PCell,UCell,VCell,WCell,
 *                         vofO,TVar,TPred,PU,PV,PW,BCp,PoCoef,Proj,dt)
 *        (src/ComputePUVW.f90:127-129, signature src/ProjectionP.f90:33-50)
 *   followed by the velocity-correction block (src/ComputePUVW.f90:179-232).
 * The GFM face coefficients (ProjectionP.f90:123-265), the Poisson matrix and
 * right-hand side (ProjectionP.f90:301-576), the HYPRE PCG/BoomerAMG solve
 * (ProjectionP.f90:95-116, 267-299), the pressure gather and ghosts
 * (ProjectionP.f90:578-643) and the correction all run on the GPU.
 *
 * Array conventions (exactly the host's Fortran arrays, passed by address):
 *   halo arrays      (0:Imax+1, 0:Jmax+1, 0:Kmax+1) float64, column-major,
 *                    element (i,j,k) at i + (Imax+2)*(j + (Jmax+2)*k)
 *   cell arrays      (1:Imax, 1:Jmax, 1:Kmax) int32 (Cell_Type, MoExCell)
 *   PoCoef           (1:Imax, 1:Jmax, 1:Kmax, 6) float64, faces W,S,B,E,N,T
 *   grid arrays      (1:Imax, 1:Jmax, 1:Kmax) float64 (Grid%dx, %dy, %dz)
 *   BC face values   VarW/VarE (Jmax,Kmax), VarS/VarN (Imax,Kmax),
 *                    VarB/VarT (Imax,Jmax)
 * Kinds: real(c_double) = real(dp); integer(c_int32_t) = integer(it4b).
 *
 * Threading: calls on one context are synchronous and not thread-safe; use one
 * context per host thread.  Every call returns an sp_status; none aborts.
 */
#ifndef SIMCO_PROJ_H
#define SIMCO_PROJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sp_status {
  SP_OK = 0,
  SP_ERR_INVALID_ARG = -1,  /* null pointer, bad size or option */
  SP_ERR_HIP = -2,          /* a HIP runtime call failed */
  SP_ERR_NOT_CONVERGED = -3,/* outputs written, stats.converged == 0 */
  SP_ERR_NONFINITE = -4,    /* NaN/Inf in the inputs or during the solve */
  SP_ERR_OOM = -5,          /* device allocation failed */
  SP_ERR_GRID = -6,         /* grid missing or not a tensor-product grid */
  SP_ERR_STATE = -7,        /* adjoint without a preceding forward call */
  SP_ERR_NO_DEVICE = -8,    /* no usable gfx950 device */
  SP_ERR_CAPACITY = -9      /* an engine work list is full: more GFM faces next to
                               solid cells than min(6N, max(2^20, N/2)) */
} sp_status;

typedef enum sp_precision {
  SP_PREC_FP64 = 0,  /* V-cycle and outer CG in FP64 */
  SP_PREC_MIXED = 1  /* FP32 V-cycle inside an FP64 outer CG */
} sp_precision;

typedef enum sp_precond {
  SP_PRECOND_MG = 0,     /* geometric multigrid V-cycle */
  SP_PRECOND_JACOBI = 1  /* diagonal (for comparison) */
} sp_precond;

/* Solver and physics parameters.  sp_params_default() fills the values the
 * reference hard-codes: densities (StateVariables.f90:10), epsi
 * (CutCell.f90:11), Roref = row (ProjectionP.f90:53). */
typedef struct sp_params {
  int32_t imax, jmax, kmax;
  int32_t precision;       /* sp_precision */
  int32_t precond;         /* sp_precond */
  int32_t mg_levels;       /* 0: coarsen while every dimension >= 4 and a
                              level holds more than 64 cells */
  int32_t mg_pre, mg_post; /* red-black Gauss-Seidel sweeps */
  int32_t mg_coarse_sweeps;
  double mg_coarse_scale;  /* coarse operator = scale * Galerkin face sums */
  int32_t max_iter;        /* outer CG iterations */
  double div_tol;          /* stop when max |div u| <= div_tol (1/time units) */
  double rho_l, rho_g, rho_ref; /* row, roa, Roref */
  double epsi;
  int32_t pin_i, pin_j, pin_k;  /* 1-based gauge cell of a singular system;
                                   0,0,0 -> zero mean over fluid cells */
  int32_t flags;                /* SP_FLAG_* bits, 0 by default */
  double mg_omega;              /* over-relaxation of the red-black Gauss-Seidel
                                   updates on the finest level,
                                   x += omega (x_gs - x); (0, 2) */
  double mg_omega_coarse;       /* the same on every coarser level; (0, 2) */
} sp_params;

/* sp_params.flags bits */
enum {
  SP_FLAG_UNFUSED_SMOOTHER = 1, /* one kernel per half-sweep instead of the
                                   temporally blocked smoother (same result) */
  SP_FLAG_CHEBYSHEV = 2         /* Chebyshev-Jacobi smoothing instead of
                                   red-black Gauss-Seidel on the levels above
                                   the single-workgroup tail: mg_pre / mg_post
                                   steps of the Chebyshev semi-iteration for
                                   D^-1 A on [lambda_max / 8, lambda_max],
                                   lambda_max = 2 (Gershgorin bound of a
                                   diagonally dominant operator); mg_omega is
                                   not used there.  Every lane updates every
                                   cell (no colours); one kernel per step. */
};

typedef struct sp_ctx sp_ctx;

/* The host's velocity and pressure boundary conditions (BCu, BCv, BCw, BCp:
 * %flag(1..6) and %const(1..6), BoundaryInterface.f90:7-21, set in the
 * drivers as Main.f90:92-105 does) for the wall step that follows the
 * correction in UpdatePUVW:
 *   BoundaryConditionVarNew (ComputePUVW.f90:255, 418-694): TypicalBC face
 *     values (BoundaryFunction.f90:760-781), the wall-normal faces
 *     u(0), u(Imax), v(:,0,:), v(:,Jmax,:), w(:,:,0), w(:,:,Kmax) and the
 *     top pressure layer p(:,:,Kmax) set to them, ghosts by
 *     ComputeGhostVarBoundary (:696-708);
 *   the mass fix of the first Dirichlet pressure face (:256-384);
 *   VariablesInternalCellCondition (Solver.f90:223, 467-501).
 * The pressure flags are sp_inputs.bc_flag.  Flags: 0 Dirichlet, 1 Neumann,
 * faces W,E,S,N,B,T. */
typedef struct sp_wall_bc {
  int32_t flag_u[6], flag_v[6], flag_w[6];
  double const_u[6], const_v[6], const_w[6], const_p[6];
} sp_wall_bc;

/* Inputs of one projection.  Pointers are host memory for sp_project and
 * device memory for sp_project_dev. */
typedef struct sp_inputs {
  const double *u_star, *v_star, *w_star;  /* TPred%u,v,w (Predictor)     halo */
  const double *dp_u, *dp_v, *dp_w;        /* PU%Dp, PV%Dp, PW%Dp          halo */
  const double *ee_area, *ne_area, *te_area; /* PCell%EEArea,NEArea,TEArea halo */
  const double *vof, *vof_old;             /* PCell%vof, PCellO%vof        halo */
  const double *vof_l, *phi_l;             /* PCell%vofL, PCell%phiL       halo */
  const int32_t *cell_type_p;              /* PCell%Cell_Type              cell */
  const int32_t *cell_type_u, *cell_type_v, *cell_type_w; /* U/V/WCell%Cell_Type */
  const int32_t *moex_u, *moex_v, *moex_w; /* U/V/WCell%MoExCell           cell */
  int32_t bc_flag[6];                      /* BCp%flag: W,E,S,N,B,T; 0 Dirichlet, 1 Neumann */
  const double *bc_val[6];                 /* BCp%VarW..VarT, NULL for Neumann faces */
  double dt;
  /* Optional: face coefficients beta = 1/rho' on the E,N,T faces
   * (cell arrays).  When set the GFM step is skipped and
   * PoCoef(:,4..6) = Dp/delta * beta; W,S,B faces take the neighbour's value.
   * This is the parametrisation the adjoint differentiates. */
  const double *beta_e, *beta_n, *beta_t;
  /* Optional: the wall step after the correction (NULL: the host keeps it). */
  const sp_wall_bc *wall;
  /* sp_project only (host arrays): 1 when ee/ne/te_area, vof, vof_old, the
   * four Cell_Type arrays and the three MoExCell arrays hold the same values
   * as in the previous sp_project call on this context (the solid has not
   * moved): the copies already on the device are reused instead of being
   * sent over PCIe again.  Ignored (everything is copied) when the previous
   * call on the context was not an sp_project call that copied them; 0
   * always copies.  The caller answers for the claim: stale geometry is not
   * detected. */
  int32_t static_geometry;
} sp_inputs;

typedef struct sp_outputs {
  double *p;          /* Proj%Pp with ghosts (halo array, every entry written) */
  double *u, *v, *w;  /* TVar%u,v,w: entries (1:I,1:J,1:K) written, others kept */
  double *po_coef;    /* optional PoCoef (cell array x6), NULL to skip */
  /* Optional face values after the wall step, each NULL to skip:
   * BCu%VarW,VarE,VarS,VarN,VarB,VarT, then BCv, BCw, BCp (24 arrays of
   * (Jmax,Kmax) for W/E, (Imax,Kmax) for S/N, (Imax,Jmax) for B/T). */
  double *bc_var[24];
} sp_outputs;

typedef struct sp_stats {
  int32_t iterations;
  int32_t converged;      /* final_div <= div_tol, or the true residual max|b - A p|/V
                             is: with Dirichlet pressure faces, a vof - vofO source
                             or masked open faces the reference's divergence of the
                             corrected field is not the residual of the system */
  int32_t levels;
  int32_t singular;       /* all-Neumann operator (gauge applied) */
  int64_t active_cells;
  int64_t stale_faces;    /* GFM faces that reused the previous beta */
  double initial_div;     /* max |b|/V before the solve */
  double final_div;       /* max |div u| of the corrected field */
  double residual_div;    /* max |A p - b|/V of the recursive residual */
  double t_assemble_ms, t_solve_ms, t_correct_ms, t_total_ms;
  double wall_div;        /* max |mres|/V after the wall step (Solver.f90:236-250),
                             0 when no wall step ran */
} sp_stats;

/* Adjoint inputs: dL/du, dL/dv, dL/dw on the corrected faces (halo arrays,
 * entries (1:I,1:J,1:K) read) and, optionally, dL/dPp for the output
 * pressure Proj%Pp (halo array: the interior and the face ghosts are read,
 * since each ghost is flag * Pp of its inner cell, ProjectionP.f90:620-643;
 * the gradient follows the gauge of the forward call, pin cell or zero mean).
 * gp = NULL: L does not depend on the pressure. */
typedef struct sp_adj_inputs {
  const double *gu, *gv, *gw;
  const double* gp;
} sp_adj_inputs;

/* Adjoint outputs: dL/du*, dL/dv*, dL/dw* (halo arrays, every entry written)
 * and dL/dK for K = PoCoef(:,4..6) (cell arrays); optionally dL/dbeta for the
 * beta parametrisation (cell arrays, NULL to skip). */
typedef struct sp_adj_outputs {
  double *gu_star, *gv_star, *gw_star;
  double *gk_e, *gk_n, *gk_t;
  double *gbeta_e, *gbeta_n, *gbeta_t;
} sp_adj_outputs;

/* Fill the defaults: FP64, multigrid, automatic levels, 2+2 sweeps, coarse
 * scale 0.5, omega 1.25 (coarse levels 1.5), 200 iterations, div_tol 1e-10,
 * rho 998.3/1.205,
 * epsi 1e-3. */
void sp_params_default(sp_params* p);

/* Create a context on HIP device `device`.  The grid needs at least 2 cells
 * per direction and fewer than 2^29 cells (SP_ERR_INVALID_ARG otherwise). */
sp_status sp_create(sp_ctx** ctx, const sp_params* params, int device);
sp_status sp_destroy(sp_ctx* ctx);

/* Replace the solver parameters (same grid size). */
sp_status sp_set_params(sp_ctx* ctx, const sp_params* params);

/* Grid (Mesh.f90:14-18): PGrid%dx,dy,dz and the staggered centre distances
 * UGrid%dx, VGrid%dy, WGrid%dz (Mesh.f90:160-247), as the host's 3-D arrays.
 * The grid must be a tensor product (dx depends on i only, ...). */
sp_status sp_set_grid(sp_ctx* ctx, const double* dxp, const double* dyp, const double* dzp,
                      const double* dxu, const double* dyv, const double* dzw);
/* Same from 1-D spacings (lengths Imax, Jmax, Kmax). */
sp_status sp_set_grid_1d(sp_ctx* ctx, const double* dxp, const double* dyp, const double* dzp,
                         const double* dxu, const double* dyv, const double* dzw);

/* Projection with host arrays: the drop-in for PoissonEquationSolver + the
 * correction block.  stats may be NULL. */
sp_status sp_project(sp_ctx* ctx, const sp_inputs* in, sp_outputs* out, sp_stats* stats);
/* Same with device arrays (no host<->device copies). */
sp_status sp_project_dev(sp_ctx* ctx, const sp_inputs* in, sp_outputs* out, sp_stats* stats);

/* The wall step alone (sp_wall_bc) on out->u, v, w, p, updated in place from
 * their current values.  Reads ee_area, ne_area, te_area, cell_type_p/u/v/w,
 * bc_flag and wall of `in` (the other input pointers may be NULL); writes
 * out->bc_var and stats->wall_div.  Host and device variants. */
sp_status sp_wall_step(sp_ctx* ctx, const sp_inputs* in, sp_outputs* out, sp_stats* stats);
sp_status sp_wall_step_dev(sp_ctx* ctx, const sp_inputs* in, sp_outputs* out, sp_stats* stats);

/* Reverse-mode derivative of the most recent projection on this context.
 * Reuses the multigrid-preconditioned CG on the backward source (the
 * operator is symmetric); only the forward pressure is kept from the
 * forward call.  It differentiates the projection without the wall step:
 * after a forward call with sp_inputs.wall set it returns SP_ERR_STATE.
 * Host and device variants. */
sp_status sp_adjoint(sp_ctx* ctx, const sp_adj_inputs* in, sp_adj_outputs* out, sp_stats* stats);
sp_status sp_adjoint_dev(sp_ctx* ctx, const sp_adj_inputs* in, sp_adj_outputs* out,
                         sp_stats* stats);

/* Fields of PrintResultVTR3D (PrintResult.f90:97-147).  x, y, z are the
 * cell-centre coordinates TGrid%x(:,1,1), %y(1,:,1), %z(1,1,:) (lengths Imax,
 * Jmax, Kmax); the nine fields are read on (1:Imax,1:Jmax,1:Kmax).  Each field
 * is either a halo array (0:Imax+1, 0:Jmax+1, 0:Kmax+1) or a cell array
 * (1:Imax, 1:Jmax, 1:Kmax): bit q of halo_mask set marks field q (in the
 * order u, v, w, p, phi, phi_l, vof, vof_l, mres) as a halo array. */
typedef struct sp_vtr_fields {
  const double *x, *y, *z;
  const double *u, *v, *w, *p;   /* TVar%u, v, w, p */
  const double *phi, *phi_l;     /* TCell%phi, phiL (FluidLvs, LiquidLvs) */
  const double *vof, *vof_l;     /* TCell%vof, vofL (FluidVof, LiquidVof) */
  const double *mres;            /* TVar%mres (Mass error) */
  uint32_t halo_mask;
} sp_vtr_fields;

/* Write <dir>Paraview/<prefix>_<itt as 8 digits>.vtr, the file
 * PrintResultVTR3D writes through VTR_mod.f90 (file name VTR_mod.f90:116-130):
 * an ASCII VTK RectilinearGrid with the same elements, array names and order
 * (Velocity, X_/Y_/Z_Velocity, Pressure, FluidLvs, LiquidLvs, FluidVof,
 * LiquidVof, Mass error).  Every value is the field rounded to float32, as
 * real(., kind=4) does, printed with 10 significant digits so that a reader
 * gets back exactly that float32; the numbers are formatted on the GPU.
 * `dir` ends with '/'; missing folders of <dir>Paraview/ are created.  `device`: 1
 * when the array pointers are device memory, 0 for host memory. */
sp_status sp_write_vtr(sp_ctx* ctx, const char* dir, const char* prefix, int64_t itt,
                       const sp_vtr_fields* fields, int device);

/* Fields of PrintResultVTK (PrintResult.f90:328-420).  x, y, z are the
 * cell-centre coordinates TGrid%x(:,1,1), %y(1,:,1), %z(1,1,:) and dx, dy, dz
 * the cell widths TGrid%dx(:,1,1), %dy(1,:,1), %dz(1,1,:) (lengths Imax, Jmax,
 * Kmax); the eight fields are read on (1:Imax,1:Jmax,1:Kmax), each a halo
 * array (0:Imax+1, 0:Jmax+1, 0:Kmax+1) or a cell array (1:Imax, 1:Jmax,
 * 1:Kmax): bit q of halo_mask set marks field q (in the order p, u, v, w,
 * phi_l, phi, vof_l, vof) as a halo array. */
typedef struct sp_vtk_fields {
  const double *x, *y, *z;
  const double *dx, *dy, *dz;
  const double *p, *u, *v, *w;   /* TVar%p, u, v, w */
  const double *phi_l, *phi;     /* TCell%phiL, phi (LiquidLvs, FluidLvs) */
  const double *vof_l, *vof;     /* TCell%vofL, vof (LiquidVof, FluidVof) */
  uint32_t halo_mask;
} sp_vtk_fields;

/* Write <dir>PCell_Binary_3D<itt as i8.8>.vtr, the file PrintResultVTK writes
 * through VTKFortran's binary XML writer, byte for byte: a RectilinearGrid
 * with WholeExtent/Extent "+1 +Imax+1 +1 +Jmax+1 +1 +Kmax+1", Coordinates X,
 * Y, Z at the cell faces (centre -/+ half width, PrintResult.f90:347-361) and
 * CellData p, u, v, w, LiquidLvs, FluidLvs, LiquidVof, FluidVof, each array
 * the base64 code of its int32 byte count followed by its float32 values
 * (i fastest).  The base64 text is produced on the GPU.  itt outside
 * [0, 99999999] gives "********" in the name, as the i8.8 edit descriptor
 * does.  `dir` is "" (the working folder, as in the reference) or ends with
 * '/'; missing folders are created.  `device`: 1 when the array pointers are
 * device memory, 0 for host memory.  Grids of 2^29 cells or more are rejected
 * (SP_ERR_INVALID_ARG): the byte count is a 32-bit integer. */
sp_status sp_write_vtk(sp_ctx* ctx, const char* dir, int64_t itt, const sp_vtk_fields* fields,
                       int device);

/* One staggered cell grid (PGrid/PCell, UGrid/UCell, VGrid/VCell or
 * WGrid/WCell) as GridPreprocess (CutCell.f90:62-325) reads and writes it.
 * Cell arrays are (1:Imax,1:Jmax,1:Kmax); ee/ne/te_area are halo arrays;
 * fce/fcn/fct are TCell%FCE/FCN/FCT (0:Imax+1,0:Jmax+1,0:Kmax+1,3), the
 * component index slowest; cell_cent is TCell%Cell_Cent (1:Imax,..,3).
 * phi, nx, ny, nz, vof and vof_l are cell arrays as AllocateVar allocates
 * them (Main.f90:229-399) when halo_fields = 0, and halo arrays
 * (0:Imax+1,0:Jmax+1,0:Kmax+1) as the test drivers allocate them
 * (Test/FlowSphere2Phases/FlowSphere2Phases.f90:316-322, 431-440) when
 * halo_fields = 1; only their interior entries are read or written. */
typedef struct sp_cell_geom {
  const double *dx, *dy, *dz;         /* TGrid%dx, dy, dz (cell arrays) */
  const double *phi, *nx, *ny, *nz;   /* TCell%phi, nx, ny, nz: solid level set, normal */
  double *vof, *vof_l;                /* TCell%vof, vofL: read, then set by DefineCell */
  int32_t* cell_type;                 /* TCell%Cell_Type: written (kept where vof is NaN) */
  double *ee_area, *ne_area, *te_area;/* TCell%EEArea, NEArea, TEArea: written */
  double *fce, *fcn, *fct;            /* updated; entries the reference never sets keep
                                         their values */
  double* cell_cent;                  /* TCell%Cell_Cent: written */
  double* wllh;                       /* TCell%WlLh: written */
  double* delh;                       /* TCell%delh: updated where the reference sets it */
  int32_t halo_fields;                /* 1: phi, nx, ny, nz, vof, vof_l carry the halo;
                                         0: cell arrays; other values are rejected
                                         (SP_ERR_INVALID_ARG): initialise the struct */
} sp_cell_geom;

/* GridPreprocess(PGrid,UGrid,VGrid,WGrid, PCell,UCell,VCell,WCell)
 * (CutCell.f90:62-325): DefineCell and CellGeoCal on the four grids (the
 * MaxFace demotion on P only, whose MoExCell is never allocated), then the
 * U/V/W and P face and cell centre corrections.  Called at start-up
 * (Main.f90:203) and after each solid move (SolidBody.f90:815).  Every
 * output is bit-identical to the reference's loops.  The grid size is the
 * context's (sp_create); sp_set_grid is not needed.  Host and device
 * variants. */
sp_status sp_grid_preprocess(sp_ctx* ctx, const sp_cell_geom* p, const sp_cell_geom* u,
                             const sp_cell_geom* v, const sp_cell_geom* w);
sp_status sp_grid_preprocess_dev(sp_ctx* ctx, const sp_cell_geom* p, const sp_cell_geom* u,
                                 const sp_cell_geom* v, const sp_cell_geom* w);

/* The members of one U/V/W cell that NewCellFace reads or writes besides
 * those of sp_cell_geom: TCell%MoExCell (read), TCell%MsCe (1:Imax,..,3,
 * written for linked cells) and the interpolation data of the E, N, T
 * faces, TCell%SxE, EtaE, AlE, DAlE, SyN, ..., DAlT (halo arrays, written
 * where the reference writes them). */
typedef struct sp_cell_interp {
  const int32_t* moex;
  int32_t* msce;
  double *sxe, *etae, *ale, *dale;
  double *syn, *etan, *aln, *daln;
  double *szt, *etat, *alt, *dalt;
} sp_cell_interp;

/* NewCellFace(PGrid,UGrid,VGrid,WGrid, PCell,UCell,VCell,WCell)
 * (CutCell.f90:788-881): the U/V/W cell centres from the P face centres,
 * CellLinking (:1269-1355) of the MoExCell cells with the reference's
 * sequential semantics, delh, and East/North/TopFaceInterpolationInf
 * (:983-1267) for the three grids.  Reads P's dx, dy, dz, ee/ne/te_area
 * and fce/fcn/fct, and U/V/W's dx, dy, dz, phi, nx, ny, nz, cell_type,
 * ee/ne/te_area and fce/fcn/fct; updates U/V/W's cell_cent and delh.  A
 * link whose master would lie at index 0 (outside the grid: the reference
 * reads outside its arrays there) leaves the cell unlinked.  `device`: 1
 * when the pointers are device memory, 0 for host memory. */
sp_status sp_new_cell_face(sp_ctx* ctx, const sp_cell_geom* p, const sp_cell_geom* u,
                           const sp_cell_geom* v, const sp_cell_geom* w, const sp_cell_interp* iu,
                           const sp_cell_interp* iv, const sp_cell_interp* iw, int device);

/* DefineMomentumExchangeCell(PCell, UCell, VCell, WCell)
 * (CutCell.f90:673-751): a cut U (V, W) cell whose P cell or +1 P neighbour
 * along its axis is solid (its own P cell only, in the last cell of a line)
 * gets MoExCell = 1; the other MoExCell entries keep their values.  Cell
 * arrays (1:Imax,1:Jmax,1:Kmax).  `device`: 1 when the pointers are device
 * memory, 0 for host memory. */
sp_status sp_define_momentum_exchange_cell(sp_ctx* ctx, const int32_t* cell_type_p,
                                           const int32_t* cell_type_u,
                                           const int32_t* cell_type_v,
                                           const int32_t* cell_type_w, int32_t* moex_u,
                                           int32_t* moex_v, int32_t* moex_w, int device);

/* NumberExternalCell(iu, iv, iw, TCell) (CutCell.f90:753-786): posnu
 * (TCell%PosNu, shape (Imax-iu, Jmax-iv, Kmax-iw)) numbers the non-solid
 * cells (Cell_Type /= 2) from 0 in the loop order i, j, k with k
 * innermost, -1 for solid cells; *ext_cell (host memory) receives
 * TCell%ExtCell = count - 1.  iu, iv, iw are 0 or 1.  `device` as above,
 * for cell_type and posnu. */
sp_status sp_number_external_cell(sp_ctx* ctx, int iu, int iv, int iw, const int32_t* cell_type,
                                  int32_t* posnu, int32_t* ext_cell, int device);

/* Allocation-free variant of sp_number_external_cell on device arrays.  It
 * has no counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  cell_type, posnu and ext_cell
 * are device memory; the same numbering as sp_number_external_cell is
 * written to posnu and ExtCell to *ext_cell.  The scan's scratch is a
 * static 16 MiB device array (2^22 int32) that each GPU the library runs on
 * holds until the process ends, shared by every call on that GPU: calls
 * that overlap on one GPU (other host threads, other contexts' streams)
 * race on it.  The call only enqueues work on sp_stream(ctx) and returns;
 * the outputs are valid once that stream has synchronized.  A grid whose
 * numbering needs more scratch (more than about 4.19 M pieces of 64 planes,
 * (Imax-iu)*(Jmax-iv)*ceil((Kmax-iw)/64)) returns SP_ERR_INVALID_ARG.  No
 * hipMalloc, hipFree or allocating library routine is called. */
sp_status sp_number_external_cell_noalloc(sp_ctx* ctx, int iu, int iv, int iw,
                                          const int32_t* cell_type, int32_t* posnu,
                                          int32_t* ext_cell);

/* The liquid fields of one grid for ComputeUVWLiquidField: TGrid%dx, dy,
 * dz (cell arrays (1:Imax,1:Jmax,1:Kmax)) and TCell%phiL, nxL, nyL, nzL,
 * vofL: cell arrays as AllocateVar allocates them (Main.f90:269-270,
 * 383-387) when halo_fields = 0, halo arrays (0:Imax+1,..) as the test
 * drivers and the projection's sp_inputs take them when halo_fields = 1
 * (only the interior entries are read or written). */
typedef struct sp_liquid_field {
  const double *dx, *dy, *dz;
  double *phi_l, *nx_l, *ny_l, *nz_l, *vof_l;
  int32_t halo_fields;  /* 0 or 1 as in sp_cell_geom; other values are rejected */
} sp_liquid_field;

/* ComputeUVWLiquidField(PGrid,PCell,UCell,VCell,WCell,UGrid,VGrid,WGrid)
 * (Clsvof.f90:174-298): U/V/W phiL, nxL, nyL, nzL averaged from the P cells
 * (DirectionAverageArray, :328-353) and U/V/W vofL from the liquid volume
 * of the two P half cells of each face (Volume_Fraction_Calc, :1475-1550).
 * Reads P's members and U/V/W's dx, dy, dz; writes U/V/W's phi_l, nx_l,
 * ny_l, nz_l, vof_l.  The reference's x**3.d0 is a pow call, here as
 * there, so vof_l can differ from a host build by the last bits of pow;
 * everything else is bit-identical.  `device`: 1 when the pointers are
 * device memory, 0 for host memory. */
sp_status sp_compute_uvw_liquid_field(sp_ctx* ctx, const sp_liquid_field* p,
                                      const sp_liquid_field* u, const sp_liquid_field* v,
                                      const sp_liquid_field* w, int device);

/* A solid body (TSolidBody, SolidBody.f90:25-54): its uniform local mesh of
 * imax x jmax x kmax cells (centres this%x, y, z; spacings this%dx, dy, dz),
 * its level set this%phi (imax, jmax, kmax), the centre of mass
 * this%CMpoint and the rotation this%rotMatLtoG (3x3, column-major as in
 * Fortran). */
typedef struct sp_solid_body {
  int32_t imax, jmax, kmax;
  double dx, dy, dz;
  const double *x, *y, *z;
  const double* phi;
  double cm[3];
  double rot[9];
} sp_solid_body;

/* One grid for interpolateLvS: TGrid%x, y, z (cell centres) and dx, dy, dz
 * (cell arrays (1:Imax,1:Jmax,1:Kmax)); TCell%phi, vof (halo arrays,
 * written over the whole array) and nx, ny, nz (halo arrays, written where
 * the reference writes them and kept elsewhere). */
typedef struct sp_lvs_grid {
  const double *x, *y, *z;
  const double *dx, *dy, *dz;
  double *phi, *vof;
  double *nx, *ny, *nz;
} sp_lvs_grid;

/* TSolidBody%interpolateLvS(grid, cell) (SolidBody.f90:646-742), the first
 * step of the moving-solid geometry update (SolidBody.f90:810-813): phi =
 * 1e4 and vof = 1 everywhere, then at every cell centre that falls inside
 * the body mesh the trilinear interpolation of the body's level set; in the
 * band |phi| <= 4 max(dx, dy, dz) the normal (Normal_Vector_Irre,
 * Clsvof.f90:2361-2456, or clamped differences on the outer layer),
 * normalised, and vof from Volume_Fraction_Calc.  Outside the band vof stays
 * 1 (inside the body too) and the normals keep their values, as in the
 * reference.  phi and the normals are bit-identical to the reference's
 * loops; vof can differ by the last bits of pow (x**3.d0).  Call once per
 * grid (P, U, V, W).  `device`: 1 when the pointers (body arrays included)
 * are device memory, 0 for host memory. */
sp_status sp_interpolate_lvs(sp_ctx* ctx, const sp_solid_body* body, const sp_lvs_grid* grid,
                             int device);

/* Inputs of Clsvof_Scheme(PGrid,PCell,TVar,BCu,BCv,BCw,BCLvs,BCvof,Time,dt,
 * itt) (Clsvof.f90:355-798).  Halo arrays: TVar%u, v, w (face velocities)
 * and PCell%nx, ny, nz (the solid normals, which set the level-set ghosts);
 * cell arrays: PGrid%x, y, z, dx, dy, dz; BCu%VarW (Jmax,Kmax), BCv%VarS
 * (Imax,Kmax), BCw%VarB (Imax,Jmax) (the inflow velocities of the level-set
 * fluxes); BCLvs%flag/const and BCVof%flag/const, faces W,E,S,N,B,T (0
 * Dirichlet, 1 Neumann). */
typedef struct sp_clsvof_inputs {
  const double *u, *v, *w;
  const double *x, *y, *z;
  const double *dx, *dy, *dz;
  const double *nx, *ny, *nz;
  const double *bcu_w, *bcv_s, *bcw_b;
  int32_t lvs_flag[6];
  double lvs_const[6];
  int32_t vof_flag[6];
  double vof_const[6];
  double dt;
} sp_clsvof_inputs;

/* The state Clsvof_Scheme advances: PCell%vofL, phiL (halo, interior and
 * face ghosts written), PCell%nxL, nyL, nzL (halo, interior written; read
 * by the first sub-step's boundary values), BCLvs%VarW..VarT (read: VarW,
 * VarS, VarB feed the first sweeps; all written) and BCVof%VarW..VarT
 * (written; NULL entries are skipped). */
typedef struct sp_clsvof_state {
  double *vof_l, *phi_l;
  double *nx_l, *ny_l, *nz_l;
  double* lvs_var[6];
  double* vof_var[6];
} sp_clsvof_state;

/* Clsvof_Scheme (Clsvof.f90:355-798): six sub-steps of dt/6, each three
 * directional sweeps of the liquid volume fraction (geometric fluxes of the
 * reconstructed plane, East_Flux..Bottom_Flux) and level set (the
 * reference's Lax-Wendroff fluxes), rotating the sweep order, then
 * BoundaryConditionLvsVof, Interface_Reconstruct (Normal_Vector_Irre,
 * Find_Distance) and Redistance_Levelset.  The reference's loops are kept
 * with their boundary forms (the level-set face 1 of the x sweep divides by
 * 2 and multiplies by dx(2), the z sweep's face 1 adds and its face kmax
 * divides by dy, the north ghost subtracts the normal term).  Values are
 * those of the reference's loops except where pow (x**3.d0 in
 * Volume_Fraction_Calc and Newton_Raphson, x**(1/3) in Find_Distance) rounds
 * differently on the GPU.  Newton_Raphson stops after 200 iterations (the
 * reference has no limit).  Every dimension must be at least 3.  `device`:
 * 1 when the pointers are device memory, 0 for host memory. */
sp_status sp_clsvof_scheme(sp_ctx* ctx, const sp_clsvof_inputs* in, sp_clsvof_state* state,
                           int device);

/* Inputs of one momentum predictor step, PredictingUVW(PGrid, UGrid, VGrid,
 * WGrid, PCell, UCell, VCell, WCell, PCellO, UCellO, VCellO, WCellO,
 * FluxDivOld, TVar_n, TVar, BCu, BCv, BCw, PU, PV, PW, Pred, dt, itt)
 * (PredictUVW.f90:49-803), with SpaceOrder 1 (first-order upwind convection,
 * :1338-1794; every driver's input file selects it) or 2 (central,
 * ModifiedConvectiveFlux :2253-2644).  The grid is the
 * context's (sp_set_grid); the cells are the host's own arrays as the
 * geometry update writes them (sp_cell_geom / sp_cell_interp layouts). */
typedef struct sp_pred_inputs {
  /* PCellO%EEArea, NEArea, TEArea (halo) */
  const double *p_ee_area, *p_ne_area, *p_te_area;
  /* UCellO, VCellO, WCellO: phi, nx, ny, nz, vof, cell_type, ee/ne/te_area,
   * fce/fcn/fct and cell_cent are read */
  const sp_cell_geom *u_old, *v_old, *w_old;
  /* their moex, sxe, syn, szt, etae, etan, etat, ale, aln, alt are read */
  const sp_cell_interp *u_old_x, *v_old_x, *w_old_x;
  /* UCell, VCell, WCell: vof, vof_l, cell_type, wllh, delh are read */
  const sp_cell_geom *u_cur, *v_cur, *w_cur;
  const double *u, *v, *w;        /* TVar%u, v, w (halo) */
  const double *u_n, *v_n, *w_n;  /* TVar_n%u, v, w (halo); read for TimeOrder 2 only */
  double* flux_div_old;           /* FluxDivOld (Imax,Jmax,Kmax,3): read for TimeOrder 3
                                     after the first step, updated where the reference
                                     updates it */
  int32_t bc_flag[18];            /* BCu%flag(1:6), BCv%flag, BCw%flag: W,E,S,N,B,T */
  const double* bc_var[18];       /* BCu%VarW..VarT, BCv%..., BCw%... face arrays */
  double dt;
  int64_t itt;                    /* time step number (1: first-order start) */
  int32_t time_order;             /* TimeOrder: 1 Euler, 2 extrapolated velocity, 3
                                     Adams-Bashforth (StateVariables.f90:13) */
  int32_t space_order;            /* SpaceOrder: 1 or 2 (3 is not built); the first
                                     step (itt = 1) takes the first-order fluxes
                                     either way (:172-185) */
  double mu_w_ref, mu_a_ref;      /* muWRef = muw/muref, muARef = mua/muref */
  double rey, fr;                 /* Rey, Fr (StateVariables.f90:12) */
  double gx, gy, gz, g;           /* gravity components and g */
} sp_pred_inputs;

/* Outputs: Pred%u, v, w (halo arrays: the entries the reference writes are
 * written, the others kept) and PU%Dp, PV%Dp, PW%Dp (halo arrays, every
 * entry written). */
typedef struct sp_pred_outputs {
  double *u, *v, *w;
  double *dp_u, *dp_v, *dp_w;
} sp_pred_outputs;

/* PredictingUVW (PredictUVW.f90:49-803) for SpaceOrder 1 and 2.
 * DiffusiveFlux zeroes its coefficients (:3672-3673), so the three HYPRE
 * momentum solves (:692-759) are diagonal: u* = rhs / aP.  Every output
 * equals a sequential evaluation of the reference's expressions bit for
 * bit; cut-cell tests that read a cell array one past its end (index
 * Jmax + 1 at :1416-1430 and :1744-1751, and their SpaceOrder 2 siblings)
 * read "no such neighbour".  SpaceOrder 2 keeps the reference's carry of
 * its local delh from one cut face to the next in loop order (i outermost,
 * k innermost); before the first cut face of a call that sets it, the
 * face's own delhec is taken (the reference reads an undefined value
 * there).  Returns SP_ERR_INVALID_ARG for SpaceOrder 3.
 * `device`: 1 when every pointer is device memory, 0 for host memory. */
sp_status sp_predict_uvw(sp_ctx* ctx, const sp_pred_inputs* in, sp_pred_outputs* out,
                         int device);

/* HIP stream the context runs on (hipStream_t as void*). */
void* sp_stream(sp_ctx* ctx);

/* Per-kernel device timings of the last call, measured with HIP events:
 * fills up to `cap` records, returns how many exist. */
typedef struct sp_kernel_time {
  char name[48];
  int32_t launches;
  double total_ms;
  double bytes;  /* algorithmic bytes moved by all launches */
} sp_kernel_time;
int sp_kernel_times(sp_ctx* ctx, sp_kernel_time* out, int cap);
/* Enable (1) / disable (0) the per-kernel event timing (off by default). */
void sp_set_profiling(sp_ctx* ctx, int on);
/* Restrict the timing to launches of the kernel `name` (as reported by
 * sp_kernel_times); NULL or "" times every kernel.  Clears the records. */
void sp_set_profiling_filter(sp_ctx* ctx, const char* name);

const char* sp_strerror(sp_status s);
const char* sp_last_error(sp_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SIMCO_PROJ_H */
