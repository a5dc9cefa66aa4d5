/*
 * CPU FP64 restatement of SIMCOFlow's pressure projection -- TEST
 * INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or as the
 * timed CPU baseline; the engine never calls it.
 *
 * Layouts (all Fortran column-major):
 *   halo arrays      (0:I+1, 0:J+1, 0:K+1)  index i + (I+2)*(j + (J+2)*k)
 *   interior arrays  (1:I, 1:J, 1:K)        index (i-1) + I*((j-1) + J*(k-1))
 *   PoCoef           (1:I, 1:J, 1:K, 6)     faces W,S,B,E,N,T
 *   1-D spacings     dxp/dyp/dzp = P-cell widths, dxu/dyv/dzw = centre
 *                    distances of the staggered U/V/W grids (Mesh.f90:160-247)
 *
 * Pinning (the checks that would fail on a dropped term, wrong sign, index or
 * transposed operand; tests/test_oracle_*.py and tests/test_golden.py):
 *   orc_gfm_coef        closed forms of the four wet/dry cases, layered
 *                       series average, face symmetry, boundary copy, the
 *                       stale-beta rule on a hand-built column
 *   orc_gfm_coef_fast   bit-identical to orc_gfm_coef on every config
 *   orc_matrix_rows,    textbook 7-point Laplacian (single phase), symmetry
 *   orc_rhs             and zero row sums, b = -V for u* = (x,0,0), b = 0 for a
 *                       discrete curl, solid rows, Dirichlet wall terms
 *   orc_sym_*           equal to the reference rows on the active cells
 *   orc_pcg             sparse direct (SuperLU) and dense solves, O(h^2) on a
 *                       manufactured cosine solution
 *   orc_correct,        divergence-free output, idempotence, div-free input
 *   orc_pressure_field, returned unchanged, masks and pressure ghosts
 *   orc_mres
 *   orc_adjoint         dot-product test of the linear map, central
 *                       differences in u* and in PoCoef(:,4..6)
 *   orc_wall_step       a transcription of BoundaryConditionVarNew in numpy
 *                       slice semantics (the assumed-shape halo offset comes
 *                       out of the slicing itself), Dirichlet faces equal to
 *                       their constants, ghost rules, zero divergence of the
 *                       mass-fixed boundary layer for uniform apertures
 *   orc_face_geo        exact rational clipping of the face rectangle by the
 *                       half-plane (an independent algorithm), closed forms of
 *                       full, empty, half and corner-triangle faces
 *   orc_grid_preprocess planar level sets: shared faces of neighbours agree,
 *                       WlLh equals the exact cut area of the plane in the
 *                       cell, U/V/W faces through the P centre equal the exact
 *                       mid-plane fraction; every face next to a solid cell is
 *                       zero; DefineCell classes and vofL scaling; the
 *                       stretched-grid face centre rule and the FCT(:,2)
 *                       quirk on hand-built cells
 *   orc_volume_fraction the exact volume of a plane-cut box by independent
 *                       integration (numpy) and closed forms (empty, full,
 *                       corner tetrahedron, axis-aligned slab)
 *   orc_uvw_liquid_field  planar liquid level sets: U vofL equals the exact
 *                       liquid fraction of the U cell; lambda-weighted
 *                       averages; the V/W quirks on hand-built columns
 *   orc_define_moex,    vectorised restatements over shifted arrays (numpy),
 *   orc_number_external a cumulative count in C order of the (i,j,k) array
 * No reference golden vector exists for this path and HYPRE is absent, so the
 * HYPRE iterate itself is not pinned (any converged solve of the pinned,
 * uniquely solvable system is).
 */
#ifndef SIMCO_ORACLE_H
#define SIMCO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
  int32_t I, J, K;
  const double *dxp, *dyp, *dzp;
  const double *dxu, *dyv, *dzw;
} orc_grid;

/* Compute1DGFMCoefficient x6 (ProjectionP.f90:53-64, 123-265). */
void orc_gfm_coef(const orc_grid* g, const double* phiL, const double* vofL, const double* vof,
                  const double* dpu, const double* dpv, const double* dpw,
                  double rho_l, double rho_g, double rho_ref, double epsi, double* pocoef);

/* Parallel evaluation with the same result as orc_gfm_coef (CPU baseline). */
void orc_gfm_coef_fast(const orc_grid* g, const double* phiL, const double* vofL, const double* vof,
                       const double* dpu, const double* dpv, const double* dpw,
                       double rho_l, double rho_g, double rho_ref, double epsi, double* pocoef);

/* SetPoissonMatrix rows (ProjectionP.f90:301-482): for every cell, 7 values
 * in the order W,S,B,diag,E,N,T and the neighbour's interior index (-1 when the
 * reference omits the entry).  Solid rows (cell_type==2) get all -1 / 0. */
void orc_matrix_rows(const orc_grid* g, const double* pocoef, const double* ee, const double* ne,
                     const double* te, const int32_t* ct, const int32_t* bc_flag,
                     double* vals, int64_t* cols);

/* SetPoissonVectors RHS (ProjectionP.f90:484-576), interior layout, 0 in solid cells.
 * bc_val[f] may be NULL for Neumann faces; sizes (J,K) for W/E, (I,K) for S/N, (I,J) for B/T. */
void orc_rhs(const orc_grid* g, const double* pocoef, const double* ee, const double* ne,
             const double* te, const int32_t* ct, const int32_t* bc_flag, const double* const* bc_val,
             const double* us, const double* vs, const double* ws, const double* vof,
             const double* vof_old, double dt, double* b);

/* y = A x with the reference rows. */
void orc_rows_apply(int64_t n, const double* vals, const int64_t* cols, const double* x, double* y);

/* Symmetric face operator used by the engine: c_x/c_y/c_z hold the coupling
 * of cell (i,j,k) to its +x/+y/+z neighbour taken from the correction's own
 * coefficient PoCoef(:,4..6) (ComputePUVW.f90:187,199,211); e holds the
 * diagonal-only terms (Dirichlet walls, faces towards solid cells) with the
 * row's own coefficient as in ProjectionP.f90:379-430. */
void orc_sym_operator(const orc_grid* g, const double* pocoef, const double* ee, const double* ne,
                      const double* te, const int32_t* ct, const int32_t* bc_flag,
                      double* cx, double* cy, double* cz, double* e);

/* Diagonal of the symmetric operator; returns the number of active cells. */
int64_t orc_sym_diag(int I, int J, int K, const double* cx, const double* cy, const double* cz,
                     const double* e, double* d);
void orc_sym_apply(int I, int J, int K, const double* cx, const double* cy, const double* cz,
                   const double* d, const double* x, double* y);

typedef struct orc_solver_opts {
  int32_t levels;       /* 0 = automatic */
  int32_t pre, post;    /* RB-GS sweeps */
  int32_t coarse_sweeps;
  int32_t max_iter;
  double  coarse_scale; /* coarse operator = scale * Galerkin sum */
  double  div_tol;      /* stop when max |r|/V <= div_tol */
  int32_t use_rows;     /* 1: outer operator = reference rows (vals/cols) */
  int32_t precond;      /* 0 none, 1 Jacobi, 2 multigrid */
  double  omega;        /* RB-GS over-relaxation on the finest level: x += omega (x_gs - x) */
  double  omega_coarse; /* the same on every coarser level */
} orc_solver_opts;

typedef struct orc_solver_stats {
  int32_t iterations, converged, levels;
  double  initial_div, final_div;
  double  seconds;
} orc_solver_stats;

/* Flexible PCG with a cell-centred multigrid V-cycle preconditioner.
 * Singular (all-Neumann) systems are solved for the mean-projected b.
 * x is overwritten (interior layout); 0 in inactive cells. */
int orc_pcg(const orc_grid* g, const double* cx, const double* cy, const double* cz, const double* e,
            const double* vals, const int64_t* cols, const double* b, const orc_solver_opts* o,
            double* x, orc_solver_stats* st);

/* Gauge: pin cell (1-based) set to 0, or mean over active cells = 0 when pin
 * is (0,0,0) or inactive.  Only applied when the operator is singular. */
void orc_gauge(int I, int J, int K, const double* d, int singular, int pi, int pj, int pk, double* x);

/* DeltaPressureGetValues + DeltaPressureBoundaryCondition
 * (ProjectionP.f90:578-643): Pp halo array from the interior solution. */
void orc_pressure_field(const orc_grid* g, const int32_t* ct, const int32_t* bc_flag,
                        const double* x, double* Pp);

/* Velocity correction (ComputePUVW.f90:181-232).  u/v/w are written on
 * (1:I,1:J,1:K); other entries are left untouched. */
void orc_correct(const orc_grid* g, const double* pocoef, const double* Pp,
                 const int32_t* ctu, const int32_t* ctv, const int32_t* ctw,
                 const int32_t* mu, const int32_t* mv, const int32_t* mw,
                 const double* us, const double* vs, const double* ws,
                 double* u, double* v, double* w);

/* mres (Solver.f90:236-250), interior layout; returns max |mres|/V. */
double orc_mres(const orc_grid* g, const double* ee, const double* ne, const double* te,
                const double* u, const double* v, const double* w, double* mres);

/* Post-correction wall step of UpdatePUVW, in the reference's order:
 * BoundaryConditionVarNew (ComputePUVW.f90:418-694) with TypicalBC
 * (BoundaryFunction.f90:760-781) and ComputeGhostVarBoundary
 * (ComputePUVW.f90:696-708), the mass fix of the first Dirichlet pressure
 * face (ComputePUVW.f90:256-384), VariablesInternalCellCondition
 * (Solver.f90:467-501).  fl/cst: 24 flags and constants, BCu W..T, BCv, BCw,
 * BCp.  u, v, w, p (halo arrays) are updated in place; var (24 pointers,
 * each may be NULL) receives the BCx%VarW..VarT face values. */
void orc_wall_step(const orc_grid* g, const double* ee, const double* ne, const double* te,
                   const int32_t* ctp, const int32_t* ctu, const int32_t* ctv, const int32_t* ctw,
                   const int32_t* fl, const double* cst, double* u, double* v, double* w,
                   double* p, double* const* var);

/* Adjoint of the projection for L with dL/du = gu, dL/dv = gv, dL/dw = gw
 * (halo arrays, read on the corrected faces (1:I,1:J,1:K)) and dL/dPp = gp
 * (halo array with its face ghosts, or NULL) for the pressure in the gauge
 * pin_i, pin_j, pin_k of orc_gauge.  Gradients are returned for u*,v*,w*
 * (halo arrays, written on the faces the forward reads) and for the E/N/T
 * face coefficients K = PoCoef(:,4..6) (interior arrays). */
int orc_adjoint(const orc_grid* g, const double* pocoef, const double* ee, const double* ne,
                const double* te, const int32_t* ct, const int32_t* bc_flag,
                const int32_t* ctu, const int32_t* ctv, const int32_t* ctw,
                const int32_t* mu, const int32_t* mv, const int32_t* mw,
                const double* Pp, const double* gu, const double* gv, const double* gw,
                const double* gp, int pin_i, int pin_j, int pin_k,
                const orc_solver_opts* o, double* gus, double* gvs, double* gws,
                double* gkx, double* gky, double* gkz, double* lambda_out);

/* OpenMP threads the parallel functions use. */
int orc_num_threads(void);

/* One staggered cell grid of GridPreprocess (CutCell.f90:62-325): TGrid%dx,
 * dy, dz and the TCell members it reads or writes.  Cell arrays are
 * (1:I,1:J,1:K); ee/ne/te are halo arrays; fce/fcn/fct are halo arrays x 3
 * components (component slowest); cent is a cell array x 3. */
typedef struct orc_cell {
  const double *dx, *dy, *dz, *phi, *nx, *ny, *nz;
  double *vof, *vofl;
  int32_t* type;
  double *ee, *ne, *te, *fce, *fcn, *fct, *cent, *wllh, *delh;
} orc_cell;

/* NewCellFace data of one U/V/W grid: MoExCell (read), MsCe (cell array x
 * 3, written for linked cells) and the interpolation arrays (halo arrays)
 * of the E, N, T faces: s = SxE/SyN/SzT, eta = EtaE/N/T, al = AlE/N/T,
 * dal = DAlE/N/T. */
typedef struct orc_interp {
  const int32_t* moex;
  int32_t* msce;
  double *s[3], *eta[3], *al[3], *dal[3];
} orc_interp;

/* FaceGeoCal (CutCell.f90:883-981). */
void orc_face_geo(double d1, double d2, double n1, double n2, double n3, double d3, double phi,
                  double* area, double cent[3]);
/* GridPreprocess (CutCell.f90:62-325): DefineCell and CellGeoCal on the four
 * grids, then the face and cell centre corrections of U, V, W and P. */
void orc_grid_preprocess(int I, int J, int K, orc_cell* P, orc_cell* U, orc_cell* V,
                         orc_cell* W);

/* DefineMomentumExchangeCell (CutCell.f90:673-751). */
void orc_define_moex(int I, int J, int K, const int32_t* ctp, const int32_t* ctu,
                     const int32_t* ctv, const int32_t* ctw, int32_t* mu, int32_t* mv,
                     int32_t* mw);
/* NumberExternalCell (CutCell.f90:753-786); returns ExtCell. */
int32_t orc_number_external(int I, int J, int K, int iu, int iv, int iw, const int32_t* ct,
                            int32_t* posnu);

/* NewCellFace (CutCell.f90:788-881): U/V/W cell centres from the P face
 * centres, CellLinking of the MoExCell cells in the reference's sequential
 * order, delh, and the nine Face*InterpolationInf calls. */
void orc_new_cell_face(int I, int J, int K, const orc_cell* P, orc_cell* U, orc_cell* V,
                       orc_cell* W, orc_interp* iu, orc_interp* iv, orc_interp* iw);

/* The liquid fields of one grid for ComputeUVWLiquidField: TGrid%dx, dy,
 * dz and TCell%phiL, nxL, nyL, nzL, vofL (cell arrays). */
typedef struct orc_liquid {
  const double *dx, *dy, *dz;
  double *phi, *nx, *ny, *nz, *vof;
} orc_liquid;
/* Volume_Fraction_Calc (Clsvof.f90:1475-1550). */
double orc_volume_fraction(double dx, double dy, double dz, double nxx, double nyy, double nzz,
                           double s);
/* ComputeUVWLiquidField (Clsvof.f90:174-298): U/V/W phiL, nxL, nyL, nzL
 * by DirectionAverageArray and vofL from the two halves of the P cells. */
void orc_uvw_liquid_field(int I, int J, int K, const orc_liquid* P, orc_liquid* U, orc_liquid* V,
                          orc_liquid* W);

/* ---- Momentum predictor (PredictUVW.f90), SpaceOrder 1 and 2 (orc_pred.c) ---- */
/* The members of one staggered cell set (UCell, VCell or WCell) that
 * PredictingUVW reads: cell arrays (1:I,1:J,1:K) and halo arrays
 * (0:I+1,0:J+1,0:K+1); fce/fcn/fct halo x3, cent cell x3. */
typedef struct orc_pred_cell {
  const double *phi, *nx, *ny, *nz, *vof, *vof_l;
  const int32_t *cell_type, *moex;
  const double *ee, *ne, *te, *fce, *fcn, *fct, *cent;
  const double *sxe, *syn, *szt, *etae, *etan, *etat, *ale, *aln, *alt;
  const double *wllh, *delh;
} orc_pred_cell;

typedef struct orc_pred_params {
  double dt;
  int64_t itt;
  int32_t time_order, space_order;
  double rho_a_ref, rho_w_ref, mu_a_ref, mu_w_ref, rey, fr, gx, gy, gz, g;
} orc_pred_params;

/* ModifiedConvectiveFluxFirstOrder x3 (PredictUVW.f90:1338-1794): CFEW
 * (0:I+1,J,K,3), CFNS (I,0:J+1,K,3), CFTB (I,J,0:K+1,3); bcv: BCu%VarW..VarT,
 * BCv%..., BCw%... (18 face arrays). */
void orc_conv_flux_first(const orc_grid* g, const double* pee, const double* pne,
                         const double* pte, const orc_pred_cell* U, const orc_pred_cell* V,
                         const orc_pred_cell* W, const double* un, const double* vn,
                         const double* wn, const double* const* bcv, double* cfew, double* cfns,
                         double* cftb);

/* ModifiedConvectiveFlux x3 (PredictUVW.f90:2253-2644), SpaceOrder 2: the
 * same arrays; the local delh carried between loop iterations as the
 * reference's loop leaves it. */
void orc_conv_flux_second(const orc_grid* g, const double* pee, const double* pne,
                          const double* pte, const orc_pred_cell* U, const orc_pred_cell* V,
                          const orc_pred_cell* W, const double* un, const double* vn,
                          const double* wn, const double* const* bcv, double* cfew, double* cfns,
                          double* cftb);

/* Of the last orc_conv_flux_second call: cut faces that read a carried
 * delh, and cut faces with none carried yet (test instrumentation). */
void orc_conv_flux_second_carries(int64_t* carried, int64_t* uncarried);

/* PredictingUVW (PredictUVW.f90:49-803) for SpaceOrder 1 and 2 (the first
 * step, itt = 1, takes the first-order fluxes either way): pred = Pred%u,v,w
 * (halo, updated where the reference writes), dp = PU/PV/PW%Dp (halo, every
 * entry written), flux_div_old (I,J,K,3) read and updated.  old = UCellO..,
 * cur = UCell..; vel = TVar%u,v,w, vel_n = TVar_n%u,v,w; bc_flag: BCu%flag,
 * BCv%flag, BCw%flag (18).  Returns -1 for SpaceOrder 3 (not restated). */
int orc_predict(const orc_grid* g, const double* pee, const double* pne, const double* pte,
                const orc_pred_cell* const old[3], const orc_pred_cell* const cur[3],
                const double* const vel[3], const double* const vel_n[3], double* flux_div_old,
                const int32_t* bc_flag, const double* const* bcv, const orc_pred_params* p,
                double* const pred[3], double* const dp[3]);

/* ---- Interface transport and solid level set (orc_clsvof.c) ---- */
/* A solid body of SolidBody.f90:25-54: its uniform local mesh (cell
 * centres x, y, z, spacings dx, dy, dz), its level set phi (imax, jmax,
 * kmax), the centre of mass and rotMatLtoG (column-major). */
typedef struct orc_solid_body {
  int32_t imax, jmax, kmax;
  double dx, dy, dz;
  const double *x, *y, *z;
  const double* phi;
  double cm[3];
  double rot[9];
} orc_solid_body;

/* Normal_Vector_Irre (Clsvof.f90:2361-2456) at interior cell (i, j, k):
 * phi a halo array, x, y, z cell arrays of the grid's cell centres. */
void orc_normal_vector_irre(int I, int J, int K, const double* x, const double* y,
                            const double* z, const double* phi, int i, int j, int k, double* nxx,
                            double* nyy, double* nzz);
/* interpolateLvS (SolidBody.f90:646-742) onto one grid: gx..gdz cell
 * arrays (TGrid%x, y, z, dx, dy, dz), phi/vof written over the whole halo
 * array, nx/ny/nz (halo) written in the band |phi| <= 4 max(d). */
void orc_interpolate_lvs(const orc_solid_body* B, int I, int J, int K, const double* gx,
                         const double* gy, const double* gz, const double* gdx, const double* gdy,
                         const double* gdz, double* phi, double* vof, double* nx, double* ny,
                         double* nz);

/* The inputs of Clsvof_Scheme (Clsvof.f90:355-798): TVar%u, v, w (halo),
 * PGrid%x, y, z, dx, dy, dz (cell arrays), PCell%nx, ny, nz (halo, the solid
 * normals used for the level-set ghosts), BCu%VarW (J,K), BCv%VarS (I,K),
 * BCw%VarB (I,J), BCLvs and BCVof %flag and %const (W,E,S,N,B,T), dt. */
typedef struct orc_clsvof_in {
  const double *u, *v, *w;
  const double *x, *y, *z, *dx, *dy, *dz;
  const double *nx, *ny, *nz;
  const double *bcu_w, *bcv_s, *bcw_b;
  int32_t lvs_flag[6];
  double lvs_const[6];
  int32_t vof_flag[6];
  double vof_const[6];
  double dt;
} orc_clsvof_in;
/* The state it updates: PCell%vofL, phiL, nxL, nyL, nzL (halo) and the face
 * values BCLvs%VarW..VarT, BCVof%VarW..VarT (null: not kept). */
typedef struct orc_clsvof_state {
  double *vof, *phi;
  double *nx, *ny, *nz;
  double* lvs_var[6];
  double* vof_var[6];
} orc_clsvof_state;
/* Find_Distance (Clsvof.f90:1311-1408). */
double orc_find_distance(double dx, double dy, double dz, double nxx, double nyy, double nzz,
                         double f);
/* One candidate distance of Redistance_Levelset for the cell at offset
 * (ii, jj, kk) from an interface cell (plane n, diss; size d); phit is the
 * target's level set. */
double orc_redist_candidate(double nxx, double nyy, double nzz, double diss, double dx, double dy,
                            double dz, int ii, int jj, int kk, double phit);
/* Clsvof_Scheme; returns -1 when a dimension is below 3. */
int orc_clsvof_scheme(int I, int J, int K, const orc_clsvof_in* in, orc_clsvof_state* st);

#ifdef __cplusplus
}
#endif
#endif
