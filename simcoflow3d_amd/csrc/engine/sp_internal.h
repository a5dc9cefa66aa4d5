// Internal declarations of the projection engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "simco_proj.h"

namespace sp {

constexpr int kMaxLevels = 24;
constexpr int64_t kTailCells = 512;  // levels this small run in one workgroup
constexpr int kTailLdsBytes = 144 * 1024;  // LDS of that workgroup (7 arrays per level)
constexpr int kRedBlockThreads = 256;

// Scalars of the outer CG, resident in device memory so no host round trip
// is needed between the kernels of one iteration.
struct Scalars {
  double rz;        // (r, z) of the current iteration
  double alpha, beta;
  double pq, zq;
  double rmax;      // max |r_i| / V_i of the recursive residual
  double bmean;     // mean of b over active cells (singular systems)
  double bsum, esum, bmax, shift, divmax;
  double walldiv;   // max |mres|/V after the wall step
  double dsum;      // sum of the diagonal over active cells
  double zshift;    // diagonal-weighted mean of z removed in singular systems
  double src_max;   // max |pbar_i| / V_i of the adjoint source
  double rsum, qsum;  // sums of r and q over the cells (for the z shift)
  double npart[8];  // scratch results of small reductions
  int64_t active;
  int32_t done;     // 1 once converged (kernels then return immediately)
  int32_t iter;
  int32_t first;
  int32_t nonfinite;
  int32_t singular;
  int32_t pad;
  uint32_t counter[16];  // last-block counters, one per reduction site
  uint32_t stale_count;  // GFM faces that need the previous-beta fixup
  uint32_t pad2[3];
};

// Reduction sites (each owns its partials)
enum RedSiteId : int {
  kSiteOperator = 0,
  kSiteResidual,
  kSiteDots,
  kSiteSpmv,
  kSiteUpdate,
  kSiteShift,
  kSiteDiv,
  kSiteAdj,
  kSiteAdj2,
  kSiteFold,  // first-stage folds of sites with one partial per 256 cells
  kNumSites
};
constexpr int kSiteValues = 8;  // values per site at most

// Counter slots (legacy single-level reductions)
enum CounterSlot : int {
  kCtrOperator = 0,
  kCtrInit = 1,
  kCtrDots = 2,
  kCtrSpmv = 3,
  kCtrUpdate = 4,
  kCtrShift = 5,
  kCtrDiv = 6,
  kCtrAdj = 7,
  kCtrAdj2 = 8,
  kCtrCheck = 9,
};

// One multigrid level (views; the storage is owned by sp_ctx::allocs).
template <typename T>
struct Level {
  int nx, ny, nz;
  int64_t n;
  T *cx, *cy, *cz;
  T* e;        // diagonal-only terms, nullptr when they are all zero
  T* e_store;  // storage for e
  T *x, *b;
  T* xs;       // pre-smoothed x (input of the fused post-smoother); the
               // Chebyshev smoother's second iterate buffer
  T* cd;       // the Chebyshev smoother's direction
};

struct Dims {
  int I, J, K;
  int64_t N, Nh;
  int SX, SY;  // halo strides: SX = I+2, SY = J+2
};

struct Timer {
  std::string name;
  int launches = 0;
  double total_ms = 0;
  double bytes = 0;
};

struct DeviceBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct HostPipe;  // pinned staging ring for the host-array entry points

}  // namespace sp

struct sp_ctx {
  sp_params prm;
  int device = 0;
  int num_cu = 256;  // compute units of the device (workgroup sizing)
  int desc_state = -1;      // has_e of the uploaded level table, -1: stale
  bool fuse_dots = false;  // the finest post-smoother also forms (r,z), (z,q), (dw,z)
  uint32_t dots_nb = 0;    // workgroups of that post-smoother (its partials)
  hipStream_t stream = nullptr;
  sp::Dims dm{};
  bool grid_set = false;
  std::string last_error;

  // grid (device)
  double *dxp = nullptr, *dyp = nullptr, *dzp = nullptr;
  double *dxu = nullptr, *dyv = nullptr, *dzw = nullptr;

  // staged inputs for the host-pointer entry points (device copies)
  std::vector<sp::DeviceBuf> stage;
  sp::HostPipe* pipe = nullptr;  // created on the first host-array call

  // assembled operator (FP64, interior layout)
  double* pocoef = nullptr;  // 6*N, faces W,S,B,E,N,T
  double *cx = nullptr, *cy = nullptr, *cz = nullptr, *ee = nullptr;
  uint8_t* active = nullptr;
  float* dw = nullptr;       // diagonal as float: weight of the null-space projection
  double* b = nullptr;
  bool has_e = false;

  // CG vectors (FP64)
  double *x = nullptr, *r = nullptr, *q = nullptr, *p0 = nullptr, *p1 = nullptr, *z64 = nullptr;
  // iterative refinement of non-singular systems: pressure = x + xd, xe the
  // increment being solved for, p1 its right-hand side
  double *xd = nullptr, *xe = nullptr;
  bool xd_active = false;
  float *r32 = nullptr, *z32 = nullptr;

  // multigrid hierarchies (only one is populated, by precision)
  int levels = 0;
  int tail_level = 0;
  sp::Level<float> lf[sp::kMaxLevels];
  sp::Level<double> ld[sp::kMaxLevels];
  void* mg_desc = nullptr;  // device copy of the level table for the tail kernel

  // GFM stale-beta machinery
  int32_t* stale_list = nullptr;  // (cell, face, kind) triples
  int64_t stale_cap = 0;
  int64_t* col_last = nullptr;    // 12 * I*J loop positions
  int64_t* col_prefix = nullptr;  // 12 * I*J
  uint8_t* gfm_cls = nullptr;     // GFM class of every interior cell (halo layout)

  // two-level reductions
  double* red_part = nullptr;
  uint32_t red_max_blocks = 0;
  int32_t* hflags_dev = nullptr;  // device view of hflags

  // scalars + partials
  sp::Scalars* S = nullptr;
  sp::Scalars* S_host = nullptr;  // pinned
  double* partials = nullptr;

  // adjoint checkpoint
  bool have_forward = false;
  double* x_ckpt = nullptr;       // forward pressure (interior, gauge applied)
  double* dp_save[3] = {nullptr, nullptr, nullptr};  // Dp on E,N,T faces (cell arrays)
  const void* fwd_inputs_tag = nullptr;
  // inputs of the last forward kept on device for the adjoint (staged copies
  // or the caller's device pointers)
  sp_inputs last_in{};
  bool last_in_host = false;
  bool geom_staged = false;  // sp_project's geometry staging slots hold the last host geometry

  // adjoint work
  double *pbar = nullptr, *lam = nullptr;

  // bc values staged
  double* bcv[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

  // wall step: face values of u, v, w, p (4 blocks of the six boundary planes)
  double* wall_var = nullptr;
  bool last_wall = false;  // the last forward applied the wall step

  // profiling
  bool profiling = false;
  std::string prof_filter;  // time only launches of this kernel (empty: all)
  std::vector<sp::Timer> timers;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  std::vector<std::pair<int, int>> ev_used;  // (timer index, event index)
  double bytes_scale = 1.0;

  size_t allocated_for = 0;  // N the buffers were sized for
  int allocated_prec = -1;
  int allocated_precond = -1;
  int allocated_levels = -1;
  int allocated_cheb = -1;  // Chebyshev step buffers allocated (flag at allocation)
  std::vector<void*> allocs;    // every device allocation of the buffers above
  int32_t* hflags = nullptr;    // pinned convergence flags (2 slots)
};

namespace sp {

struct RedSite;
// ---- host helpers (sp_api.hip) ----
sp_status fail(sp_ctx* c, sp_status s, const std::string& msg);
sp_status hip_check(sp_ctx* c, hipError_t e, const char* what);
void timer_begin(sp_ctx* c, const char* name, double bytes, int* slot);
void timer_end(sp_ctx* c, int slot);

// ---- host <-> device copies of pageable arrays (sp_hostio.hip) ----
sp_status pipe_upload(sp_ctx* c, void* dst, const void* src, size_t bytes);
// synchronous: dst holds the data on return
sp_status pipe_download(sp_ctx* c, void* dst, const void* src, size_t bytes);
// Download of a halo array (0:I+1, 0:J+1, 0:K+1) of doubles that writes only
// its entries (1:I, 1:J, 1:K) into dst.
sp_status pipe_download_interior(sp_ctx* c, double* dst, const double* src);
void pipe_destroy(sp_ctx* c);

// ---- assembly (sp_assemble.hip) ----
sp_status assemble(sp_ctx* c, const sp_inputs& in, int64_t* stale_count);

// ---- solver (sp_solver.hip) ----
sp_status mg_setup(sp_ctx* c);
sp_status solve(sp_ctx* c, const double* rhs, double tol, sp_stats* st, bool check_true,
                bool resume);
sp_status refine(sp_ctx* c, double tol, double* rmax, int* iters);
sp_status fold_refinement(sp_ctx* c);

// ---- correction (sp_correct.hip) ----
sp_status correct(sp_ctx* c, const sp_inputs& in, sp_outputs& out, sp_stats* st);

// ---- wall step after the correction (sp_wall.hip) ----
// Updates out.u, v, w, p in place; `host_var`: out.bc_var are host pointers.
sp_status wall_step(sp_ctx* c, const sp_inputs& in, const sp_outputs& out, bool host_var);

// GridPreprocess on the four device grids P, U, V, W (sp_cutcell.hip).
sp_status grid_preprocess(sp_ctx* c, const sp_cell_geom* const g[4]);
// DefineMomentumExchangeCell and NumberExternalCell (sp_celllinks.hip);
// cnt holds number_scratch(Imax-iu, Jmax-iv, Kmax-iw) int32, ext one int32
// (device memory).
sp_status define_moex(sp_ctx* c, const int32_t* ctp, const int32_t* ctu, const int32_t* ctv,
                      const int32_t* ctw, int32_t* mu, int32_t* mv, int32_t* mw);
sp_status number_external(sp_ctx* c, int iu, int iv, int iw, const int32_t* ct, int32_t* pos,
                          int32_t* cnt, int32_t* ext, bool timed = true);
size_t number_scratch(int I2, int J2, int K2);
// The allocation-free entry points' static device scratch (sp_noalloc.hip):
// its address on `device` (the current device) is looked up once per device
// by sp_create.
hipError_t noalloc_init(int device);
// NewCellFace (sp_newcellface.hip); scratch holds 9 N doubles.
sp_status new_cell_face(sp_ctx* c, const sp_cell_geom* const g[4],
                        const sp_cell_interp* const x[3], double* scratch);
// ComputeUVWLiquidField (sp_liquid.hip) on device arrays P, U, V, W.
sp_status uvw_liquid_field(sp_ctx* c, const sp_liquid_field* const g[4]);
// interpolateLvS (sp_clsvof.hip) of a solid body onto one grid
sp_status interpolate_lvs(sp_ctx* c, const sp_solid_body& b, const sp_lvs_grid& g, bool device);
// Clsvof_Scheme (sp_clsvof.hip) on device arrays; work holds
// clsvof_work_doubles(D) doubles.
sp_status clsvof_scheme(sp_ctx* c, const sp_clsvof_inputs& in, const sp_clsvof_state& st,
                        double* work);
size_t clsvof_work_doubles(const Dims& D);
// Momentum predictor (sp_predict.hip) on device arrays; faces holds
// predict_face_doubles(D, second) doubles (second: SpaceOrder 2 after the
// first step, which also needs the delh carries).
sp_status predict_uvw(sp_ctx* c, const sp_pred_inputs& in, const sp_pred_outputs& out,
                      double* faces);
size_t predict_face_doubles(const Dims& D, bool second);
// doubles of the face fluxes alone (the start of that buffer)
size_t predict_flux_doubles(const Dims& D);
// ---- Layout conversion at the entry points (sp_layout.hip) ----
// Interior of a halo array (Nh doubles) into a cell array (N) and back, on
// the context's stream.
sp_status halo_to_cell(sp_ctx* c, const double* halo, double* cell);
sp_status cell_to_halo(sp_ctx* c, const double* cell, double* halo);
// ---- VTR output (sp_vtr.hip) ----
sp_status write_vtr(sp_ctx* c, const char* dir, const char* prefix, int64_t itt,
                    const sp_vtr_fields& f, bool device);
sp_status write_vtk(sp_ctx* c, const char* dir, int64_t itt, const sp_vtk_fields& f,
                    bool device);

// ---- adjoint (sp_adjoint.hip) ----
sp_status adjoint(sp_ctx* c, const sp_adj_inputs& in, sp_adj_outputs& out, sp_stats* st);

}  // namespace sp
