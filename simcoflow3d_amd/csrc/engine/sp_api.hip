// C ABI of the projection engine (include/simco_proj.h): context, device
// buffers, the host- and device-pointer projection entry points, the adjoint
// and per-kernel timing.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "sp_internal.h"

namespace sp {


sp_status fail(sp_ctx* c, sp_status s, const std::string& msg) {
  if (c) c->last_error = msg;
  return s;
}

sp_status hip_check(sp_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return SP_OK;
  if (c) c->last_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? SP_ERR_OOM : SP_ERR_HIP;
}

void timer_begin(sp_ctx* c, const char* name, double bytes, int* slot) {
  *slot = -1;
  if (!c->profiling) return;
  if (!c->prof_filter.empty() && c->prof_filter != name) return;
  int ti = -1;
  for (size_t q = 0; q < c->timers.size(); ++q)
    if (c->timers[q].name == name) { ti = (int)q; break; }
  if (ti < 0) {
    c->timers.push_back(Timer{name, 0, 0.0, 0.0});
    ti = (int)c->timers.size() - 1;
  }
  c->timers[ti].launches += 1;
  c->timers[ti].bytes += bytes;
  const size_t ei = c->ev_used.size();
  if (ei >= c->ev_pool.size()) {
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    c->ev_pool.push_back({a, b});
  }
  hipEventRecord(c->ev_pool[ei].first, c->stream);
  c->ev_used.push_back({ti, (int)ei});
  *slot = (int)ei;
}

void timer_end(sp_ctx* c, int slot) {
  if (slot < 0) return;
  hipEventRecord(c->ev_pool[slot].second, c->stream);
}

static void timers_collect(sp_ctx* c) {
  if (!c->profiling) return;
  hipStreamSynchronize(c->stream);
  for (auto& u : c->ev_used) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev_pool[u.second].first, c->ev_pool[u.second].second);
    c->timers[u.first].total_ms += ms;
  }
  c->ev_used.clear();
}

// Coarsen while every dimension is at least 4 and the level holds more than
// 64 cells (256^3: 7 levels, down to 4^3; a level more, 2^3, costs tail
// time without saving CG iterations, while the 512x256x256 tank keeps its
// 4x2x2 level: stopping at 8x4x4 costs it an iteration).
static int auto_levels(int I, int J, int K) {
  int L = 1, nx = I, ny = J, nz = K;
  while (nx >= 4 && ny >= 4 && nz >= 4 && (int64_t)nx * ny * nz > 64 && L < kMaxLevels) {
    nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
    ++L;
  }
  return L;
}

template <typename T>
static sp_status dmalloc(sp_ctx* c, T** p, size_t count) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) { *p = nullptr; return hip_check(c, e, "hipMalloc"); }
  c->allocs.push_back(v);
  *p = static_cast<T*>(v);
  return SP_OK;
}

static void free_buffers(sp_ctx* c) {
  for (void* p : c->allocs) hipFree(p);
  c->allocs.clear();
  c->allocated_for = 0;
  c->allocated_prec = -1;
  c->allocated_precond = -1;
  c->allocated_levels = -1;
  c->allocated_cheb = -1;
  c->have_forward = false;
}

template <typename T>
static sp_status alloc_levels(sp_ctx* c, Level<T>* lv, bool level0_owned) {
  const Dims& D = c->dm;
  for (int l = 0; l < c->levels; ++l) {
    Level<T>& L = lv[l];
    if (l == 0) { L.nx = D.I; L.ny = D.J; L.nz = D.K; }
    else {
      L.nx = (lv[l - 1].nx + 1) / 2; L.ny = (lv[l - 1].ny + 1) / 2; L.nz = (lv[l - 1].nz + 1) / 2;
    }
    L.n = (int64_t)L.nx * L.ny * L.nz;
    sp_status s;
    if ((s = dmalloc(c, &L.xs, L.n))) return s;
    // Chebyshev step buffer only where that smoother is selected
    L.cd = nullptr;
    if ((c->prm.flags & SP_FLAG_CHEBYSHEV) && (s = dmalloc(c, &L.cd, L.n))) return s;
    if (l == 0 && !level0_owned) continue;
    if ((s = dmalloc(c, &L.cx, L.n)) || (s = dmalloc(c, &L.cy, L.n)) || (s = dmalloc(c, &L.cz, L.n)) ||
        (s = dmalloc(c, &L.e_store, L.n)) || (s = dmalloc(c, &L.b, L.n)) || (s = dmalloc(c, &L.x, L.n)))
      return s;
    L.e = L.e_store;
  }
  return SP_OK;
}

static sp_status ensure_buffers(sp_ctx* c) {
  const Dims& D = c->dm;
  const int prec = c->prm.precision, pc = c->prm.precond;
  int L = auto_levels(D.I, D.J, D.K);
  if (c->prm.mg_levels > 0) L = std::min(L, c->prm.mg_levels);
  if (pc != SP_PRECOND_MG) L = 1;
  const int cheb = (c->prm.flags & SP_FLAG_CHEBYSHEV) ? 1 : 0;
  if (c->allocated_for == (size_t)D.N && c->allocated_prec == prec && c->allocated_precond == pc &&
      c->allocated_levels == L && c->allocated_cheb == cheb)
    return SP_OK;
  free_buffers(c);
  c->levels = L;
  c->desc_state = -1;  // level table must be uploaded again
  const int64_t N = D.N;
  sp_status s;
  if ((s = dmalloc(c, &c->pocoef, 6 * N)) || (s = dmalloc(c, &c->cx, N)) || (s = dmalloc(c, &c->cy, N)) ||
      (s = dmalloc(c, &c->cz, N)) || (s = dmalloc(c, &c->ee, N)) || (s = dmalloc(c, &c->b, N)) ||
      (s = dmalloc(c, &c->active, N)) || (s = dmalloc(c, &c->dw, N)) || (s = dmalloc(c, &c->x, N)) || (s = dmalloc(c, &c->r, N)) ||
      (s = dmalloc(c, &c->q, N)) || (s = dmalloc(c, &c->p0, N)) || (s = dmalloc(c, &c->p1, N)) ||
      (s = dmalloc(c, &c->pbar, N)) || (s = dmalloc(c, &c->x_ckpt, N)) ||
      (s = dmalloc(c, &c->xd, N)) || (s = dmalloc(c, &c->xe, N)))
    return s;
  const bool mixed_mg = prec == SP_PREC_MIXED && pc == SP_PRECOND_MG;
  if (mixed_mg) {
    if ((s = dmalloc(c, &c->r32, N)) || (s = dmalloc(c, &c->z32, N))) return s;
    if ((s = alloc_levels<float>(c, c->lf, true))) return s;
    // level 0 works on r32 -> z32
    c->lf[0].b = c->r32;
    c->lf[0].x = c->z32;
  } else {
    if ((s = dmalloc(c, &c->z64, N))) return s;
    c->r32 = nullptr;
    c->z32 = nullptr;
    if ((s = alloc_levels<double>(c, c->ld, false))) return s;
    Level<double>& L0 = c->ld[0];
    L0.cx = c->cx; L0.cy = c->cy; L0.cz = c->cz; L0.e_store = c->ee; L0.e = c->ee;
    L0.b = c->r; L0.x = c->z64;
  }
  // tail: the levels from which every coarser level fits the tail kernel's
  // LDS (seven arrays per level), each finer than the coarsest holding at
  // most kTailCells cells.  c->levels when not even the coarsest level fits
  // (a hierarchy cut short by mg_levels, or a large grid that cannot
  // coarsen): the V-cycle then smooths the coarsest level in global memory.
  {
    const int64_t esz = mixed_mg ? 4 : 8;
    c->tail_level = c->levels;
    int64_t bytes = 0;
    for (int l = c->levels - 1; l >= 0; --l) {
      const int64_t n = mixed_mg ? c->lf[l].n : c->ld[l].n;
      bytes += 7 * esz * n;
      if (bytes > kTailLdsBytes || (l < c->levels - 1 && n > kTailCells)) break;
      c->tail_level = l;
    }
  }
  if ((s = dmalloc(c, reinterpret_cast<char**>(&c->mg_desc), 128 * kMaxLevels))) return s;
  c->stale_cap = std::min<int64_t>(6 * N, std::max<int64_t>(1 << 20, N / 2));
  if ((s = dmalloc(c, &c->stale_list, 3 * c->stale_cap))) return s;
  if ((s = dmalloc(c, &c->col_last, 12 * (int64_t)D.I * D.J)) ||
      (s = dmalloc(c, &c->col_prefix, 12 * (int64_t)D.I * D.J)) ||
      (s = dmalloc(c, &c->gfm_cls, (int64_t)D.Nh)))
    return s;
  c->red_max_blocks = (uint32_t)std::max<int64_t>(
      4096, (int64_t)((D.I + 2 + 63) / 64) * ((D.J + 2 + 3) / 4) * (D.K + 2));
  if ((s = dmalloc(c, &c->red_part, (size_t)kNumSites * kSiteValues * c->red_max_blocks))) return s;
  for (int f = 0; f < 6; ++f) {
    const int64_t m = std::max<int64_t>({(int64_t)D.J * D.K, (int64_t)D.I * D.K, (int64_t)D.I * D.J});
    if ((s = dmalloc(c, &c->bcv[f], m))) return s;
  }
  // face values of the wall step: u, v, w, p x the six boundary planes
  if ((s = dmalloc(c, &c->wall_var,
                   8 * ((int64_t)D.J * D.K + (int64_t)D.I * D.K + (int64_t)D.I * D.J))))
    return s;
  c->allocated_for = (size_t)N;
  c->allocated_prec = prec;
  c->allocated_precond = pc;
  c->allocated_levels = L;
  c->allocated_cheb = cheb;
  return SP_OK;
}

static bool check_params(const sp_params* p, std::string& why) {
  if (p->imax < 2 || p->jmax < 2 || p->kmax < 2) { why = "imax, jmax, kmax must be >= 2"; return false; }
  // the smoother kernels address level arrays with 32-bit byte offsets, and
  // the FP64 ones reach 8 bytes per cell
  if ((int64_t)p->imax * p->jmax * p->kmax >= ((int64_t)1 << 29)) {
    why = "grid too large (at most 2^29 - 1 cells)";
    return false;
  }
  if (p->precision != SP_PREC_FP64 && p->precision != SP_PREC_MIXED) { why = "precision"; return false; }
  if (p->precond != SP_PRECOND_MG && p->precond != SP_PRECOND_JACOBI) { why = "precond"; return false; }
  if (p->mg_pre < 1 || p->mg_post < 0 || p->mg_coarse_sweeps < 1) { why = "sweeps"; return false; }
  if (p->max_iter < 1) { why = "max_iter"; return false; }
  if (!(p->div_tol >= 0.0)) { why = "div_tol"; return false; }
  if (!(p->rho_l > 0) || !(p->rho_g > 0) || !(p->rho_ref > 0)) { why = "densities"; return false; }
  if (!(p->mg_coarse_scale > 0)) { why = "mg_coarse_scale"; return false; }
  if (!(p->mg_omega > 0.0 && p->mg_omega < 2.0) ||
      !(p->mg_omega_coarse > 0.0 && p->mg_omega_coarse < 2.0)) {
    why = "mg_omega and mg_omega_coarse must be in (0, 2)";
    return false;
  }
  return true;
}

static bool check_wall(const sp_wall_bc* w) {
  for (int f = 0; f < 6; ++f)
    for (int32_t v : {w->flag_u[f], w->flag_v[f], w->flag_w[f]})
      if (v != 0 && v != 1) return false;
  return true;
}

static sp_status check_inputs(sp_ctx* c, const sp_inputs* in, const sp_outputs* out) {
  if (!in || !out) return fail(c, SP_ERR_INVALID_ARG, "null inputs/outputs");
  const void* req[] = {in->u_star, in->v_star, in->w_star, in->dp_u, in->dp_v, in->dp_w,
                       in->ee_area, in->ne_area, in->te_area, in->vof, in->vof_old,
                       in->cell_type_p, in->cell_type_u, in->cell_type_v, in->cell_type_w,
                       in->moex_u, in->moex_v, in->moex_w, out->p, out->u, out->v, out->w};
  for (const void* p : req)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "required input/output pointer is null");
  const bool beta = in->beta_e || in->beta_n || in->beta_t;
  if (beta && !(in->beta_e && in->beta_n && in->beta_t))
    return fail(c, SP_ERR_INVALID_ARG, "beta_e, beta_n, beta_t must be given together");
  if (!beta && (!in->vof_l || !in->phi_l))
    return fail(c, SP_ERR_INVALID_ARG, "vof_l and phi_l are required for the GFM coefficients");
  for (int f = 0; f < 6; ++f) {
    if (in->bc_flag[f] != 0 && in->bc_flag[f] != 1)
      return fail(c, SP_ERR_INVALID_ARG, "bc_flag must be 0 (Dirichlet) or 1 (Neumann)");
    if (in->bc_flag[f] == 0 && !in->bc_val[f])
      return fail(c, SP_ERR_INVALID_ARG, "bc_val required for a Dirichlet face");
  }
  if (!(in->dt > 0.0) || !std::isfinite(in->dt)) return fail(c, SP_ERR_INVALID_ARG, "dt must be > 0");
  if (!c->grid_set) return fail(c, SP_ERR_GRID, "sp_set_grid has not been called");
  if (in->wall && !check_wall(in->wall))
    return fail(c, SP_ERR_INVALID_ARG, "sp_wall_bc flags must be 0 (Dirichlet) or 1 (Neumann)");
  return SP_OK;
}

static size_t bc_len(const Dims& D, int f) {
  return f < 2 ? (size_t)D.J * D.K : (f < 4 ? (size_t)D.I * D.K : (size_t)D.I * D.J);
}

// Forward on device arrays.  `host_bc`: bc_val pointers are host memory.
static sp_status project_device(sp_ctx* c, const sp_inputs* in, sp_outputs* out, sp_stats* st,
                                bool host_bc) {
  auto t0 = std::chrono::steady_clock::now();
  sp_status s = ensure_buffers(c);
  if (s) return s;
  const Dims& D = c->dm;
  hipError_t e = hipMemsetAsync(c->S, 0, sizeof(Scalars), c->stream);
  if (e != hipSuccess) return hip_check(c, e, "hipMemsetAsync");
  for (int f = 0; f < 6; ++f) {
    if (in->bc_flag[f] != 0) continue;
    e = hipMemcpyAsync(c->bcv[f], in->bc_val[f], sizeof(double) * bc_len(D, f),
                       host_bc ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(bc)");
  }
  c->have_forward = false;
  c->xd_active = false;
  int64_t stale = 0;
  if ((s = assemble(c, *in, &stale))) return s;
  e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "assembly");
  auto t1 = std::chrono::steady_clock::now();
  stale = c->S_host->stale_count;
  if (stale > c->stale_cap)
    return fail(c, SP_ERR_CAPACITY,
                "too many GFM faces next to unclassified (solid) cells for the stale-face "
                "list (" + std::to_string(stale) + " > " + std::to_string(c->stale_cap) + ")");
  if (c->S_host->nonfinite) return fail(c, SP_ERR_NONFINITE, "non-finite value in the assembled system");
  c->has_e = c->S_host->esum > 0.0;
  if ((s = mg_setup(c))) return s;
  sp_stats local{};
  sp_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof(*S));
  S->initial_div = c->S_host->bmax;
  S->singular = c->S_host->singular;
  S->active_cells = c->S_host->active;
  S->stale_faces = stale;
  s = solve(c, c->b, c->prm.div_tol, S, false, false);
  if (s != SP_OK && s != SP_ERR_NOT_CONVERGED) return s;
  auto t2 = std::chrono::steady_clock::now();
  // Correction and divergence of the corrected field.  With Neumann walls,
  // no volume source and no masked open faces the divergence is the true
  // residual of the solve, so a drift of the recursive residual shows here:
  // the solve then resumes from the true residual (at most 3 times).  If the
  // true residual is already below the tolerance, the solve has converged
  // and the remaining divergence comes from the inputs (Dirichlet pressure
  // faces, vof - vofO, masked faces), which the reference shows as well.
  bool res_converged = false;
  for (int restart = 0;; ++restart) {
    sp_status sc = correct(c, *in, *out, S);
    if (sc) return sc;
    e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "correction");
    if (c->S_host->divmax <= c->prm.div_tol || c->S_host->nonfinite || restart >= 3 ||
        c->S_host->iter >= c->prm.max_iter)
      break;
    if (!c->S_host->singular) {
      // Non-singular systems (a Dirichlet face fixes the pressure level):
      // refine the pressure as x + xd, so that the corrected velocity is
      // not limited by the rounding of that level
      double rmax = 0.0;
      int it_ref = 0;
      const int32_t it_before = S->iterations;
      s = refine(c, c->prm.div_tol, &rmax, &it_ref);
      if (s != SP_OK && s != SP_ERR_NOT_CONVERGED) return s;
      S->iterations = it_before + it_ref;
      if (!(rmax > c->prm.div_tol)) {
        res_converged = true;  // the residual is below tol: the inputs carry the rest
        break;
      }
      continue;
    }
    const int32_t it_before = c->S_host->iter;
    s = solve(c, c->b, c->prm.div_tol, S, false, true);
    if (s != SP_OK && s != SP_ERR_NOT_CONVERGED) return s;
    if (c->S_host->done && c->S_host->iter == it_before) {
      res_converged = true;  // x unchanged: the corrected field stands
      break;
    }
  }
  if (out->po_coef) {
    e = hipMemcpyAsync(out->po_coef, c->pocoef, sizeof(double) * 6 * D.N, hipMemcpyDeviceToDevice,
                       c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(po_coef)");
  }
  // the wall step of UpdatePUVW on the corrected field (the convergence
  // test above stays on the projection's own output)
  if (in->wall && (s = wall_step(c, *in, *out, host_bc))) return s;
  // checkpoint of the forward pressure for the adjoint: the iterate buffer
  // (with a refinement folded in) becomes the checkpoint and the old
  // checkpoint the next iterate
  if (c->xd_active && (s = fold_refinement(c))) return s;
  std::swap(c->x, c->x_ckpt);
  e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "correction");
  auto t3 = std::chrono::steady_clock::now();
  S->final_div = c->S_host->divmax;
  S->wall_div = in->wall ? c->S_host->walldiv : 0.0;
  S->converged = !c->S_host->nonfinite && (S->final_div <= c->prm.div_tol || res_converged);
  S->t_assemble_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  S->t_solve_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  S->t_correct_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
  S->t_total_ms = std::chrono::duration<double, std::milli>(t3 - t0).count();
  c->last_in = *in;
  c->last_wall = in->wall != nullptr;
  c->have_forward = true;
  timers_collect(c);
  return S->converged ? SP_OK : SP_ERR_NOT_CONVERGED;
}

// Device staging of host arrays
static sp_status stage(sp_ctx* c, size_t idx, size_t bytes, void** out) {
  if (c->stage.size() <= idx) c->stage.resize(idx + 1);
  DeviceBuf& b = c->stage[idx];
  if (b.bytes < bytes) {
    if (b.ptr) hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) return hip_check(c, e, "hipMalloc(stage)");
    b.bytes = bytes;
  }
  *out = b.ptr;
  return SP_OK;
}

template <typename T>
static sp_status upload(sp_ctx* c, size_t idx, const T* src, size_t count, const T** dst) {
  if (!src) { *dst = nullptr; return SP_OK; }
  void* d;
  sp_status s = stage(c, idx, count * sizeof(T), &d);
  if (s) return s;
  if ((s = pipe_upload(c, d, src, count * sizeof(T)))) return s;
  *dst = static_cast<const T*>(d);
  return SP_OK;
}

}  // namespace sp

using namespace sp;

extern "C" {

void sp_params_default(sp_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->precision = SP_PREC_FP64;
  p->precond = SP_PRECOND_MG;
  p->mg_levels = 0;
  p->mg_pre = 2;
  p->mg_post = 2;
  // two forward and two backward red-black sweep pairs on the coarsest level
  // (a few cells): one pair costs the 512x256x256 tank a CG iteration, more
  // than two cost tail time without saving any
  p->mg_coarse_sweeps = 2;
  p->mg_coarse_scale = 0.5;
  p->mg_omega = 1.25;
  p->mg_omega_coarse = 1.5;
  p->max_iter = 200;
  p->div_tol = 1.0e-10;
  p->rho_l = 998.3;   // row, StateVariables.f90:10
  p->rho_g = 1.205;   // roa
  p->rho_ref = 998.3; // Roref = row (ProjectionP.f90:53)
  p->epsi = 1.0e-3;   // CutCell.f90:11
  p->pin_i = p->pin_j = p->pin_k = 0;
}

sp_status sp_create(sp_ctx** out, const sp_params* params, int device) {
  if (!out || !params) return SP_ERR_INVALID_ARG;
  *out = nullptr;
  std::string why;
  if (!check_params(params, why)) return SP_ERR_INVALID_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev)
    return SP_ERR_NO_DEVICE;
  sp_ctx* c = new (std::nothrow) sp_ctx();
  if (!c) return SP_ERR_OOM;
  c->prm = *params;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SP_ERR_HIP;
  }
  // the allocation-free entry points' scratch address on this device
  if (noalloc_init(device) != hipSuccess) {
    hipStreamDestroy(c->stream);
    delete c;
    return SP_ERR_HIP;
  }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      ncu > 0)
    c->num_cu = ncu;
  Dims& D = c->dm;
  D.I = params->imax; D.J = params->jmax; D.K = params->kmax;
  D.N = (int64_t)D.I * D.J * D.K;
  D.SX = D.I + 2; D.SY = D.J + 2;
  D.Nh = (int64_t)D.SX * D.SY * (D.K + 2);
  if (hipMalloc(&c->S, sizeof(Scalars)) != hipSuccess ||
      hipHostMalloc(&c->S_host, sizeof(Scalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->hflags, 4 * sizeof(int32_t), hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c->hflags_dev, c->hflags, 0) != hipSuccess ||
      hipMalloc(&c->dxp, sizeof(double) * D.I) != hipSuccess ||
      hipMalloc(&c->dyp, sizeof(double) * D.J) != hipSuccess ||
      hipMalloc(&c->dzp, sizeof(double) * D.K) != hipSuccess ||
      hipMalloc(&c->dxu, sizeof(double) * D.I) != hipSuccess ||
      hipMalloc(&c->dyv, sizeof(double) * D.J) != hipSuccess ||
      hipMalloc(&c->dzw, sizeof(double) * D.K) != hipSuccess) {
    sp_destroy(c);
    return SP_ERR_OOM;
  }
  std::memset(c->S_host, 0, sizeof(Scalars));
  *out = c;
  return SP_OK;
}

sp_status sp_destroy(sp_ctx* c) {
  if (!c) return SP_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_buffers(c);
  pipe_destroy(c);
  for (auto& s : c->stage) if (s.ptr) hipFree(s.ptr);
  c->stage.clear();
  for (auto& p : c->ev_pool) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  if (c->S) hipFree(c->S);
  if (c->S_host) hipHostFree(c->S_host);
  if (c->hflags) hipHostFree(c->hflags);
  for (double* p : {c->dxp, c->dyp, c->dzp, c->dxu, c->dyv, c->dzw}) if (p) hipFree(p);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SP_OK;
}

sp_status sp_set_params(sp_ctx* c, const sp_params* p) {
  if (!c || !p) return SP_ERR_INVALID_ARG;
  std::string why;
  if (!check_params(p, why)) return fail(c, SP_ERR_INVALID_ARG, why);
  if (p->imax != c->dm.I || p->jmax != c->dm.J || p->kmax != c->dm.K)
    return fail(c, SP_ERR_INVALID_ARG, "sp_set_params cannot change the grid size");
  c->prm = *p;
  c->have_forward = false;
  return SP_OK;
}

sp_status sp_set_grid_1d(sp_ctx* c, const double* dxp, const double* dyp, const double* dzp,
                         const double* dxu, const double* dyv, const double* dzw) {
  if (!c || !dxp || !dyp || !dzp || !dxu || !dyv || !dzw) return SP_ERR_INVALID_ARG;
  const Dims& D = c->dm;
  const double* in[6] = {dxp, dyp, dzp, dxu, dyv, dzw};
  const int len[6] = {D.I, D.J, D.K, D.I, D.J, D.K};
  for (int q = 0; q < 6; ++q)
    for (int i = 0; i < len[q]; ++i)
      if (!(in[q][i] > 0.0) || !std::isfinite(in[q][i]))
        return fail(c, SP_ERR_GRID, "grid spacings must be finite and > 0");
  hipSetDevice(c->device);
  double* dst[6] = {c->dxp, c->dyp, c->dzp, c->dxu, c->dyv, c->dzw};
  for (int q = 0; q < 6; ++q) {
    hipError_t e = hipMemcpy(dst[q], in[q], sizeof(double) * len[q], hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpy(grid)");
  }
  c->grid_set = true;
  c->have_forward = false;
  return SP_OK;
}

sp_status sp_set_grid(sp_ctx* c, const double* dxp, const double* dyp, const double* dzp,
                      const double* dxu, const double* dyv, const double* dzw) {
  if (!c || !dxp || !dyp || !dzp || !dxu || !dyv || !dzw) return SP_ERR_INVALID_ARG;
  const Dims& D = c->dm;
  std::vector<double> x(D.I), y(D.J), z(D.K), xu(D.I), yv(D.J), zw(D.K);
  auto at = [&](const double* a, int i, int j, int k) {
    return a[i + (int64_t)D.I * (j + (int64_t)D.J * k)];
  };
  for (int i = 0; i < D.I; ++i) { x[i] = at(dxp, i, 0, 0); xu[i] = at(dxu, i, 0, 0); }
  for (int j = 0; j < D.J; ++j) { y[j] = at(dyp, 0, j, 0); yv[j] = at(dyv, 0, j, 0); }
  for (int k = 0; k < D.K; ++k) { z[k] = at(dzp, 0, 0, k); zw[k] = at(dzw, 0, 0, k); }
  for (int k = 0; k < D.K; ++k)
    for (int j = 0; j < D.J; ++j)
      for (int i = 0; i < D.I; ++i)
        if (at(dxp, i, j, k) != x[i] || at(dxu, i, j, k) != xu[i] || at(dyp, i, j, k) != y[j] ||
            at(dyv, i, j, k) != yv[j] || at(dzp, i, j, k) != z[k] || at(dzw, i, j, k) != zw[k])
          return fail(c, SP_ERR_GRID, "grid is not a tensor product");
  return sp_set_grid_1d(c, x.data(), y.data(), z.data(), xu.data(), yv.data(), zw.data());
}

sp_status sp_project_dev(sp_ctx* c, const sp_inputs* in, sp_outputs* out, sp_stats* st) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  sp_status s = check_inputs(c, in, out);
  if (s) return s;
  c->last_in_host = false;
  return project_device(c, in, out, st, false);
}

sp_status sp_project(sp_ctx* c, const sp_inputs* in, sp_outputs* out, sp_stats* st) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  sp_status s = check_inputs(c, in, out);
  if (s) return s;
  if ((s = ensure_buffers(c))) return s;
  const Dims& D = c->dm;
  const size_t Nh = (size_t)D.Nh, N = (size_t)D.N;
  sp_inputs d = *in;
  size_t k = 0;
  const double** hp[] = {&d.u_star, &d.v_star, &d.w_star, &d.dp_u, &d.dp_v, &d.dp_w,
                         &d.vof_l, &d.phi_l};
  for (const double** p : hp)
    if ((s = upload(c, k++, *p, Nh, p))) return s;
  // The geometry (apertures, vof, vofO, cell types, MoExCell) keeps staging
  // slots of its own, used by no other entry point, so a caller whose solid
  // has not moved can leave it on the device (sp_inputs.static_geometry).
  const double** gh[] = {&d.ee_area, &d.ne_area, &d.te_area, &d.vof, &d.vof_old};
  const int32_t** gc[] = {&d.cell_type_p, &d.cell_type_u, &d.cell_type_v, &d.cell_type_w,
                          &d.moex_u, &d.moex_v, &d.moex_w};
  const size_t kGeo = 400;  // no other entry point stages at or above 400
  if (in->static_geometry && c->geom_staged) {
    for (size_t q = 0; q < 5; ++q) *gh[q] = static_cast<const double*>(c->stage[kGeo + q].ptr);
    for (size_t q = 0; q < 7; ++q) *gc[q] = static_cast<const int32_t*>(c->stage[kGeo + 5 + q].ptr);
  } else {
    c->geom_staged = false;
    for (size_t q = 0; q < 5; ++q)
      if ((s = upload(c, kGeo + q, *gh[q], Nh, gh[q]))) return s;
    for (size_t q = 0; q < 7; ++q)
      if ((s = upload(c, kGeo + 5 + q, *gc[q], N, gc[q]))) return s;
    c->geom_staged = true;
  }
  const double** bp[] = {&d.beta_e, &d.beta_n, &d.beta_t};
  for (const double** p : bp)
    if ((s = upload(c, k++, *p, N, p))) return s;
  // Outputs.  Without the wall step only the entries (1:I,1:J,1:K) of u, v,
  // w are written (and only those are copied back, so the caller's other
  // entries are kept); the wall step also writes boundary entries, so then
  // the caller's u, v, w are uploaded first and copied back whole.
  const bool whole = in->wall != nullptr;
  sp_outputs o{};
  double* outs_h[4] = {out->p, out->u, out->v, out->w};
  double* outs_d[4];
  for (int q = 0; q < 4; ++q) {
    void* p;
    if ((s = stage(c, k++, Nh * sizeof(double), &p))) return s;
    outs_d[q] = static_cast<double*>(p);
    if (whole && q > 0 && (s = pipe_upload(c, p, outs_h[q], Nh * sizeof(double)))) return s;
  }
  o.p = outs_d[0]; o.u = outs_d[1]; o.v = outs_d[2]; o.w = outs_d[3];
  o.po_coef = nullptr;
  for (int q = 0; q < 24; ++q) o.bc_var[q] = out->bc_var[q];  // host memory
  c->last_in_host = true;
  sp_status r = project_device(c, &d, &o, st, true);
  if (r != SP_OK && r != SP_ERR_NOT_CONVERGED) return r;
  for (int q = 0; q < 4; ++q) {
    s = (whole || q == 0) ? pipe_download(c, outs_h[q], outs_d[q], Nh * sizeof(double))
                          : pipe_download_interior(c, outs_h[q], outs_d[q]);
    if (s) return s;
  }
  if (out->po_coef && (s = pipe_download(c, out->po_coef, c->pocoef, 6 * N * sizeof(double))))
    return s;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_project");
  return r;
}

static sp_status adjoint_common(sp_ctx* c, const sp_adj_inputs* in, sp_adj_outputs* out,
                                sp_stats* st) {
  sp_stats local{};
  sp_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof(*S));
  auto t0 = std::chrono::steady_clock::now();
  hipError_t e0 = hipMemsetAsync(&c->S->iter, 0, sizeof(int32_t), c->stream);
  if (e0 == hipSuccess) e0 = hipMemsetAsync(&c->S->nonfinite, 0, sizeof(int32_t), c->stream);
  if (e0 != hipSuccess) return hip_check(c, e0, "hipMemsetAsync");
  sp_status s = adjoint(c, *in, *out, S);
  if (s != SP_OK && s != SP_ERR_NOT_CONVERGED) return s;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "adjoint");
  S->t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  S->singular = c->S_host->singular;
  S->active_cells = c->S_host->active;
  timers_collect(c);
  // the forward's x has been reused as the adjoint variable; the checkpoint
  // keeps the pressure, so further adjoint calls stay valid
  return S->converged ? SP_OK : SP_ERR_NOT_CONVERGED;
}

sp_status sp_adjoint_dev(sp_ctx* c, const sp_adj_inputs* in, sp_adj_outputs* out, sp_stats* st) {
  if (!c || !in || !out) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!c->have_forward) return fail(c, SP_ERR_STATE, "no forward projection to differentiate");
  if (c->last_wall)
    return fail(c, SP_ERR_STATE, "the adjoint covers the projection without the wall step");
  if (!in->gu || !in->gv || !in->gw || !out->gu_star || !out->gv_star || !out->gw_star ||
      !out->gk_e || !out->gk_n || !out->gk_t)
    return fail(c, SP_ERR_INVALID_ARG, "null adjoint pointer");
  return adjoint_common(c, in, out, st);
}

sp_status sp_adjoint(sp_ctx* c, const sp_adj_inputs* in, sp_adj_outputs* out, sp_stats* st) {
  if (!c || !in || !out) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!c->have_forward) return fail(c, SP_ERR_STATE, "no forward projection to differentiate");
  if (!c->last_in_host)
    return fail(c, SP_ERR_STATE, "forward ran on device arrays; use sp_adjoint_dev");
  if (c->last_wall)
    return fail(c, SP_ERR_STATE, "the adjoint covers the projection without the wall step");
  if (!in->gu || !in->gv || !in->gw || !out->gu_star || !out->gv_star || !out->gw_star ||
      !out->gk_e || !out->gk_n || !out->gk_t)
    return fail(c, SP_ERR_INVALID_ARG, "null adjoint pointer");
  const Dims& D = c->dm;
  const size_t Nh = (size_t)D.Nh, N = (size_t)D.N;
  sp_status s;
  size_t k = 40;  // staging slots after the forward's
  sp_adj_inputs di{};
  if ((s = upload(c, k++, in->gu, Nh, &di.gu)) || (s = upload(c, k++, in->gv, Nh, &di.gv)) ||
      (s = upload(c, k++, in->gw, Nh, &di.gw)))
    return s;
  // slots 43-51 hold the outputs below, 52 the pressure cotangent
  if (in->gp && (s = upload(c, 52, in->gp, Nh, &di.gp))) return s;
  sp_adj_outputs dout{};
  double* hs[9] = {out->gu_star, out->gv_star, out->gw_star, out->gk_e, out->gk_n, out->gk_t,
                   out->gbeta_e, out->gbeta_n, out->gbeta_t};
  double* ds[9];
  for (int q = 0; q < 9; ++q) {
    ds[q] = nullptr;
    if (!hs[q]) continue;
    void* p;
    if ((s = stage(c, k++, (q < 3 ? Nh : N) * sizeof(double), &p))) return s;
    ds[q] = static_cast<double*>(p);
  }
  dout.gu_star = ds[0]; dout.gv_star = ds[1]; dout.gw_star = ds[2];
  dout.gk_e = ds[3]; dout.gk_n = ds[4]; dout.gk_t = ds[5];
  dout.gbeta_e = ds[6]; dout.gbeta_n = ds[7]; dout.gbeta_t = ds[8];
  sp_status r = adjoint_common(c, &di, &dout, st);
  if (r != SP_OK && r != SP_ERR_NOT_CONVERGED) return r;
  for (int q = 0; q < 9; ++q) {
    if (!hs[q]) continue;
    if ((s = pipe_download(c, hs[q], ds[q], (q < 3 ? Nh : N) * sizeof(double)))) return s;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_adjoint");
  return r;
}

static sp_status check_wall_call(sp_ctx* c, const sp_inputs* in, const sp_outputs* out) {
  if (!in || !out || !in->wall) return fail(c, SP_ERR_INVALID_ARG, "null inputs/outputs/wall");
  const void* req[] = {in->ee_area, in->ne_area, in->te_area, in->cell_type_p, in->cell_type_u,
                       in->cell_type_v, in->cell_type_w, out->u, out->v, out->w, out->p};
  for (const void* p : req)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "required input/output pointer is null");
  for (int f = 0; f < 6; ++f)
    if (in->bc_flag[f] != 0 && in->bc_flag[f] != 1)
      return fail(c, SP_ERR_INVALID_ARG, "bc_flag must be 0 (Dirichlet) or 1 (Neumann)");
  if (!check_wall(in->wall))
    return fail(c, SP_ERR_INVALID_ARG, "sp_wall_bc flags must be 0 (Dirichlet) or 1 (Neumann)");
  if (!c->grid_set) return fail(c, SP_ERR_GRID, "sp_set_grid has not been called");
  return ensure_buffers(c);
}

static sp_status wall_finish(sp_ctx* c, sp_stats* st) {
  hipError_t e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "wall step");
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->wall_div = c->S_host->walldiv;
    st->converged = 1;
  }
  timers_collect(c);
  return SP_OK;
}

sp_status sp_wall_step_dev(sp_ctx* c, const sp_inputs* in, sp_outputs* out, sp_stats* st) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  sp_status s = check_wall_call(c, in, out);
  if (s) return s;
  if ((s = wall_step(c, *in, *out, false))) return s;
  return wall_finish(c, st);
}

sp_status sp_wall_step(sp_ctx* c, const sp_inputs* in, sp_outputs* out, sp_stats* st) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  sp_status s = check_wall_call(c, in, out);
  if (s) return s;
  const Dims& D = c->dm;
  const size_t Nh = (size_t)D.Nh, N = (size_t)D.N;
  // staging slots past the forward's and the adjoint's, so a later adjoint
  // still finds the forward's inputs
  size_t k = 60;
  sp_inputs d{};
  std::memcpy(d.bc_flag, in->bc_flag, sizeof(d.bc_flag));
  d.wall = in->wall;
  if ((s = upload(c, k++, in->ee_area, Nh, &d.ee_area)) ||
      (s = upload(c, k++, in->ne_area, Nh, &d.ne_area)) ||
      (s = upload(c, k++, in->te_area, Nh, &d.te_area)) ||
      (s = upload(c, k++, in->cell_type_p, N, &d.cell_type_p)) ||
      (s = upload(c, k++, in->cell_type_u, N, &d.cell_type_u)) ||
      (s = upload(c, k++, in->cell_type_v, N, &d.cell_type_v)) ||
      (s = upload(c, k++, in->cell_type_w, N, &d.cell_type_w)))
    return s;
  sp_outputs o{};
  double* hs[4] = {out->u, out->v, out->w, out->p};
  double* ds[4];
  for (int q = 0; q < 4; ++q) {
    const double* p;
    if ((s = upload(c, k++, static_cast<const double*>(hs[q]), Nh, &p))) return s;
    ds[q] = const_cast<double*>(p);
  }
  o.u = ds[0]; o.v = ds[1]; o.w = ds[2]; o.p = ds[3];
  for (int q = 0; q < 24; ++q) o.bc_var[q] = out->bc_var[q];  // host memory
  if ((s = wall_step(c, d, o, true))) return s;
  for (int q = 0; q < 4; ++q)
    if ((s = pipe_download(c, hs[q], ds[q], Nh * sizeof(double)))) return s;
  return wall_finish(c, st);
}

sp_status sp_write_vtr(sp_ctx* c, const char* dir, const char* prefix, int64_t itt,
                       const sp_vtr_fields* f, int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!dir || !prefix || !f) return fail(c, SP_ERR_INVALID_ARG, "null VTR argument");
  const sp_status s = write_vtr(c, dir, prefix, itt, *f, device != 0);
  timers_collect(c);
  return s;
}

sp_status sp_write_vtk(sp_ctx* c, const char* dir, int64_t itt, const sp_vtk_fields* f,
                       int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!dir || !f) return fail(c, SP_ERR_INVALID_ARG, "null VTK argument");
  const sp_status s = write_vtk(c, dir, itt, *f, device != 0);
  timers_collect(c);
  return s;
}

// The members of sp_cell_geom a host may allocate with the halo
// (halo_fields = 1), in struct order: phi, nx, ny, nz, vof, vof_l.  The
// kernels index cell arrays, so the entry points convert at the boundary.
constexpr int kGeomHalo = 6;
constexpr unsigned kGeomVof = 1u << 4, kGeomVofL = 1u << 5;

static const double* geom_halo_get(const sp_cell_geom& g, int q) {
  const double* const m[kGeomHalo] = {g.phi, g.nx, g.ny, g.nz, g.vof, g.vof_l};
  return m[q];
}
static void geom_halo_set(sp_cell_geom& g, int q, double* p) {
  switch (q) {
    case 0: g.phi = p; break;
    case 1: g.nx = p; break;
    case 2: g.ny = p; break;
    case 3: g.nz = p; break;
    case 4: g.vof = p; break;
    default: g.vof_l = p; break;
  }
}
// Host arrays of the halo members are uploaded whole (Nh values).
static size_t geom_count(const sp_cell_geom& g, size_t N, size_t Nh) {
  return g.halo_fields ? Nh : N;
}
// Device pointers: the halo members of g replaced by cell-layout copies in
// staging slots slot .. slot + 5.
static sp_status geom_cells_in(sp_ctx* c, size_t slot, sp_cell_geom& g) {
  if (!g.halo_fields) return SP_OK;
  for (int q = 0; q < kGeomHalo; ++q) {
    const double* h = geom_halo_get(g, q);
    if (!h) continue;
    void* b;
    sp_status s = stage(c, slot + q, (size_t)c->dm.N * sizeof(double), &b);
    if (s || (s = halo_to_cell(c, h, static_cast<double*>(b)))) return s;
    geom_halo_set(g, q, static_cast<double*>(b));
  }
  return SP_OK;
}
// After the kernels: the members in `mask` (bit q) back into the halo
// arrays of `orig` (interior entries only).
static sp_status geom_cells_out(sp_ctx* c, const sp_cell_geom& orig, const sp_cell_geom& conv,
                                unsigned mask) {
  if (!orig.halo_fields) return SP_OK;
  for (int q = 0; q < kGeomHalo; ++q) {
    const double* h = geom_halo_get(orig, q);
    if (!((mask >> q) & 1u) || !h) continue;
    sp_status s = cell_to_halo(c, geom_halo_get(conv, q), const_cast<double*>(h));
    if (s) return s;
  }
  return SP_OK;
}

// The same for sp_liquid_field: phi_l, nx_l, ny_l, nz_l, vof_l.
constexpr int kLiqHalo = 5;
static double* liq_halo_get(const sp_liquid_field& g, int q) {
  double* const m[kLiqHalo] = {g.phi_l, g.nx_l, g.ny_l, g.nz_l, g.vof_l};
  return m[q];
}
static void liq_halo_set(sp_liquid_field& g, int q, double* p) {
  switch (q) {
    case 0: g.phi_l = p; break;
    case 1: g.nx_l = p; break;
    case 2: g.ny_l = p; break;
    case 3: g.nz_l = p; break;
    default: g.vof_l = p; break;
  }
}
static sp_status liq_cells_in(sp_ctx* c, size_t slot, sp_liquid_field& g) {
  if (!g.halo_fields) return SP_OK;
  for (int q = 0; q < kLiqHalo; ++q) {
    void* b;
    sp_status s = stage(c, slot + q, (size_t)c->dm.N * sizeof(double), &b);
    if (s || (s = halo_to_cell(c, liq_halo_get(g, q), static_cast<double*>(b)))) return s;
    liq_halo_set(g, q, static_cast<double*>(b));
  }
  return SP_OK;
}
static sp_status liq_cells_out(sp_ctx* c, const sp_liquid_field& orig,
                               const sp_liquid_field& conv) {
  if (!orig.halo_fields) return SP_OK;
  for (int q = 0; q < kLiqHalo; ++q) {
    sp_status s = cell_to_halo(c, liq_halo_get(conv, q), liq_halo_get(orig, q));
    if (s) return s;
  }
  return SP_OK;
}

// halo_fields selects the indexing of caller arrays, so a value from an
// uninitialised struct must not reach the kernels
static sp_status check_layout(sp_ctx* c, int32_t halo_fields, const char* what) {
  if (halo_fields == 0 || halo_fields == 1) return SP_OK;
  return fail(c, SP_ERR_INVALID_ARG, what);
}

static sp_status check_geom(sp_ctx* c, const sp_cell_geom* const g[4]) {
  for (int m = 0; m < 4; ++m) {
    if (!g[m]) return fail(c, SP_ERR_INVALID_ARG, "null sp_cell_geom");
    const sp_cell_geom& a = *g[m];
    if (sp_status s = check_layout(c, a.halo_fields, "sp_cell_geom.halo_fields must be 0 or 1"))
      return s;
    const void* req[] = {a.dx, a.dy, a.dz, a.phi, a.nx, a.ny, a.nz, a.vof, a.vof_l,
                         a.cell_type, a.ee_area, a.ne_area, a.te_area, a.fce, a.fcn, a.fct,
                         a.cell_cent, a.wllh, a.delh};
    for (const void* p : req)
      if (!p) return fail(c, SP_ERR_INVALID_ARG, "sp_cell_geom member is null");
  }
  return SP_OK;
}

sp_status sp_grid_preprocess_dev(sp_ctx* c, const sp_cell_geom* p, const sp_cell_geom* u,
                                 const sp_cell_geom* v, const sp_cell_geom* w) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  const sp_cell_geom* g[4] = {p, u, v, w};
  sp_status s = check_geom(c, g);
  if (s) return s;
  // halo-allocated level sets, normals and volume fractions as cell arrays
  sp_cell_geom d[4];
  const sp_cell_geom* dg[4];
  for (int m = 0; m < 4; ++m) {
    d[m] = *g[m];
    if ((s = geom_cells_in(c, 600 + 6 * m, d[m]))) return s;
    dg[m] = &d[m];
  }
  if ((s = grid_preprocess(c, dg))) return s;
  for (int m = 0; m < 4; ++m)
    if ((s = geom_cells_out(c, *g[m], d[m], kGeomVof | kGeomVofL))) return s;
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_grid_preprocess");
  timers_collect(c);
  return SP_OK;
}

sp_status sp_grid_preprocess(sp_ctx* c, const sp_cell_geom* p, const sp_cell_geom* u,
                             const sp_cell_geom* v, const sp_cell_geom* w) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  const sp_cell_geom* g[4] = {p, u, v, w};
  sp_status s = check_geom(c, g);
  if (s) return s;
  const size_t N = (size_t)c->dm.N, Nh = (size_t)c->dm.Nh;
  // staging slots past those of the projection, adjoint and wall step
  size_t k = 100;
  sp_cell_geom d[4], up[4];
  for (int m = 0; m < 4; ++m) {
    const sp_cell_geom& h = *g[m];
    sp_cell_geom& o = d[m];
    o.halo_fields = h.halo_fields;
    const size_t nl = geom_count(h, N, Nh);  // phi, nx, ny, nz, vof, vof_l
    const double* t;
    const int32_t* ti;
    if ((s = upload(c, k++, h.dx, N, &o.dx)) || (s = upload(c, k++, h.dy, N, &o.dy)) ||
        (s = upload(c, k++, h.dz, N, &o.dz)) || (s = upload(c, k++, h.phi, nl, &o.phi)) ||
        (s = upload(c, k++, h.nx, nl, &o.nx)) || (s = upload(c, k++, h.ny, nl, &o.ny)) ||
        (s = upload(c, k++, h.nz, nl, &o.nz)))
      return s;
    // read-write arrays: uploaded, written back below
    if ((s = upload(c, k++, static_cast<const double*>(h.vof), nl, &t))) return s;
    o.vof = const_cast<double*>(t);
    if ((s = upload(c, k++, static_cast<const double*>(h.vof_l), nl, &t))) return s;
    o.vof_l = const_cast<double*>(t);
    if ((s = upload(c, k++, static_cast<const int32_t*>(h.cell_type), N, &ti))) return s;
    o.cell_type = const_cast<int32_t*>(ti);
    if ((s = upload(c, k++, static_cast<const double*>(h.fce), 3 * Nh, &t))) return s;
    o.fce = const_cast<double*>(t);
    if ((s = upload(c, k++, static_cast<const double*>(h.fcn), 3 * Nh, &t))) return s;
    o.fcn = const_cast<double*>(t);
    if ((s = upload(c, k++, static_cast<const double*>(h.fct), 3 * Nh, &t))) return s;
    o.fct = const_cast<double*>(t);
    if ((s = upload(c, k++, static_cast<const double*>(h.delh), N, &t))) return s;
    o.delh = const_cast<double*>(t);
    // write-only arrays
    void* b;
    if ((s = stage(c, k++, Nh * sizeof(double), &b))) return s;
    o.ee_area = static_cast<double*>(b);
    if ((s = stage(c, k++, Nh * sizeof(double), &b))) return s;
    o.ne_area = static_cast<double*>(b);
    if ((s = stage(c, k++, Nh * sizeof(double), &b))) return s;
    o.te_area = static_cast<double*>(b);
    if ((s = stage(c, k++, 3 * N * sizeof(double), &b))) return s;
    o.cell_cent = static_cast<double*>(b);
    if ((s = stage(c, k++, N * sizeof(double), &b))) return s;
    o.wllh = static_cast<double*>(b);
    up[m] = o;
    if ((s = geom_cells_in(c, 600 + 6 * m, o))) return s;
  }
  const sp_cell_geom* dg[4] = {&d[0], &d[1], &d[2], &d[3]};
  if ((s = grid_preprocess(c, dg))) return s;
  for (int m = 0; m < 4; ++m) {
    const sp_cell_geom& h = *g[m];
    const sp_cell_geom& o = d[m];
    if ((s = geom_cells_out(c, up[m], o, kGeomVof | kGeomVofL))) return s;
    const size_t nl = geom_count(h, N, Nh);
    const struct { void* dst; const void* src; size_t bytes; } back[] = {
        {h.vof, up[m].vof, nl * 8},      {h.vof_l, up[m].vof_l, nl * 8},
        {h.cell_type, o.cell_type, N * 4}, {h.ee_area, o.ee_area, Nh * 8},
        {h.ne_area, o.ne_area, Nh * 8},  {h.te_area, o.te_area, Nh * 8},
        {h.fce, o.fce, 3 * Nh * 8},      {h.fcn, o.fcn, 3 * Nh * 8},
        {h.fct, o.fct, 3 * Nh * 8},      {h.cell_cent, o.cell_cent, 3 * N * 8},
        {h.wllh, o.wllh, N * 8},         {h.delh, o.delh, N * 8}};
    for (const auto& x : back)
      if ((s = pipe_download(c, x.dst, x.src, x.bytes))) return s;
  }
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_grid_preprocess");
  timers_collect(c);
  return SP_OK;
}

sp_status sp_new_cell_face(sp_ctx* c, const sp_cell_geom* p, const sp_cell_geom* u,
                           const sp_cell_geom* v, const sp_cell_geom* w, const sp_cell_interp* iu,
                           const sp_cell_interp* iv, const sp_cell_interp* iw, int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  const sp_cell_geom* g[4] = {p, u, v, w};
  const sp_cell_interp* x[3] = {iu, iv, iw};
  for (int m = 0; m < 4; ++m) {
    if (!g[m]) return fail(c, SP_ERR_INVALID_ARG, "null sp_cell_geom");
    const sp_cell_geom& a = *g[m];
    if (sp_status s = check_layout(c, a.halo_fields, "sp_cell_geom.halo_fields must be 0 or 1"))
      return s;
    // P's spacing is read too: the face along a grid's own axis crosses a P cell
    const void* req[] = {a.dx, a.dy, a.dz, a.ee_area, a.ne_area, a.te_area, a.fce, a.fcn, a.fct};
    for (const void* q : req)
      if (!q) return fail(c, SP_ERR_INVALID_ARG, "sp_cell_geom member is null");
    if (m == 0) continue;
    const void* req2[] = {a.phi, a.nx, a.ny, a.nz, a.cell_type, a.cell_cent, a.delh};
    for (const void* q : req2)
      if (!q) return fail(c, SP_ERR_INVALID_ARG, "sp_cell_geom member is null");
  }
  for (int m = 0; m < 3; ++m) {
    if (!x[m]) return fail(c, SP_ERR_INVALID_ARG, "null sp_cell_interp");
    const sp_cell_interp& b = *x[m];
    const void* req[] = {b.moex, b.msce, b.sxe, b.etae, b.ale, b.dale, b.syn, b.etan,
                         b.aln, b.daln, b.szt, b.etat, b.alt, b.dalt};
    for (const void* q : req)
      if (!q) return fail(c, SP_ERR_INVALID_ARG, "sp_cell_interp member is null");
  }
  const size_t N = (size_t)c->dm.N, Nh = (size_t)c->dm.Nh;
  void* scratch;  // the U/V/W centres before CellLinking
  sp_status s = stage(c, 150, 9 * N * sizeof(double), &scratch);
  if (s) return s;
  if (device) {
    sp_cell_geom d[4];
    const sp_cell_geom* dg[4];
    for (int m = 0; m < 4; ++m) {
      d[m] = *g[m];
      if ((s = geom_cells_in(c, 630 + 6 * m, d[m]))) return s;
      dg[m] = &d[m];
    }
    if ((s = new_cell_face(c, dg, x, static_cast<double*>(scratch)))) return s;
  } else {
    size_t k = 151;
    sp_cell_geom d[4];
    sp_cell_interp xi[3];
    auto up = [&](const double* h, size_t n, const double** o) { return upload(c, k++, h, n, o); };
    for (int m = 0; m < 4; ++m) {
      const sp_cell_geom& h = *g[m];
      sp_cell_geom& o = d[m];
      o = sp_cell_geom{};
      const double* t;
      if ((s = up(h.ee_area, Nh, &t))) return s;
      o.ee_area = const_cast<double*>(t);
      if ((s = up(h.ne_area, Nh, &t))) return s;
      o.ne_area = const_cast<double*>(t);
      if ((s = up(h.te_area, Nh, &t))) return s;
      o.te_area = const_cast<double*>(t);
      if ((s = up(h.fce, 3 * Nh, &t))) return s;
      o.fce = const_cast<double*>(t);
      if ((s = up(h.fcn, 3 * Nh, &t))) return s;
      o.fcn = const_cast<double*>(t);
      if ((s = up(h.fct, 3 * Nh, &t))) return s;
      o.fct = const_cast<double*>(t);
      if ((s = up(h.dx, N, &o.dx)) || (s = up(h.dy, N, &o.dy)) || (s = up(h.dz, N, &o.dz)))
        return s;
      if (m == 0) continue;
      const size_t nl = geom_count(h, N, Nh);
      o.halo_fields = h.halo_fields;
      if ((s = up(h.phi, nl, &o.phi)) || (s = up(h.nx, nl, &o.nx)) || (s = up(h.ny, nl, &o.ny)) ||
          (s = up(h.nz, nl, &o.nz)))
        return s;
      if ((s = geom_cells_in(c, 630 + 6 * m, o))) return s;
      const int32_t* ti;
      if ((s = upload(c, k++, static_cast<const int32_t*>(h.cell_type), N, &ti))) return s;
      o.cell_type = const_cast<int32_t*>(ti);
      if ((s = up(h.cell_cent, 3 * N, &t))) return s;
      o.cell_cent = const_cast<double*>(t);
      void* b;
      if ((s = stage(c, k++, N * sizeof(double), &b))) return s;
      o.delh = static_cast<double*>(b);
    }
    for (int m = 0; m < 3; ++m) {
      const sp_cell_interp& h = *x[m];
      sp_cell_interp& o = xi[m];
      const int32_t* ti;
      if ((s = upload(c, k++, h.moex, N, &o.moex))) return s;
      if ((s = upload(c, k++, static_cast<const int32_t*>(h.msce), 3 * N, &ti))) return s;
      o.msce = const_cast<int32_t*>(ti);
      double* const hs[12] = {h.sxe, h.etae, h.ale, h.dale, h.syn, h.etan,
                              h.aln, h.daln, h.szt, h.etat, h.alt, h.dalt};
      double** os[12] = {&o.sxe, &o.etae, &o.ale, &o.dale, &o.syn, &o.etan,
                         &o.aln, &o.daln, &o.szt, &o.etat, &o.alt, &o.dalt};
      for (int q = 0; q < 12; ++q) {
        const double* t;
        if ((s = up(hs[q], Nh, &t))) return s;
        *os[q] = const_cast<double*>(t);
      }
    }
    const sp_cell_geom* dg[4] = {&d[0], &d[1], &d[2], &d[3]};
    const sp_cell_interp* dxi[3] = {&xi[0], &xi[1], &xi[2]};
    if ((s = new_cell_face(c, dg, dxi, static_cast<double*>(scratch)))) return s;
    for (int m = 0; m < 3; ++m) {
      const sp_cell_geom& h = *g[1 + m];
      const sp_cell_geom& o = d[1 + m];
      if ((s = pipe_download(c, h.cell_cent, o.cell_cent, 3 * N * sizeof(double))) ||
          (s = pipe_download(c, h.delh, o.delh, N * sizeof(double))))
        return s;
      const sp_cell_interp& hx = *x[m];
      const sp_cell_interp& ox = xi[m];
      if ((s = pipe_download(c, hx.msce, ox.msce, 3 * N * sizeof(int32_t)))) return s;
      double* const hs[12] = {hx.sxe, hx.etae, hx.ale, hx.dale, hx.syn, hx.etan,
                              hx.aln, hx.daln, hx.szt, hx.etat, hx.alt, hx.dalt};
      double* const os[12] = {ox.sxe, ox.etae, ox.ale, ox.dale, ox.syn, ox.etan,
                              ox.aln, ox.daln, ox.szt, ox.etat, ox.alt, ox.dalt};
      for (int q = 0; q < 12; ++q)
        if ((s = pipe_download(c, hs[q], os[q], Nh * sizeof(double)))) return s;
    }
  }
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_new_cell_face");
  timers_collect(c);
  return SP_OK;
}

sp_status sp_interpolate_lvs(sp_ctx* c, const sp_solid_body* body, const sp_lvs_grid* grid,
                             int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!body || !grid) return fail(c, SP_ERR_INVALID_ARG, "null sp_solid_body or sp_lvs_grid");
  const sp_status s = interpolate_lvs(c, *body, *grid, device != 0);
  timers_collect(c);
  return s;
}

sp_status sp_clsvof_scheme(sp_ctx* c, const sp_clsvof_inputs* in, sp_clsvof_state* st,
                           int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!in || !st) return fail(c, SP_ERR_INVALID_ARG, "null sp_clsvof_inputs or sp_clsvof_state");
  const Dims& D = c->dm;
  if (D.I < 3 || D.J < 3 || D.K < 3)
    return fail(c, SP_ERR_INVALID_ARG, "Clsvof_Scheme needs at least 3 cells in each direction");
  if (D.K > 65535 || (D.J + 3) / 4 > 65535)
    return fail(c, SP_ERR_INVALID_ARG, "Clsvof_Scheme: Kmax or Jmax beyond the launch grid");
  const void* req[] = {in->u, in->v, in->w, in->x, in->y, in->z, in->dx, in->dy, in->dz,
                       in->nx, in->ny, in->nz, in->bcu_w, in->bcv_s, in->bcw_b,
                       st->vof_l, st->phi_l, st->nx_l, st->ny_l, st->nz_l};
  for (const void* q : req)
    if (!q) return fail(c, SP_ERR_INVALID_ARG, "required Clsvof_Scheme array is null");
  for (int f = 0; f < 6; ++f) {
    if (!st->lvs_var[f]) return fail(c, SP_ERR_INVALID_ARG, "sp_clsvof_state.lvs_var entry is null");
    if ((in->lvs_flag[f] != 0 && in->lvs_flag[f] != 1) || (in->vof_flag[f] != 0 && in->vof_flag[f] != 1))
      return fail(c, SP_ERR_INVALID_ARG, "BCLvs/BCVof flags must be 0 (Dirichlet) or 1 (Neumann)");
  }
  if (!(in->dt >= 0.0)) return fail(c, SP_ERR_INVALID_ARG, "dt must be non-negative");
  const size_t N = (size_t)D.N, Nh = (size_t)D.Nh;
  const size_t nface[6] = {(size_t)D.J * D.K, (size_t)D.J * D.K, (size_t)D.I * D.K,
                           (size_t)D.I * D.K, (size_t)D.I * D.J, (size_t)D.I * D.J};
  void* work;
  sp_status s = stage(c, 500, clsvof_work_doubles(D) * sizeof(double), &work);
  if (s) return s;
  if (device) {
    s = clsvof_scheme(c, *in, *st, static_cast<double*>(work));
    if (s) return s;
    const hipError_t e = hipStreamSynchronize(c->stream);
    timers_collect(c);
    return e == hipSuccess ? SP_OK : hip_check(c, e, "Clsvof_Scheme");
  }
  // host arrays: staging slots 510..559 belong to Clsvof_Scheme alone
  size_t k = 510;
  sp_clsvof_inputs d = *in;
  auto upd = [&](const double* h, size_t n, const double** o) { return upload(c, k++, h, n, o); };
  const double** ins[15] = {&d.u, &d.v, &d.w, &d.x, &d.y, &d.z, &d.dx, &d.dy, &d.dz,
                            &d.nx, &d.ny, &d.nz, &d.bcu_w, &d.bcv_s, &d.bcw_b};
  const size_t nin[15] = {Nh, Nh, Nh, N, N, N, N, N, N, Nh, Nh, Nh, nface[0], nface[2], nface[4]};
  for (int q = 0; q < 15; ++q)
    if ((s = upd(*ins[q], nin[q], ins[q]))) return s;
  sp_clsvof_state ds = *st;
  double** sts[5] = {&ds.vof_l, &ds.phi_l, &ds.nx_l, &ds.ny_l, &ds.nz_l};
  for (int q = 0; q < 5; ++q) {
    const double* t;
    if ((s = upd(*sts[q], Nh, &t))) return s;
    *sts[q] = const_cast<double*>(t);
  }
  for (int f = 0; f < 6; ++f) {
    const double* t;
    if ((s = upd(st->lvs_var[f], nface[f], &t))) return s;
    ds.lvs_var[f] = const_cast<double*>(t);
    if (st->vof_var[f]) {
      void* o;
      if ((s = stage(c, k++, nface[f] * sizeof(double), &o))) return s;
      ds.vof_var[f] = static_cast<double*>(o);
    }
  }
  if ((s = clsvof_scheme(c, d, ds, static_cast<double*>(work)))) return s;
  double* const hs[5] = {st->vof_l, st->phi_l, st->nx_l, st->ny_l, st->nz_l};
  const double* const dsrc[5] = {ds.vof_l, ds.phi_l, ds.nx_l, ds.ny_l, ds.nz_l};
  for (int q = 0; q < 5; ++q)
    if ((s = pipe_download(c, hs[q], dsrc[q], Nh * sizeof(double)))) return s;
  for (int f = 0; f < 6; ++f) {
    if ((s = pipe_download(c, st->lvs_var[f], ds.lvs_var[f], nface[f] * sizeof(double)))) return s;
    if (st->vof_var[f] &&
        (s = pipe_download(c, st->vof_var[f], ds.vof_var[f], nface[f] * sizeof(double))))
      return s;
  }
  timers_collect(c);
  return SP_OK;
}

sp_status sp_compute_uvw_liquid_field(sp_ctx* c, const sp_liquid_field* p,
                                      const sp_liquid_field* u, const sp_liquid_field* v,
                                      const sp_liquid_field* w, int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  const sp_liquid_field* g[4] = {p, u, v, w};
  for (int m = 0; m < 4; ++m) {
    if (!g[m]) return fail(c, SP_ERR_INVALID_ARG, "null sp_liquid_field");
    const sp_liquid_field& a = *g[m];
    if (sp_status s = check_layout(c, a.halo_fields, "sp_liquid_field.halo_fields must be 0 or 1"))
      return s;
    const void* req[] = {a.dx, a.dy, a.dz, a.phi_l, a.nx_l, a.ny_l, a.nz_l, a.vof_l};
    for (const void* q : req)
      if (!q) return fail(c, SP_ERR_INVALID_ARG, "sp_liquid_field member is null");
  }
  const size_t N = (size_t)c->dm.N, Nh = (size_t)c->dm.Nh;
  sp_status s;
  if (device) {
    // halo-allocated liquid fields as cell arrays, the U/V/W results back
    sp_liquid_field d[4];
    const sp_liquid_field* dg[4];
    for (int m = 0; m < 4; ++m) {
      d[m] = *g[m];
      if ((s = liq_cells_in(c, 700 + 5 * m, d[m]))) return s;
      dg[m] = &d[m];
    }
    if ((s = uvw_liquid_field(c, dg))) return s;
    for (int m = 1; m < 4; ++m)
      if ((s = liq_cells_out(c, *g[m], d[m]))) return s;
  } else {
    size_t k = 250;  // staging slots shared with the other geometry entry points' range
    sp_liquid_field d[4], up[4];
    for (int m = 0; m < 4; ++m) {
      const sp_liquid_field& h = *g[m];
      sp_liquid_field& o = d[m];
      o.halo_fields = h.halo_fields;
      if ((s = upload(c, k++, h.dx, N, &o.dx)) || (s = upload(c, k++, h.dy, N, &o.dy)) ||
          (s = upload(c, k++, h.dz, N, &o.dz)))
        return s;
      double* const hs[5] = {h.phi_l, h.nx_l, h.ny_l, h.nz_l, h.vof_l};
      double** os[5] = {&o.phi_l, &o.nx_l, &o.ny_l, &o.nz_l, &o.vof_l};
      for (int f = 0; f < 5; ++f) {
        // P's fields are read; U/V/W's are written (their halo entries kept)
        if (m == 0 || h.halo_fields) {
          const double* t;
          if ((s = upload(c, k++, static_cast<const double*>(hs[f]), h.halo_fields ? Nh : N, &t)))
            return s;
          *os[f] = const_cast<double*>(t);
        } else {
          void* b;
          if ((s = stage(c, k++, N * sizeof(double), &b))) return s;
          *os[f] = static_cast<double*>(b);
        }
      }
      up[m] = o;
      if ((s = liq_cells_in(c, 700 + 5 * m, o))) return s;
    }
    const sp_liquid_field* dg[4] = {&d[0], &d[1], &d[2], &d[3]};
    if ((s = uvw_liquid_field(c, dg))) return s;
    for (int m = 1; m < 4; ++m) {
      if ((s = liq_cells_out(c, up[m], d[m]))) return s;
      double* const hs[5] = {g[m]->phi_l, g[m]->nx_l, g[m]->ny_l, g[m]->nz_l, g[m]->vof_l};
      double* const os[5] = {up[m].phi_l, up[m].nx_l, up[m].ny_l, up[m].nz_l, up[m].vof_l};
      const size_t n = g[m]->halo_fields ? Nh : N;
      for (int f = 0; f < 5; ++f)
        if ((s = pipe_download(c, hs[f], os[f], n * sizeof(double)))) return s;
    }
  }
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_compute_uvw_liquid_field");
  timers_collect(c);
  return SP_OK;
}

sp_status sp_define_momentum_exchange_cell(sp_ctx* c, const int32_t* ctp, const int32_t* ctu,
                                           const int32_t* ctv, const int32_t* ctw, int32_t* mu,
                                           int32_t* mv, int32_t* mw, int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!ctp || !ctu || !ctv || !ctw || !mu || !mv || !mw)
    return fail(c, SP_ERR_INVALID_ARG, "null Cell_Type or MoExCell pointer");
  const size_t N = (size_t)c->dm.N;
  sp_status s;
  if (device) {
    if ((s = define_moex(c, ctp, ctu, ctv, ctw, mu, mv, mw))) return s;
  } else {
    // staging slots past those the projection, adjoint and wall step keep
    // between calls; the geometry entry points share theirs within a call
    size_t k = 130;
    const int32_t *d[4], *m[3];
    const int32_t* hin[4] = {ctp, ctu, ctv, ctw};
    int32_t* hm[3] = {mu, mv, mw};
    for (int q = 0; q < 4; ++q)
      if ((s = upload(c, k++, hin[q], N, &d[q]))) return s;
    for (int q = 0; q < 3; ++q)
      if ((s = upload(c, k++, static_cast<const int32_t*>(hm[q]), N, &m[q]))) return s;
    if ((s = define_moex(c, d[0], d[1], d[2], d[3], const_cast<int32_t*>(m[0]),
                         const_cast<int32_t*>(m[1]), const_cast<int32_t*>(m[2]))))
      return s;
    for (int q = 0; q < 3; ++q)
      if ((s = pipe_download(c, hm[q], m[q], N * sizeof(int32_t)))) return s;
  }
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_define_momentum_exchange_cell");
  timers_collect(c);
  return SP_OK;
}

sp_status sp_number_external_cell(sp_ctx* c, int iu, int iv, int iw, const int32_t* ct,
                                  int32_t* posnu, int32_t* ext_cell, int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  if (!ct || !posnu || !ext_cell)
    return fail(c, SP_ERR_INVALID_ARG, "null Cell_Type, PosNu or ExtCell pointer");
  if ((iu != 0 && iu != 1) || (iv != 0 && iv != 1) || (iw != 0 && iw != 1))
    return fail(c, SP_ERR_INVALID_ARG, "iu, iv, iw must be 0 or 1");
  const Dims& D = c->dm;
  const size_t N = (size_t)D.N;
  const size_t NP = (size_t)(D.I - iu) * (D.J - iv) * (D.K - iw);
  const size_t NC = number_scratch(D.I - iu, D.J - iv, D.K - iw);
  void *cnt, *ext;
  sp_status s;
  if ((s = stage(c, 140, NC * sizeof(int32_t), &cnt)) ||
      (s = stage(c, 141, sizeof(int32_t), &ext)))
    return s;
  const int32_t* dct = ct;
  int32_t* dpos = posnu;
  if (!device) {
    void* p;
    if ((s = upload(c, 142, ct, N, &dct)) || (s = stage(c, 143, NP * sizeof(int32_t), &p)))
      return s;
    dpos = static_cast<int32_t*>(p);
  }
  if ((s = number_external(c, iu, iv, iw, dct, dpos, static_cast<int32_t*>(cnt),
                           static_cast<int32_t*>(ext))))
    return s;
  if (!device && (s = pipe_download(c, posnu, dpos, NP * sizeof(int32_t)))) return s;
  hipError_t e = hipMemcpyAsync(ext_cell, ext, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_number_external_cell");
  timers_collect(c);
  return SP_OK;
}

static sp_status check_pred(sp_ctx* c, const sp_pred_inputs* in, const sp_pred_outputs* out) {
  if (!in || !out) return fail(c, SP_ERR_INVALID_ARG, "null predictor inputs/outputs");
  if (in->space_order != 1 && in->space_order != 2)
    return fail(c, SP_ERR_INVALID_ARG,
                "SpaceOrder 1 (first-order upwind) and 2 (central) are built, not 3");
  if (in->time_order < 1 || in->time_order > 3)
    return fail(c, SP_ERR_INVALID_ARG, "time_order must be 1, 2 or 3");
  if (!(in->dt > 0.0) || !std::isfinite(in->dt))
    return fail(c, SP_ERR_INVALID_ARG, "dt must be > 0");
  if (!c->grid_set) return fail(c, SP_ERR_GRID, "sp_set_grid has not been called");
  const void* req[] = {in->p_ee_area, in->p_ne_area, in->p_te_area, in->u_old, in->v_old,
                       in->w_old, in->u_old_x, in->v_old_x, in->w_old_x, in->u_cur, in->v_cur,
                       in->w_cur, in->u, in->v, in->w, in->flux_div_old, out->u, out->v, out->w,
                       out->dp_u, out->dp_v, out->dp_w};
  for (const void* p : req)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "required predictor pointer is null");
  if (in->time_order == 2 && (!in->u_n || !in->v_n || !in->w_n))
    return fail(c, SP_ERR_INVALID_ARG, "TimeOrder 2 needs TVar_n (u_n, v_n, w_n)");
  for (int q = 0; q < 18; ++q) {
    if (in->bc_flag[q] != 0 && in->bc_flag[q] != 1)
      return fail(c, SP_ERR_INVALID_ARG, "predictor bc_flag must be 0 or 1");
    if (!in->bc_var[q]) return fail(c, SP_ERR_INVALID_ARG, "predictor bc_var entry is null");
  }
  const sp_cell_geom* og[3] = {in->u_old, in->v_old, in->w_old};
  const sp_cell_interp* ox[3] = {in->u_old_x, in->v_old_x, in->w_old_x};
  const sp_cell_geom* cg[3] = {in->u_cur, in->v_cur, in->w_cur};
  for (int q = 0; q < 3; ++q) {
    const sp_cell_geom& g = *og[q];
    const void* rg[] = {g.phi, g.nx, g.ny, g.nz, g.vof, g.cell_type, g.ee_area, g.ne_area,
                        g.te_area, g.fce, g.fcn, g.fct, g.cell_cent};
    for (const void* p : rg)
      if (!p) return fail(c, SP_ERR_INVALID_ARG, "a member of an old U/V/W cell is null");
    const sp_cell_interp& x = *ox[q];
    const void* rx[] = {x.moex, x.sxe, x.syn, x.szt, x.etae, x.etan, x.etat, x.ale, x.aln, x.alt};
    for (const void* p : rx)
      if (!p) return fail(c, SP_ERR_INVALID_ARG, "a member of an old U/V/W cell's interpolation data is null");
    const sp_cell_geom& h = *cg[q];
    const void* rc[] = {h.vof, h.vof_l, h.cell_type, h.wllh, h.delh};
    for (const void* p : rc)
      if (!p) return fail(c, SP_ERR_INVALID_ARG, "a member of a current U/V/W cell is null");
    sp_status s = check_layout(c, g.halo_fields, "sp_cell_geom.halo_fields must be 0 or 1");
    if (s || (s = check_layout(c, h.halo_fields, "sp_cell_geom.halo_fields must be 0 or 1")))
      return s;
  }
  return SP_OK;
}

sp_status sp_predict_uvw(sp_ctx* c, const sp_pred_inputs* in, sp_pred_outputs* out, int device) {
  if (!c) return SP_ERR_INVALID_ARG;
  hipSetDevice(c->device);
  sp_status s = check_pred(c, in, out);
  if (s) return s;
  const Dims& D = c->dm;
  const size_t N = (size_t)D.N, Nh = (size_t)D.Nh;
  void* faces;
  const bool second = in->space_order == 2 && in->itt != 1;
  if ((s = stage(c, 299, predict_face_doubles(D, second) * sizeof(double), &faces))) return s;
  if (device) {
    // halo-allocated cell fields of the old and current cells as cell arrays
    sp_pred_inputs d = *in;
    sp_cell_geom og[3] = {*in->u_old, *in->v_old, *in->w_old};
    sp_cell_geom cg[3] = {*in->u_cur, *in->v_cur, *in->w_cur};
    for (int q = 0; q < 3; ++q)
      if ((s = geom_cells_in(c, 660 + 6 * q, og[q])) || (s = geom_cells_in(c, 680 + 6 * q, cg[q])))
        return s;
    d.u_old = &og[0]; d.v_old = &og[1]; d.w_old = &og[2];
    d.u_cur = &cg[0]; d.v_cur = &cg[1]; d.w_cur = &cg[2];
    if ((s = predict_uvw(c, d, *out, static_cast<double*>(faces)))) return s;
  } else {
    // staging slots from 1000 up belong to the predictor alone (its ~120
    // arrays must not reach the projection's geometry slots at 400)
    size_t k = 1000;
    sp_pred_inputs d = *in;
    auto upd = [&](const double* h, size_t n, const double** o) { return upload(c, k++, h, n, o); };
    auto upi = [&](const int32_t* h, size_t n, const int32_t** o) { return upload(c, k++, h, n, o); };
    if ((s = upd(in->p_ee_area, Nh, &d.p_ee_area)) || (s = upd(in->p_ne_area, Nh, &d.p_ne_area)) ||
        (s = upd(in->p_te_area, Nh, &d.p_te_area)))
      return s;
    sp_cell_geom og[3], cg[3];
    sp_cell_interp ox[3];
    const sp_cell_geom* hg[3] = {in->u_old, in->v_old, in->w_old};
    const sp_cell_interp* hx[3] = {in->u_old_x, in->v_old_x, in->w_old_x};
    const sp_cell_geom* hc[3] = {in->u_cur, in->v_cur, in->w_cur};
    for (int q = 0; q < 3; ++q) {
      const sp_cell_geom& h = *hg[q];
      sp_cell_geom& o = og[q];
      o = sp_cell_geom{};
      o.halo_fields = h.halo_fields;
      const size_t nl = geom_count(h, N, Nh);
      const double* t;
      const int32_t* ti;
      if ((s = upd(h.phi, nl, &o.phi)) || (s = upd(h.nx, nl, &o.nx)) ||
          (s = upd(h.ny, nl, &o.ny)) || (s = upd(h.nz, nl, &o.nz)))
        return s;
      if ((s = upd(h.vof, nl, &t))) return s;
      o.vof = const_cast<double*>(t);
      if ((s = upi(h.cell_type, N, &ti))) return s;
      o.cell_type = const_cast<int32_t*>(ti);
      double* const hv[7] = {h.ee_area, h.ne_area, h.te_area, h.fce, h.fcn, h.fct, h.cell_cent};
      double** ov[7] = {&o.ee_area, &o.ne_area, &o.te_area, &o.fce, &o.fcn, &o.fct, &o.cell_cent};
      const size_t nv[7] = {Nh, Nh, Nh, 3 * Nh, 3 * Nh, 3 * Nh, 3 * N};
      for (int m = 0; m < 7; ++m) {
        if ((s = upd(hv[m], nv[m], &t))) return s;
        *ov[m] = const_cast<double*>(t);
      }
      const sp_cell_interp& x = *hx[q];
      sp_cell_interp& ou = ox[q];
      ou = sp_cell_interp{};
      if ((s = upi(x.moex, N, &ou.moex))) return s;
      double* const xv[9] = {x.sxe, x.syn, x.szt, x.etae, x.etan, x.etat, x.ale, x.aln, x.alt};
      double** xo[9] = {&ou.sxe, &ou.syn, &ou.szt, &ou.etae, &ou.etan, &ou.etat,
                        &ou.ale, &ou.aln, &ou.alt};
      for (int m = 0; m < 9; ++m) {
        if ((s = upd(xv[m], Nh, &t))) return s;
        *xo[m] = const_cast<double*>(t);
      }
      const sp_cell_geom& hcq = *hc[q];
      sp_cell_geom& oc = cg[q];
      oc = sp_cell_geom{};
      oc.halo_fields = hcq.halo_fields;
      const size_t nc = geom_count(hcq, N, Nh);  // vof, vof_l
      double* const cv[4] = {hcq.vof, hcq.vof_l, hcq.wllh, hcq.delh};
      double** co[4] = {&oc.vof, &oc.vof_l, &oc.wllh, &oc.delh};
      for (int m = 0; m < 4; ++m) {
        if ((s = upd(cv[m], m < 2 ? nc : N, &t))) return s;
        *co[m] = const_cast<double*>(t);
      }
      if ((s = upi(hcq.cell_type, N, &ti))) return s;
      oc.cell_type = const_cast<int32_t*>(ti);
      if ((s = geom_cells_in(c, 660 + 6 * q, o)) || (s = geom_cells_in(c, 680 + 6 * q, oc)))
        return s;
    }
    d.u_old = &og[0]; d.v_old = &og[1]; d.w_old = &og[2];
    d.u_old_x = &ox[0]; d.v_old_x = &ox[1]; d.w_old_x = &ox[2];
    d.u_cur = &cg[0]; d.v_cur = &cg[1]; d.w_cur = &cg[2];
    if ((s = upd(in->u, Nh, &d.u)) || (s = upd(in->v, Nh, &d.v)) || (s = upd(in->w, Nh, &d.w)))
      return s;
    if (in->time_order == 2 &&
        ((s = upd(in->u_n, Nh, &d.u_n)) || (s = upd(in->v_n, Nh, &d.v_n)) ||
         (s = upd(in->w_n, Nh, &d.w_n))))
      return s;
    const double* fd;
    if ((s = upd(in->flux_div_old, 3 * N, &fd))) return s;
    d.flux_div_old = const_cast<double*>(fd);
    const size_t bl[6] = {(size_t)D.J * D.K, (size_t)D.J * D.K, (size_t)D.I * D.K,
                          (size_t)D.I * D.K, (size_t)D.I * D.J, (size_t)D.I * D.J};
    for (int q = 0; q < 18; ++q)
      if ((s = upd(in->bc_var[q], bl[q % 6], &d.bc_var[q]))) return s;
    // Pred keeps the entries it does not write: upload, then copy back whole
    sp_pred_outputs o{};
    double* ho[6] = {out->u, out->v, out->w, out->dp_u, out->dp_v, out->dp_w};
    double** oo[6] = {&o.u, &o.v, &o.w, &o.dp_u, &o.dp_v, &o.dp_w};
    for (int q = 0; q < 6; ++q) {
      const double* t;
      if (q < 3) {
        if ((s = upd(ho[q], Nh, &t))) return s;
      } else {
        void* b;
        if ((s = stage(c, k++, Nh * sizeof(double), &b))) return s;
        t = static_cast<const double*>(b);
      }
      *oo[q] = const_cast<double*>(t);
    }
    if ((s = predict_uvw(c, d, o, static_cast<double*>(faces)))) return s;
    for (int q = 0; q < 6; ++q)
      if ((s = pipe_download(c, ho[q], *oo[q], Nh * sizeof(double)))) return s;
    if ((s = pipe_download(c, in->flux_div_old, d.flux_div_old, 3 * N * sizeof(double)))) return s;
  }
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "sp_predict_uvw");
  timers_collect(c);
  return SP_OK;
}

void* sp_stream(sp_ctx* c) { return c ? (void*)c->stream : nullptr; }

void sp_set_profiling(sp_ctx* c, int on) {
  if (!c) return;
  c->profiling = on != 0;
  c->timers.clear();
  c->ev_used.clear();
}

void sp_set_profiling_filter(sp_ctx* c, const char* name) {
  if (!c) return;
  c->prof_filter = name ? name : "";
  c->timers.clear();
  c->ev_used.clear();
}

int sp_kernel_times(sp_ctx* c, sp_kernel_time* out, int cap) {
  if (!c) return 0;
  const int n = (int)c->timers.size();
  for (int q = 0; q < n && q < cap && out; ++q) {
    std::memset(&out[q], 0, sizeof(out[q]));
    std::snprintf(out[q].name, sizeof(out[q].name), "%s", c->timers[q].name.c_str());
    out[q].launches = c->timers[q].launches;
    out[q].total_ms = c->timers[q].total_ms;
    out[q].bytes = c->timers[q].bytes;
  }
  return n;
}

const char* sp_strerror(sp_status s) {
  switch (s) {
    case SP_OK: return "ok";
    case SP_ERR_INVALID_ARG: return "invalid argument";
    case SP_ERR_HIP: return "HIP runtime error";
    case SP_ERR_NOT_CONVERGED: return "solver did not reach the divergence tolerance";
    case SP_ERR_NONFINITE: return "non-finite value";
    case SP_ERR_OOM: return "device out of memory";
    case SP_ERR_GRID: return "grid not set or not a tensor-product grid";
    case SP_ERR_STATE: return "no forward projection to differentiate";
    case SP_ERR_NO_DEVICE: return "no usable HIP device";
    case SP_ERR_CAPACITY: return "an engine work list is full";
  }
  return "unknown status";
}

const char* sp_last_error(sp_ctx* c) { return c ? c->last_error.c_str() : ""; }

}  // extern "C"
