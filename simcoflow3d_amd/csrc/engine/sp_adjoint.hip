// Reverse-mode derivative of the projection.
//
// Forward (per face f between cells P and N, aperture-area a_f, coefficient
// K_f = PoCoef(:,4..6), mask m_f = not MoEx and not solid):
//   b   = G^T diag(a) u*            (SetPoissonVectors, ProjectionP.f90:521-526)
//   A p = b,  A = G^T diag(a K) G    (symmetric, SetPoissonMatrix)
//   u   = m (u* - K G p)             (ComputePUVW.f90:185-219)
// Backward for dL/du = g and dL/dPp = gp:
//   pbar = -G^T (m K g)              (with the pressure ghost rule at walls)
//        + Q^T gp                    (Q: the output field Pp of the iterate x,
//                                     its gauge shift and face ghosts)
//   A lam = pbar                     (same multigrid PCG: A is self-adjoint)
//   dL/du* = m g + a G lam
//   dL/dK  = -m g (G p) - a (G lam)(G p)   (second term on faces between active cells)
// Only the forward pressure is kept from the forward call; everything else is
// recomputed.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "sp_device.h"
#include "sp_geom.h"
#include "sp_internal.h"

namespace sp {
namespace {

struct AdjArgs {
  int I, J, K, SX, SY;
  int64_t N;
  const double* xp;       // forward pressure (interior, shift applied below)
  double shift;
  const uint8_t* active;
  const double* pocoef;
  const int32_t *ctu, *ctv, *ctw, *mu, *mv, *mw;
  int32_t fl[6];
  const double *gu, *gv, *gw;
  const double* gp;       // dL/dPp (halo array) or nullptr
  int64_t pin;            // interior index of the gauge cell, or -1 (mean gauge)
  int pin_a, pin_b, pin_c;
  const double *ee, *ne, *te;
  const double *dxp, *dyp, *dzp;
  const double *dp[3];
  const double *dgrid[3];
  const double* lam;
  double* pbar;
  double *gus, *gvs, *gws, *gk[3], *gb[3];
  RedSite site;
  Scalars* S;
};

__device__ inline int64_t nidx(const AdjArgs& A, int a, int b, int c) {
  return a + (int64_t)A.I * (b + (int64_t)A.J * c);
}
__device__ inline int64_t hidx(const AdjArgs& A, int a, int b, int c) {
  return (a + 1) + (int64_t)A.SX * ((b + 1) + (int64_t)A.SY * (c + 1));
}

// m_f * g_f of the face on the + side of cell n in direction d
__device__ inline double tface(const AdjArgs& A, int d, int64_t n, int64_t h) {
  if (d == 0) return (A.mu[n] != 1 && A.ctu[n] != 2) ? A.gu[h] : 0.0;
  if (d == 1) return (A.mv[n] != 1 && A.ctv[n] != 2) ? A.gv[h] : 0.0;
  return (A.mw[n] != 1 && A.ctw[n] != 2) ? A.gw[h] : 0.0;
}

// pbar (gathered per cell) and the correction part of the gradients
__global__ void __launch_bounds__(256) k_adj_source(AdjArgs A) {
  const Cell3 cq = cell_of(A.I, A.J, A.K);
  const int a = cq.a, b = cq.b, c = cq.c;
  const int64_t N = A.N;
  double psum = 0.0, pmax = 0.0, qsum = 0.0;
  if (cq.in) {
    {
      const int64_t n = nidx(A, a, b, c), h = hidx(A, a, b, c);
      const int co[3] = {a, b, c};
      const int lim[3] = {A.I, A.J, A.K};
      const int64_t ns[3] = {1, A.I, (int64_t)A.I * A.J};
      const int64_t hs[3] = {1, A.SX, (int64_t)A.SX * A.SY};
      double* gs[3] = {A.gus, A.gvs, A.gws};
      const double pP = A.active[n] ? A.xp[n] - A.shift : 0.0;
      double pb = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double t = tface(A, d, n, h);
        const double K = A.pocoef[(3 + d) * N + n];
        gs[d][h] = t;  // direct part m g
        const bool last = co[d] == lim[d] - 1;
        const double pN = !last ? (A.active[n + ns[d]] ? A.xp[n + ns[d]] - A.shift : 0.0)
                                : (double)A.fl[2 * d + 1] * pP;
        A.gk[d][n] = -t * (pN - pP);
        // u_f depends on p(P) with +K and on p(N) (or the ghost) with -K
        pb += t * K;
        if (last) pb -= (double)A.fl[2 * d + 1] * t * K;
        if (co[d] > 0) {
          const int64_t m = n - ns[d];
          pb -= tface(A, d, m, h - hs[d]) * A.pocoef[(3 + d) * N + m];
        }
      }
      if (A.gp && A.active[n]) {
        // Pp(P) = x(P) - shift, and each face ghost is flag * Pp of its
        // inner cell (edges and corners 0)
        double q = A.gp[h];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (co[d] == 0) q += (double)A.fl[2 * d] * A.gp[h - hs[d]];
          if (co[d] == lim[d] - 1) q += (double)A.fl[2 * d + 1] * A.gp[h + hs[d]];
        }
        pb += q;
        qsum = q;
      }
      pb = A.active[n] ? pb : 0.0;
      A.pbar[n] = pb;
      psum += pb;
      pmax = fabs(pb) / (A.dxp[a] * A.dyp[b] * A.dzp[c]);
    }
  }
  double v[3] = {psum, pmax, qsum};
  block_partials<kTileThreads, 3, 2u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_adj_source_fin(AdjArgs A, uint32_t nb) {
  double t[3];
  finalize_partials<kFinThreads, 3, 2u>(A.site, nb, t);
  if (flat_tid() == 0) {
    Scalars* S = A.S;
    double bsum = t[0], smax = t[1];
    if (S->singular && A.pin >= 0 && A.active[A.pin]) {
      // pin gauge Pp = x - x(pin): its transpose takes the sum of the
      // pressure cotangent off the pin cell.  (The mean gauge's transpose is
      // the mean projection the solver applies to every singular source.)
      const double pb = A.pbar[A.pin] - t[2];
      A.pbar[A.pin] = pb;
      bsum -= t[2];
      const double r = fabs(pb) / (A.dxp[A.pin_a] * A.dyp[A.pin_b] * A.dzp[A.pin_c]);
      smax = r > smax ? r : smax;
    }
    S->bsum = bsum;
    S->src_max = smax;
    S->bmean = (S->singular && S->active > 0) ? bsum / (double)S->active : 0.0;
  }
}

// mean of lam over active cells (the gauge of a singular system)
__global__ void __launch_bounds__(256) k_adj_lam_mean(AdjArgs A) {
  const Cell3 q = cell_of(A.I, A.J, A.K);
  double v[1] = {0.0};
  if (q.in) {
    const int64_t n = nidx(A, q.a, q.b, q.c);
    if (A.active[n]) v[0] = A.lam[n];
  }
  block_partials<kTileThreads, 1, 0u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_adj_lam_mean_fin(AdjArgs A, uint32_t nb) {
  double t[1];
  finalize_partials<kFinThreads, 1, 0u>(A.site, nb, t);
  if (flat_tid() == 0) {
    Scalars* S = A.S;
    S->npart[0] = (S->singular && S->active > 0) ? t[0] / (double)S->active : 0.0;
  }
}

// dL/du* += a G lam ; dL/dK -= a (G lam)(G p) ; dL/dbeta = dL/dK * Dp/delta
__global__ void __launch_bounds__(256) k_adj_grad(AdjArgs A) {
  const Cell3 cq = cell_of(A.I, A.J, A.K);
  if (!cq.in) return;
  const int a = cq.a, b = cq.b, c = cq.c;
  const double lmean = A.S->npart[0];
  {
    const int64_t n = nidx(A, a, b, c), h = hidx(A, a, b, c);
    const int co[3] = {a, b, c};
    const int lim[3] = {A.I, A.J, A.K};
    const int64_t ns[3] = {1, A.I, (int64_t)A.I * A.J};
    const int64_t hs[3] = {1, A.SX, (int64_t)A.SX * A.SY};
    const double* ap[3] = {A.ee, A.ne, A.te};
    double* gs[3] = {A.gus, A.gvs, A.gws};
    const double dxyz[3] = {A.dxp[a], A.dyp[b], A.dzp[c]};
    const bool actP = A.active[n] != 0;
    const double lP = actP ? A.lam[n] - lmean : 0.0;
    const double pP = actP ? A.xp[n] - A.shift : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double area = dxyz[(d + 1) % 3] * dxyz[(d + 2) % 3];
      const bool last = co[d] == lim[d] - 1;
      const bool actN = !last && A.active[n + ns[d]] != 0;
      const double lN = actN ? A.lam[n + ns[d]] - lmean : 0.0;
      // + face of this cell
      gs[d][h] += area * ap[d][h] * (lN - lP);
      double gk = A.gk[d][n];
      if (actP && actN) {
        const double pN = A.xp[n + ns[d]] - A.shift;
        gk -= ap[d][h] * area * (lN - lP) * (pN - pP);
      }
      A.gk[d][n] = gk;
      if (A.gb[d]) A.gb[d][n] = gk * (A.dp[d][h] / A.dgrid[d][co[d]]);
      // the wall face on the minus side (index 0) of the first cell
      if (co[d] == 0) gs[d][h - hs[d]] = area * ap[d][h - hs[d]] * lP;
    }
  }
}

}  // namespace

sp_status adjoint(sp_ctx* c, const sp_adj_inputs& in, sp_adj_outputs& out, sp_stats* st) {
  const Dims& D = c->dm;
  const sp_inputs& fin = c->last_in;
  const dim3 fg = flat_grid(D.I, D.J, D.K), fb(64, 4, 1);
  AdjArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K; A.SX = D.SX; A.SY = D.SY; A.N = D.N;
  A.xp = c->x_ckpt;
  A.shift = c->S_host->shift;
  A.active = c->active;
  A.pocoef = c->pocoef;
  A.ctu = fin.cell_type_u; A.ctv = fin.cell_type_v; A.ctw = fin.cell_type_w;
  A.mu = fin.moex_u; A.mv = fin.moex_v; A.mw = fin.moex_w;
  for (int f = 0; f < 6; ++f) A.fl[f] = fin.bc_flag[f];
  A.gu = in.gu; A.gv = in.gv; A.gw = in.gw;
  A.gp = in.gp;
  const sp_params& P = c->prm;
  A.pin = -1;
  if (P.pin_i >= 1 && P.pin_j >= 1 && P.pin_k >= 1 && P.pin_i <= D.I && P.pin_j <= D.J &&
      P.pin_k <= D.K) {
    A.pin_a = P.pin_i - 1; A.pin_b = P.pin_j - 1; A.pin_c = P.pin_k - 1;
    A.pin = A.pin_a + (int64_t)D.I * (A.pin_b + (int64_t)D.J * A.pin_c);
  }
  A.ee = fin.ee_area; A.ne = fin.ne_area; A.te = fin.te_area;
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp;
  A.dp[0] = fin.dp_u; A.dp[1] = fin.dp_v; A.dp[2] = fin.dp_w;
  A.dgrid[0] = c->dxu; A.dgrid[1] = c->dyv; A.dgrid[2] = c->dzw;
  A.lam = c->x;
  A.pbar = c->pbar;
  A.gus = out.gu_star; A.gvs = out.gv_star; A.gws = out.gw_star;
  A.gk[0] = out.gk_e; A.gk[1] = out.gk_n; A.gk[2] = out.gk_t;
  A.gb[0] = out.gbeta_e; A.gb[1] = out.gbeta_n; A.gb[2] = out.gbeta_t;
  A.S = c->S;
  const size_t hb = sizeof(double) * (size_t)D.Nh;
  for (double* p : {out.gu_star, out.gv_star, out.gw_star}) {
    hipError_t e = hipMemsetAsync(p, 0, hb, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemsetAsync");
  }
  const double NB = (double)D.N;
  int slot;
  A.site = red_site(c, kSiteAdj);
  timer_begin(c, "adj_source",
              NB * (3 * 8 + 6 * 4 + 3 * 8 + 8 + 1 + 3 * 8 + 3 * 8 + 8 + (in.gp ? 8 : 0)), &slot);
  hipLaunchKernelGGL(k_adj_source, fg, fb, 0, c->stream, A);
  // one partial per 256 cells: fold them on many workgroups first
  const uint32_t nb = flat_blocks(D.I, D.J, D.K), chunk = 512, nf = (nb + chunk - 1) / chunk;
  const RedSite fold = red_site(c, kSiteFold);
  hipLaunchKernelGGL((k_fold_partials<3, 2u>), dim3(nf), dim3(256), 0, c->stream, A.site, fold, nb,
                     chunk);
  A.site = fold;
  hipLaunchKernelGGL(k_adj_source_fin, dim3(1), dim3(kFinThreads), 0, c->stream, A, nf);
  timer_end(c, slot);
  // Solve A lam = pbar (mean-projected when singular) with the forward's
  // operator and multigrid hierarchy until max |r|/V <= 1e-10 max |pbar|/V.
  // The adjoint source has other units than the divergence, so the
  // tolerance is relative to it; on a stiff two-phase operator the residual
  // of a source with a large net flux levels off near 1e-12 of the source
  // (earlier with the FP32 V-cycle than with the FP64 one), and gradients to
  // 1e-8 need far less.
  hipError_t e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "adjoint source");
  const double rel = 1e-10;
  sp_status s = solve(c, c->pbar, rel * c->S_host->src_max, st, true, false);
  if (s != SP_OK) return s;
  A.site = red_site(c, kSiteAdj2);
  timer_begin(c, "adj_lam_mean", NB * 9, &slot);
  hipLaunchKernelGGL(k_adj_lam_mean, fg, fb, 0, c->stream, A);
  hipLaunchKernelGGL((k_fold_partials<1, 0u>), dim3(nf), dim3(256), 0, c->stream, A.site, fold, nb,
                     chunk);
  A.site = fold;
  hipLaunchKernelGGL(k_adj_lam_mean_fin, dim3(1), dim3(kFinThreads), 0, c->stream, A, nf);
  timer_end(c, slot);
  timer_begin(c, "adj_grad", NB * (8 * 3 + 1 + 3 * 8 + 3 * 8 * 2 + 3 * 8 * 2), &slot);
  hipLaunchKernelGGL(k_adj_grad, fg, fb, 0, c->stream, A);
  timer_end(c, slot);
  return SP_OK;
}

}  // namespace sp
