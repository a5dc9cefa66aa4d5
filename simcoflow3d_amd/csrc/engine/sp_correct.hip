// Pressure field and velocity correction on the GPU.
//
//   gauge shift of a singular (all-Neumann) system,
//   DeltaPressureGetValues + DeltaPressureBoundaryCondition (ProjectionP.f90:578-643),
//   correction u = u* - (Pp(i+1)-Pp(i)) PoCoef(:,4) ... (ComputePUVW.f90:181-232),
//   the mass residual mres of Solver.f90:236-250 as the convergence gate.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "sp_device.h"
#include "sp_geom.h"
#include "sp_internal.h"

namespace sp {
namespace {

struct CorrArgs {
  int I, J, K, SX, SY;
  int64_t N;
  const double* x;
  const double* xd;  // refinement of x (non-singular systems), or nullptr
  const uint8_t* active;
  const double* pocoef;
  const int32_t *ctu, *ctv, *ctw, *mu, *mv, *mw;
  int32_t fl[6];
  const double *us, *vs, *ws;
  const double *ee, *ne, *te;
  const double *dxp, *dyp, *dzp;
  double *u, *v, *w, *Pp;
  int64_t pin;  // interior index of the gauge cell or -1
  RedSite site;
  Scalars* S;
};

// shift = x(pin) or mean of x over active cells (singular systems only)
__global__ void __launch_bounds__(256) k_shift(CorrArgs A) {
  const Col3 q = col_of(A.I, A.J, A.K);
  double v[1] = {0.0};
  if (q.in) {
    const int64_t IJ = (int64_t)A.I * A.J;
    int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
#pragma unroll 4
    for (int c = q.c0; c < q.c1; ++c, n += IJ)
      if (A.active[n]) v[0] += A.x[n];
  }
  block_partials<kTileThreads, 1, 0u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_shift_fin(CorrArgs A, uint32_t nb) {
  double t[1];
  finalize_partials<kFinThreads, 1, 0u>(A.site, nb, t);
  if (flat_tid() == 0) {
    Scalars* S = A.S;
    double sh = 0.0;
    if (S->singular) {
      if (A.pin >= 0 && A.active[A.pin]) sh = A.x[A.pin];
      else if (S->active > 0) sh = t[0] / (double)S->active;
    }
    S->shift = sh;
  }
}

__device__ inline double pval(const CorrArgs& A, int64_t n, double shift) {
  if (!A.active[n]) return 0.0;
  return A.xd ? (A.x[n] - shift) + A.xd[n] : A.x[n] - shift;
}

// Correction of the faces on the + side of every cell, and the interior of
// Proj%Pp (DeltaPressureGetValues, ProjectionP.f90:591-613).  Each thread
// handles KC cells of a column.  Every load of a cell is issued up front at
// an address that is always valid (a neighbour past the domain edge reads
// the cell itself), and the reference's conditions select among the loaded
// values: loads behind data-dependent branches would cost a memory latency
// per branch level.
template <int KC>
__global__ void __launch_bounds__(256) k_correct(CorrArgs A) {
#pragma clang fp contract(off)
  const Col3 q = col_of<KC>(A.I, A.J, A.K);
  if (!q.in) return;
  const int a = q.a, b = q.b;
  const double shift = A.S->shift;
  const int64_t IJ = (int64_t)A.I * A.J, N = A.N, SXY = (int64_t)A.SX * A.SY;
  const bool hasE = a < A.I - 1, hasN = b < A.J - 1;
  const int64_t oE = hasE ? 1 : 0, oN = hasN ? A.I : 0;
  int64_t n = a + (int64_t)A.I * (b + (int64_t)A.J * q.c0);
  int64_t h = (a + 1) + (int64_t)A.SX * ((b + 1) + (int64_t)A.SY * (q.c0 + 1));
#pragma unroll
  for (int cc = 0; cc < KC; ++cc, n += IJ, h += SXY) {
    const int c = q.c0 + cc;
    if (c >= q.c1) break;
    const bool hasT = c < A.K - 1;
    const int64_t mE = n + oE, mN = n + oN, mT = hasT ? n + IJ : n;
    const bool aP = A.active[n] != 0, aE = A.active[mE] != 0, aN = A.active[mN] != 0,
               aT = A.active[mT] != 0;
    const double xP = A.x[n], xE = A.x[mE], xN = A.x[mN], xT = A.x[mT];
    double dP = 0.0, dE_ = 0.0, dN_ = 0.0, dT_ = 0.0;
    if (A.xd) {  // uniform
      dP = A.xd[n]; dE_ = A.xd[mE]; dN_ = A.xd[mN]; dT_ = A.xd[mT];
    }
    const int32_t mu = A.mu[n], ctu = A.ctu[n], mv = A.mv[n], ctv = A.ctv[n], mw = A.mw[n],
                  ctw = A.ctw[n];
    const double cE = A.pocoef[3 * N + n], cN = A.pocoef[4 * N + n], cT = A.pocoef[5 * N + n];
    const double usv = A.us[h], vsv = A.vs[h], wsv = A.ws[h];
    // pval: 0 in inactive cells, x - shift (+ the refinement) otherwise
    auto pv = [&](bool act, double xv, double dv) {
      return act ? (A.xd ? (xv - shift) + dv : xv - shift) : 0.0;
    };
    const double pP = pv(aP, xP, dP);
    // Pp(i+1)-Pp(i): between two active cells the gauge cancels, so the
    // difference is taken on the solver's iterate, whose stiff-phase values
    // sit near zero; faces touching inactive cells or walls use the output
    // field with its ghosts (ProjectionP.f90:591-643).
    auto dpres = [&](bool inside, bool am, double xm, double dm, int fl) {
      const double both = A.xd ? (xm - xP) + (dm - dP) : xm - xP;
      const double pN = inside ? pv(am, xm, dm) : (double)fl * pP;
      return (inside && aP && am) ? both : pN - pP;
    };
    const double dE = dpres(hasE, aE, xE, dE_, A.fl[1]);
    const double dN = dpres(hasN, aN, xN, dN_, A.fl[3]);
    const double dT = dpres(hasT, aT, xT, dT_, A.fl[5]);
    const double uu = (mu != 1 && ctu != 2) ? usv - dE * cE : 0.0;
    const double vv = (mv != 1 && ctv != 2) ? vsv - dN * cN : 0.0;
    const double ww = (mw != 1 && ctw != 2) ? wsv - dT * cT : 0.0;
    A.u[h] = uu;
    A.v[h] = vv;
    A.w[h] = ww;
    A.Pp[h] = pP;
  }
}
// cells per thread of k_correct
constexpr int kCorrK = 1;

// Ghost entries of Proj%Pp: faces flag*Pp(inner) (ProjectionP.f90:620-643),
// edges and corners 0.  blockIdx.z selects the face: 0/1 bottom/top planes
// (all i, j), 2/3 south/north (all i, interior k), 4/5 west/east (interior
// j, k).
__global__ void __launch_bounds__(256) k_pressure_ghosts(CorrArgs A) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y * blockDim.y + threadIdx.y;
  const int f = blockIdx.z;
  int i, j, k;
  if (f < 2) {
    if (u > A.I + 1 || w > A.J + 1) return;
    i = u; j = w; k = f == 0 ? 0 : A.K + 1;
  } else if (f < 4) {
    if (u > A.I + 1 || w >= A.K) return;
    i = u; j = f == 2 ? 0 : A.J + 1; k = w + 1;
  } else {
    if (u >= A.J || w >= A.K) return;
    i = f == 4 ? 0 : A.I + 1; j = u + 1; k = w + 1;
  }
  const double shift = A.S->shift;
  const bool ix = i >= 1 && i <= A.I, jx = j >= 1 && j <= A.J, kx = k >= 1 && k <= A.K;
  auto P = [&](int ii, int jj, int kk) {
    return pval(A, (ii - 1) + (int64_t)A.I * ((jj - 1) + (int64_t)A.J * (kk - 1)), shift);
  };
  double v = 0.0;
  if (!ix && jx && kx) v = (double)A.fl[i == 0 ? 0 : 1] * P(i == 0 ? 1 : A.I, j, k);
  else if (ix && !jx && kx) v = (double)A.fl[j == 0 ? 2 : 3] * P(i, j == 0 ? 1 : A.J, k);
  else if (ix && jx && !kx) v = (double)A.fl[k == 0 ? 4 : 5] * P(i, j, k == 0 ? 1 : A.K);
  A.Pp[i + (int64_t)A.SX * (j + (int64_t)A.SY * k)] = v;
}

// mres of the corrected field (Solver.f90:239-247); the west/south/bottom
// wall faces (index 0) take the predictor's value, which the correction does
// not touch.
// max |div u| / V over the cells; KC cells of a column per thread, every
// load of the KC cells issued before the arithmetic (the cell index clamped
// into the column, so no load waits on a test).
template <int KC>
__global__ void __launch_bounds__(256) k_divergence(CorrArgs A) {
#pragma clang fp contract(off)
  const Col3 q = col_of<KC>(A.I, A.J, A.K);
  const int64_t SXY = (int64_t)A.SX * A.SY;
  double v[1] = {0.0};
  if (q.in) {
    const int a = q.a, b = q.b;
    const double dx = A.dxp[a], dy = A.dyp[b];
    // the W / S faces of the first cells are the caller's u* / v* (walls)
    const double* uw_src = a > 0 ? A.u : A.us;
    const double* vs_src = b > 0 ? A.v : A.vs;
    const int64_t h0 = (a + 1) + (int64_t)A.SX * ((b + 1) + (int64_t)A.SY * (q.c0 + 1));
    double uP[KC], uW[KC], vP[KC], vS[KC], wP[KC], wB[KC], eP[KC], eW[KC], nP[KC], nS[KC],
        tP[KC], tB[KC], dz[KC];
#pragma unroll
    for (int cc = 0; cc < KC; ++cc) {
      const int c = min(q.c0 + cc, q.c1 - 1);
      const int64_t h = h0 + SXY * (c - q.c0);
      const double* wb_src = c > 0 ? A.w : A.ws;
      uP[cc] = A.u[h]; uW[cc] = uw_src[h - 1];
      vP[cc] = A.v[h]; vS[cc] = vs_src[h - A.SX];
      wP[cc] = A.w[h]; wB[cc] = wb_src[h - SXY];
      eP[cc] = A.ee[h]; eW[cc] = A.ee[h - 1];
      nP[cc] = A.ne[h]; nS[cc] = A.ne[h - A.SX];
      tP[cc] = A.te[h]; tB[cc] = A.te[h - SXY];
      dz[cc] = A.dzp[c];
    }
#pragma unroll
    for (int cc = 0; cc < KC; ++cc) {
      if (q.c0 + cc >= q.c1) break;
      const double r = dy * dz[cc] * (eP[cc] * uP[cc] - eW[cc] * uW[cc]) +
                       dx * dz[cc] * (nP[cc] * vP[cc] - nS[cc] * vS[cc]) +
                       dx * dy * (tP[cc] * wP[cc] - tB[cc] * wB[cc]);
      double m = fabs(r) / (dx * dy * dz[cc]);
      if (!(m <= 1.0e300)) m = 1.0e300;  // NaN / Inf
      v[0] = m > v[0] ? m : v[0];
    }
  }
  block_partials<kTileThreads, 1, 1u>(A.site, v);
}
// cells per thread of k_divergence
constexpr int kDivK = 4;

__global__ void __launch_bounds__(kFinThreads) k_divergence_fin(CorrArgs A, uint32_t nb) {
  double t[1];
  finalize_partials<kFinThreads, 1, 1u>(A.site, nb, t);
  if (flat_tid() == 0) A.S->divmax = t[0];
}

}  // namespace

sp_status correct(sp_ctx* c, const sp_inputs& in, sp_outputs& out, sp_stats* st) {
  const Dims& D = c->dm;
  CorrArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K; A.SX = D.SX; A.SY = D.SY; A.N = D.N;
  const dim3 fb(64, 4, 1);
  A.x = c->x; A.active = c->active; A.pocoef = c->pocoef;
  A.xd = c->xd_active ? c->xd : nullptr;
  A.ctu = in.cell_type_u; A.ctv = in.cell_type_v; A.ctw = in.cell_type_w;
  A.mu = in.moex_u; A.mv = in.moex_v; A.mw = in.moex_w;
  for (int f = 0; f < 6; ++f) A.fl[f] = in.bc_flag[f];
  A.us = in.u_star; A.vs = in.v_star; A.ws = in.w_star;
  A.ee = in.ee_area; A.ne = in.ne_area; A.te = in.te_area;
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp;
  A.u = out.u; A.v = out.v; A.w = out.w; A.Pp = out.p;
  const sp_params& P = c->prm;
  if (P.pin_i >= 1 && P.pin_j >= 1 && P.pin_k >= 1 && P.pin_i <= D.I && P.pin_j <= D.J &&
      P.pin_k <= D.K)
    A.pin = (P.pin_i - 1) + (int64_t)D.I * ((P.pin_j - 1) + (int64_t)D.J * (P.pin_k - 1));
  else
    A.pin = -1;
  A.S = c->S;
  const double NB = (double)D.N;
  const dim3 cg = col_grid(D.I, D.J, D.K);
  const uint32_t ncol = col_blocks(D.I, D.J, D.K);
  int slot;
  // gauge: only singular systems are shifted (S->shift is 0 otherwise)
  if (c->S_host->singular) {
    A.site = red_site(c, kSiteShift);
    timer_begin(c, "shift", NB * 9, &slot);
    hipLaunchKernelGGL(k_shift, cg, fb, 0, c->stream, A);
    timer_end(c, slot);
    hipLaunchKernelGGL(k_shift_fin, dim3(1), dim3(kFinThreads), 0, c->stream, A, ncol);
  }
  // x, mask, masks x6, PoCoef x3, u*, v*, w* -> u, v, w, Pp
  timer_begin(c, "correct", NB * (8 + 1 + 6 * 4 + 3 * 8 + 3 * 8 + 3 * 8 + 8), &slot);
  hipLaunchKernelGGL(k_correct<kCorrK>, col_grid_k<kCorrK>(D.I, D.J, D.K), fb, 0, c->stream, A);
  timer_end(c, slot);
  {
    const int gx = std::max(D.I + 2, D.J), gy = std::max(D.J + 2, D.K);
    dim3 blk(64, 4, 1), grd((gx + 63) / 64, (gy + 3) / 4, 6);
    hipLaunchKernelGGL(k_pressure_ghosts, grd, blk, 0, c->stream, A);
  }
  A.site = red_site(c, kSiteDiv);
  timer_begin(c, "divergence", NB * 6 * 8, &slot);
  hipLaunchKernelGGL(k_divergence<kDivK>, col_grid_k<kDivK>(D.I, D.J, D.K), fb, 0, c->stream, A);
  hipLaunchKernelGGL(k_divergence_fin, dim3(1), dim3(kFinThreads), 0, c->stream, A,
                     col_blocks_k<kDivK>(D.I, D.J, D.K));
  timer_end(c, slot);
  (void)st;
  return SP_OK;
}

}  // namespace sp
