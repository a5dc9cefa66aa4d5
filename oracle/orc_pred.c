/*
 * CPU FP64 restatement of the momentum predictor PredictingUVW
 * (src/PredictUVW.f90:49-803) for SpaceOrder 1 and 2 -- TEST INFRASTRUCTURE ONLY
 * (see orc.h).  Compiled with -ffp-contract=off: every expression keeps the
 * reference's operation order, so the engine's kernels can match it bit for
 * bit.
 *
 * What the reference does, and what is restated here:
 *   - convective face fluxes, first-order upwind with the cut-cell
 *     corrections, ModifiedConvectiveFluxFirstOrder (PredictUVW.f90:1338-1794),
 *     with the solid velocity 0 (:166-168);
 *   - DiffusiveFlux computes coefficients and then sets them all to zero
 *     (:3672-3673), so the diffusive terms, the Crank-Nicolson part of the
 *     right-hand side and every off-diagonal matrix coefficient vanish: the
 *     three HYPRE systems (:692-759) are diagonal and their solution is
 *     rhs / aP (HYPRE's PCG returns that up to the rounding of its one step,
 *     which no test here can see: parity unpinned below 1e-15 relative);
 *   - the momentum source: FluxDiv from the face fluxes, Adams-Bashforth
 *     with FluxDivOld for TimeOrder 3 (:320-337), the wall friction Sp
 *     (:404-410), aP and Dp (:1025-1037), the right-hand side and the Dp
 *     scaling (:1206-1241), the Dp halo copies (:778-800), the predictor
 *     boundary conditions (:3796-3894) and the solid-cell condition
 *     (:3896-3916).
 * Reads the reference leaves undefined (a cell array at index Jmax + 1 in
 * two cut-cell branches, :1416-1430 and :1744-1751) are read as "no such
 * neighbour", as the sibling branches that guard the index do.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define HX(g, i, j, k) ((int64_t)(i) + (int64_t)((g)->I + 2) * ((int64_t)(j) + (int64_t)((g)->J + 2) * (int64_t)(k)))
#define NX(g, i, j, k) ((int64_t)((i) - 1) + (int64_t)(g)->I * ((int64_t)((j) - 1) + (int64_t)(g)->J * (int64_t)((k) - 1)))

/* staggered grid spacings (Mesh.f90:160-247): U: dxu, dyp, dzp;
 * V: dxp, dyv, dzp; W: dxp, dyp, dzw (index = the cell's own index) */
#define UDX(g, i) ((g)->dxu[(i) - 1])
#define UDY(g, j) ((g)->dyp[(j) - 1])
#define UDZ(g, k) ((g)->dzp[(k) - 1])
#define VDX(g, i) ((g)->dxp[(i) - 1])
#define VDY(g, j) ((g)->dyv[(j) - 1])
#define VDZ(g, k) ((g)->dzp[(k) - 1])
#define WDX(g, i) ((g)->dxp[(i) - 1])
#define WDY(g, j) ((g)->dyp[(j) - 1])
#define WDZ(g, k) ((g)->dzw[(k) - 1])

/* face flux arrays: CFEW(0:I+1,J,K,3), CFNS(I,0:J+1,K,3), CFTB(I,J,0:K+1,3) */
static inline int64_t ew_ix(const orc_grid* g, int i, int j, int k, int c) {
  return (int64_t)i + (int64_t)(g->I + 2) * ((j - 1) + (int64_t)g->J * (k - 1)) +
         (int64_t)(g->I + 2) * g->J * g->K * (c - 1);
}
static inline int64_t ns_ix(const orc_grid* g, int i, int j, int k, int c) {
  return (int64_t)(i - 1) + (int64_t)g->I * (j + (int64_t)(g->J + 2) * (k - 1)) +
         (int64_t)g->I * (g->J + 2) * g->K * (c - 1);
}
static inline int64_t tb_ix(const orc_grid* g, int i, int j, int k, int c) {
  return (int64_t)(i - 1) + (int64_t)g->I * ((j - 1) + (int64_t)g->J * k) +
         (int64_t)g->I * g->J * (g->K + 2) * (c - 1);
}

static const double kEps = 1.0e-20;  /* local epsi of the flux routine (:1362) */

/* |Cell_Cent . n + phi| of cell (i,j,k) */
static inline double plane_dist(const orc_grid* g, const orc_pred_cell* C, int i, int j, int k) {
  const int64_t n = NX(g, i, j, k), N = (int64_t)g->I * g->J * g->K;
  return fabs(C->cent[n] * C->nx[n] + C->cent[n + N] * C->ny[n] + C->cent[n + 2 * N] * C->nz[n] +
              C->phi[n]);
}

/* cell (i,j,k) may serve as the upwind neighbour of a cut face: not a
 * momentum-exchange cell, not solid */
static inline int usable(const orc_grid* g, const orc_pred_cell* C, int i, int j, int k) {
  const int64_t n = NX(g, i, j, k);
  return C->moex[n] != 1 && C->cell_type[n] != 2;
}

static inline double vofc(const orc_grid* g, const orc_pred_cell* C, int i, int j, int k) {
  return C->vof[NX(g, i, j, k)];
}

/* face value of a halo x3 array */
static inline double f3(const orc_grid* g, const double* a, int i, int j, int k, int c) {
  const int64_t Nh = (int64_t)(g->I + 2) * (g->J + 2) * (g->K + 2);
  return a[HX(g, i, j, k) + Nh * (c - 1)];
}

static inline double cent(const orc_grid* g, const orc_pred_cell* C, int i, int j, int k, int c) {
  const int64_t N = (int64_t)g->I * g->J * g->K;
  return C->cent[NX(g, i, j, k) + N * (c - 1)];
}

/* BC face values: q = 6*component + face (W,E,S,N,B,T); W/E (J,K), S/N (I,K),
 * B/T (I,J) */
static inline double bcval(const orc_grid* g, const double* const* bcv, int q, int a, int b) {
  const int f = q % 6;
  const int n1 = f < 2 ? g->J : g->I;
  return bcv[q][(a - 1) + (int64_t)n1 * (b - 1)];
}

void orc_conv_flux_first(const orc_grid* g, const double* pee, const double* pne,
                         const double* pte, const orc_pred_cell* U, const orc_pred_cell* V,
                         const orc_pred_cell* W, const double* un, const double* vn,
                         const double* wn, const double* const* bcv, double* cfew, double* cfns,
                         double* cftb) {
  const int I = g->I, J = g->J, K = g->K;
  const double uS = 0.0, vS = 0.0, wS = 0.0; /* uSolid, vSolid, wSolid (:166-168) */
#define H(i, j, k) HX(g, i, j, k)
  /* ---- iu = 1: faces x_{i-1/2}, i = 1..I+1 (:1366-1525) ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I + 1; ++i) {
        double uw, uwp, uwn, eta, flux;
        /* u momentum, convective u (:1367-1389) */
        uw = 0.5 * (un[H(i, j, k)] + un[H(i - 1, j, k)]);
        uwp = 0.5 * (uw + fabs(uw));
        uwn = 0.5 * (uw - fabs(uw));
        if (i == 1) {
          flux = (uwn * un[H(i, j, k)] + uwp * un[H(i - 1, j, k)]) * U->ee[H(i, j, k)] * UDY(g, j) *
                 UDZ(g, k);
        } else if (i >= I) {
          flux = (uwn * un[H(i, j, k)] + uwp * un[H(i - 1, j, k)]) * U->ee[H(i - 1, j, k)] *
                 UDY(g, j) * UDZ(g, k);
        } else {
          eta = U->etae[H(i - 1, j, k)];
          uw = (1.0 - eta) * un[H(i - 1, j, k)] + eta * un[H(i, j, k)];
          uw = (uw - uS) * U->ale[H(i - 1, j, k)] + uS;
          uwp = 0.5 * (uw + fabs(uw));
          uwn = 0.5 * (uw - fabs(uw));
          flux = (uwn * un[H(i, j, k)] + uwp * un[H(i - 1, j, k)]) * U->ee[H(i - 1, j, k)] *
                 UDY(g, j) * UDZ(g, k);
        }
        cfew[ew_ix(g, i, j, k, 1)] = flux;
        /* v momentum, convective u (:1391-1448) */
        uw = 0.5 * (un[H(i - 1, j + 1, k)] + un[H(i - 1, j, k)]);
        if (i > I) {
          flux = uw * bcval(g, bcv, 6 + 1, j, k) * V->ee[H(i - 1, j, k)] * VDY(g, j) * VDZ(g, k);
        } else if (i == 1) {
          flux = uw * bcval(g, bcv, 6 + 0, j, k) * V->ee[H(i, j, k)] * VDY(g, j) * VDZ(g, k);
        } else {
          if (pee[H(i - 1, j, k)] <= 1.0e-5 || pee[H(i - 1, j + 1, k)] <= 1.0e-5) {
            const double d1 = f3(g, V->fce, i - 1, j, k, 1) * V->nx[NX(g, i - 1, j, k)] +
                              f3(g, V->fce, i - 1, j, k, 2) * V->ny[NX(g, i - 1, j, k)] +
                              f3(g, V->fce, i - 1, j, k, 3) * V->nz[NX(g, i - 1, j, k)] +
                              V->phi[NX(g, i - 1, j, k)];
            const double d2 = (f3(g, V->fce, i - 1, j, k, 1) - UDX(g, i - 1)) * V->nx[NX(g, i, j, k)] +
                              f3(g, V->fce, i - 1, j, k, 2) * V->ny[NX(g, i, j, k)] +
                              f3(g, V->fce, i - 1, j, k, 3) * V->nz[NX(g, i, j, k)] +
                              V->phi[NX(g, i, j, k)];
            const double dhec = 0.5 * (d1 + d2);
            double dh = dhec;
            /* both tests read U cells at j + 1, which do not exist for
             * j = Jmax: both read as false there */
            if (j < J && usable(g, U, i - 1, j + 1, k) &&
                vofc(g, U, i - 1, j + 1, k) >= vofc(g, U, i - 1, j, k)) {
              dh = plane_dist(g, U, i - 1, j + 1, k);
              uw = un[H(i - 1, j + 1, k)];
            } else if (j < J && usable(g, U, i - 1, j, k) &&
                       vofc(g, U, i - 1, j, k) > vofc(g, U, i - 1, j + 1, k)) {
              dh = plane_dist(g, U, i - 1, j, k);
              uw = un[H(i - 1, j, k)];
            } else {
              dh = dhec;
            }
            uw = (uw - uS) * dhec / (dh + kEps) + uS;
          } else {
            const double Sy = U->syn[H(i - 1, j, k)];
            eta = fabs(f3(g, V->fce, i - 1, j, k, 2) + VDY(g, j) / 2.0 - cent(g, U, i - 1, j, k, 2)) /
                  Sy;
            if (fabs(eta) >= 1.0) eta = 0.5;
            uw = (1.0 - eta) * un[H(i - 1, j, k)] + eta * un[H(i - 1, j + 1, k)];
          }
          uwp = 0.5 * (uw + fabs(uw));
          uwn = 0.5 * (uw - fabs(uw));
          flux = (uwp * vn[H(i - 1, j, k)] + uwn * vn[H(i, j, k)]) * V->ee[H(i - 1, j, k)] *
                 VDY(g, j) * VDZ(g, k);
        }
        cfew[ew_ix(g, i, j, k, 2)] = flux;
        /* w momentum, convective u (:1449-1506) */
        uw = 0.5 * (un[H(i - 1, j, k + 1)] + un[H(i - 1, j, k)]);
        if (i >= I) {
          flux = uw * bcval(g, bcv, 12 + 1, j, k) * W->ee[H(i - 1, j, k)] * WDY(g, j) * WDZ(g, k);
        } else if (i == 1) {
          flux = uw * bcval(g, bcv, 12 + 0, j, k) * W->ee[H(i, j, k)] * WDY(g, j) * WDZ(g, k);
        } else {
          if (pee[H(i - 1, j, k)] <= 1.0e-5 || pee[H(i - 1, j, k + 1)] <= 1.0e-5) {
            const double d1 = f3(g, W->fce, i - 1, j, k, 1) * W->nx[NX(g, i - 1, j, k)] +
                              f3(g, W->fce, i - 1, j, k, 2) * W->ny[NX(g, i - 1, j, k)] +
                              f3(g, W->fce, i - 1, j, k, 3) * W->nz[NX(g, i - 1, j, k)] +
                              W->phi[NX(g, i - 1, j, k)];
            const double d2 = (f3(g, W->fce, i - 1, j, k, 1) - UDX(g, i - 1)) * W->nx[NX(g, i, j, k)] +
                              f3(g, W->fce, i - 1, j, k, 2) * W->ny[NX(g, i, j, k)] +
                              f3(g, W->fce, i - 1, j, k, 3) * W->nz[NX(g, i, j, k)] +
                              W->phi[NX(g, i, j, k)];
            const double dhec = 0.5 * (d1 + d2);
            double dh;
            if (k < K) {
              if (usable(g, U, i - 1, j, k + 1) &&
                  vofc(g, U, i - 1, j, k + 1) >= vofc(g, U, i - 1, j, k)) {
                dh = plane_dist(g, U, i - 1, j, k + 1);
                uw = un[H(i - 1, j, k + 1)];
              } else if (usable(g, U, i - 1, j, k) &&
                         vofc(g, U, i - 1, j, k) > vofc(g, U, i - 1, j, k + 1)) {
                dh = plane_dist(g, U, i - 1, j, k);
                uw = un[H(i - 1, j, k)];
              } else {
                dh = dhec;
              }
            } else {
              dh = dhec;
            }
            /* the scaling is applied twice (:1492-1494) */
            uw = uw * dhec / (dh + kEps);
            uw = (uw - uS) * dhec / (dh + kEps) + uS;
          } else {
            const double Sz = U->szt[H(i - 1, j, k)];
            eta = fabs(f3(g, W->fce, i - 1, j, k, 3) + WDZ(g, k) / 2.0 - cent(g, U, i - 1, j, k, 3)) /
                  Sz;
            if (fabs(eta) >= 1.0) eta = 0.5;
            uw = (1.0 - eta) * un[H(i - 1, j, k)] + eta * un[H(i - 1, j, k + 1)];
          }
          uwp = 0.5 * (uw + fabs(uw));
          uwn = 0.5 * (uw - fabs(uw));
          flux = (uwp * wn[H(i - 1, j, k)] + uwn * wn[H(i, j, k)]) * W->ee[H(i - 1, j, k)] *
                 WDY(g, j) * WDZ(g, k);
        }
        cfew[ew_ix(g, i, j, k, 3)] = flux;
      }
  /* ---- iv = 1: faces y_{j-1/2}, j = 1..J+1 (:1527-1643) ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J + 1; ++j)
      for (int i = 1; i <= I; ++i) {
        double vs, vsp, vsn, eta, flux;
        /* u momentum, convective v (:1529-1567) */
        vs = 0.5 * (vn[H(i, j - 1, k)] + vn[H(i + 1, j - 1, k)]);
        if (j >= J) {
          flux = vs * bcval(g, bcv, 0 + 3, i, k) * U->ne[H(i, j - 1, k)] * UDX(g, i) * UDZ(g, k);
        } else if (j == 1) {
          flux = vs * bcval(g, bcv, 0 + 2, i, k) * U->ne[H(i, j, k)] * UDX(g, i) * UDZ(g, k);
        } else {
          if (pne[H(i, j - 1, k)] <= 1.0e-5 || pne[H(i + 1, j - 1, k)] <= 1.0e-5) {
            const double d1 = f3(g, U->fcn, i, j - 1, k, 1) * U->nx[NX(g, i, j - 1, k)] +
                              f3(g, U->fcn, i, j - 1, k, 2) * U->ny[NX(g, i, j - 1, k)] +
                              f3(g, U->fcn, i, j - 1, k, 3) * U->nz[NX(g, i, j - 1, k)] +
                              U->phi[NX(g, i, j - 1, k)];
            const double d2 = f3(g, U->fcn, i, j - 1, k, 1) * U->nx[NX(g, i, j, k)] +
                              (f3(g, U->fcn, i, j - 1, k, 2) - VDY(g, j - 1)) * U->ny[NX(g, i, j, k)] +
                              f3(g, U->fcn, i, j - 1, k, 3) * U->nz[NX(g, i, j, k)] +
                              U->phi[NX(g, i, j, k)];
            const double dhec = 0.5 * (d1 + d2);
            double dh;
            if (i < I) {
              if (usable(g, V, i, j - 1, k) && vofc(g, V, i, j - 1, k) >= vofc(g, V, i + 1, j - 1, k)) {
                dh = plane_dist(g, V, i, j - 1, k);
                vs = vn[H(i, j - 1, k)];
              } else if (usable(g, V, i + 1, j - 1, k) &&
                         vofc(g, V, i + 1, j - 1, k) > vofc(g, V, i, j - 1, k)) {
                dh = plane_dist(g, V, i + 1, j - 1, k);
                vs = vn[H(i + 1, j - 1, k)];
              } else {
                dh = dhec;
              }
            } else {
              dh = dhec;
            }
            vs = (vs - vS) * dhec / (dh + kEps) + vS;
          } else {
            const double Sx = V->sxe[H(i, j - 1, k)];
            eta = fabs(f3(g, U->fcn, i, j - 1, k, 1) + 0.5 * UDX(g, i) - cent(g, V, i, j - 1, k, 1)) /
                  Sx;
            if (fabs(eta) >= 1.0) eta = 0.5;
            vs = (1.0 - eta) * vn[H(i, j - 1, k)] + eta * vn[H(i + 1, j - 1, k)];
          }
          vsp = 0.5 * (vs + fabs(vs));
          vsn = 0.5 * (vs - fabs(vs));
          flux = (vsn * un[H(i, j, k)] + vsp * un[H(i, j - 1, k)]) * U->ne[H(i, j - 1, k)] *
                 UDX(g, i) * UDZ(g, k);
        }
        cfns[ns_ix(g, i, j, k, 1)] = flux;
        /* v momentum, convective v (:1569-1585): the interior branch forms
         * its interpolated vs but the flux keeps the plain average's parts */
        vs = 0.5 * (vn[H(i, j, k)] + vn[H(i, j - 1, k)]);
        vsp = 0.5 * (vs + fabs(vs));
        vsn = 0.5 * (vs - fabs(vs));
        if (j >= J) {
          flux = (vsn * vn[H(i, j, k)] + vsp * vn[H(i, j - 1, k)]) * V->ne[H(i, j - 1, k)] *
                 VDX(g, i) * VDZ(g, k);
        } else if (j == 1) {
          flux = (vsn * vn[H(i, j, k)] + vsp * vn[H(i, j - 1, k)]) * V->ne[H(i, j, k)] * VDX(g, i) *
                 VDZ(g, k);
        } else {
          flux = (vsn * vn[H(i, j, k)] + vsp * vn[H(i, j - 1, k)]) * V->ne[H(i, j - 1, k)] *
                 VDX(g, i) * VDZ(g, k);
        }
        cfns[ns_ix(g, i, j, k, 2)] = flux;
        /* w momentum, convective v (:1587-1642) */
        vs = 0.5 * (vn[H(i, j - 1, k)] + vn[H(i, j - 1, k + 1)]);
        if (j >= J) {
          flux = vs * bcval(g, bcv, 12 + 3, i, k) * W->ne[H(i, j - 1, k)] * WDX(g, i) * WDZ(g, k);
        } else if (j == 1) {
          flux = vs * bcval(g, bcv, 12 + 2, i, k) * W->ne[H(i, j, k)] * WDX(g, i) * WDZ(g, k);
        } else {
          if (pne[H(i, j - 1, k)] <= 1.0e-5 || pne[H(i, j - 1, k + 1)] <= 1.0e-5) {
            const double d1 = f3(g, W->fcn, i, j - 1, k, 1) * W->nx[NX(g, i, j - 1, k)] +
                              f3(g, W->fcn, i, j - 1, k, 2) * W->ny[NX(g, i, j - 1, k)] +
                              f3(g, W->fcn, i, j - 1, k, 3) * W->nz[NX(g, i, j - 1, k)] +
                              W->phi[NX(g, i, j - 1, k)];
            const double d2 = f3(g, W->fcn, i, j - 1, k, 1) * W->nx[NX(g, i, j, k)] +
                              (f3(g, W->fcn, i, j - 1, k, 2) - WDY(g, j - 1)) * W->ny[NX(g, i, j, k)] +
                              f3(g, W->fcn, i, j - 1, k, 3) * W->nz[NX(g, i, j, k)] +
                              W->phi[NX(g, i, j, k)];
            const double dhec = 0.5 * (d1 + d2);
            double dh;
            if (k < K) {
              if (usable(g, V, i, j - 1, k) && vofc(g, V, i, j - 1, k) >= vofc(g, V, i, j - 1, k + 1)) {
                dh = plane_dist(g, V, i, j - 1, k);
                vs = vn[H(i, j - 1, k)];
              } else if (usable(g, V, i, j - 1, k + 1) &&
                         vofc(g, V, i, j - 1, k + 1) > vofc(g, V, i, j - 1, k)) {
                dh = plane_dist(g, V, i, j - 1, k + 1);
                vs = vn[H(i, j - 1, k + 1)];
              } else {
                dh = dhec;
              }
            } else {
              dh = dhec;
            }
            vs = (vs - vS) * dhec / (dh + kEps) + vS;
          } else {
            const double Sz = V->szt[H(i, j - 1, k)];
            eta = fabs(f3(g, W->fcn, i, j - 1, k, 3) + 0.5 * WDZ(g, k) - cent(g, W, i, j - 1, k, 3)) /
                  Sz;
            if (fabs(eta) >= 1.0) eta = 0.5;
            vs = (1.0 - eta) * vn[H(i, j - 1, k)] + eta * vn[H(i, j - 1, k + 1)];
          }
          vsp = 0.5 * (vs + fabs(vs));
          vsn = 0.5 * (vs - fabs(vs));
          flux = (vsp * wn[H(i, j - 1, k)] + vsn * wn[H(i, j, k)]) * W->ne[H(i, j - 1, k)] *
                 WDX(g, i) * WDZ(g, k);
        }
        cfns[ns_ix(g, i, j, k, 3)] = flux;
      }
  /* ---- iw = 1: faces z_{k-1/2}, k = 1..K+1 (:1644-1790) ---- */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K + 1; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        double wb, wbp, wbn, eta, flux;
        /* u momentum, convective w: the average takes vn12(i+1,j,k-1)
         * (:1647), and the bottom boundary BCu%VarT (:1654) */
        wb = 0.5 * (wn[H(i, j, k - 1)] + vn[H(i + 1, j, k - 1)]);
        if (k >= K) {
          flux = wb * bcval(g, bcv, 0 + 5, i, j) * U->te[H(i, j, k - 1)] * UDX(g, i) * UDY(g, j);
        } else if (k == 1) {
          flux = wb * bcval(g, bcv, 0 + 5, i, j) * U->te[H(i, j, k)] * UDX(g, i) * UDY(g, j);
        } else {
          if (pte[H(i, j, k - 1)] <= 1.0e-5 || pte[H(i + 1, j, k - 1)] <= 1.0e-5) {
            const double d1 = f3(g, U->fct, i, j, k - 1, 1) * U->nx[NX(g, i, j, k - 1)] +
                              f3(g, U->fct, i, j, k - 1, 2) * U->ny[NX(g, i, j, k - 1)] +
                              f3(g, U->fct, i, j, k - 1, 3) * U->nz[NX(g, i, j, k - 1)] +
                              U->phi[NX(g, i, j, k - 1)];
            /* (FCT(:,2) - dz) * ny, as :1662-1663 write it */
            const double d2 = f3(g, U->fct, i, j, k - 1, 1) * U->nx[NX(g, i, j, k)] +
                              (f3(g, U->fct, i, j, k - 1, 2) - WDZ(g, k - 1)) * U->ny[NX(g, i, j, k)] +
                              f3(g, U->fct, i, j, k - 1, 3) * U->nz[NX(g, i, j, k)] +
                              U->phi[NX(g, i, j, k)];
            const double dhec = 0.5 * (d1 + d2);
            double dh;
            if (i < I) {
              if (usable(g, W, i, j, k - 1) && vofc(g, W, i, j, k - 1) >= vofc(g, W, i + 1, j, k - 1)) {
                dh = plane_dist(g, W, i, j, k - 1);
                wb = wn[H(i, j, k - 1)];
              } else if (usable(g, W, i + 1, j, k - 1) &&
                         vofc(g, W, i + 1, j, k - 1) > vofc(g, W, i, j, k - 1)) {
                dh = plane_dist(g, W, i + 1, j, k - 1);
                wb = wn[H(i + 1, j, k - 1)];
              } else {
                dh = dhec;
              }
            } else {
              dh = dhec;
            }
            wb = (wb - wS) * dhec / (dh + kEps) + wS;
          } else {
            const double Sx = W->sxe[H(i, j, k - 1)];
            eta = fabs(f3(g, U->fct, i, j, k - 1, 1) + 0.5 * UDX(g, i) - cent(g, W, i, j, k - 1, 1)) /
                  Sx;
            if (fabs(eta) >= 1.0) eta = 0.5;
            wb = (1.0 - eta) * wn[H(i, j, k - 1)] + eta * wn[H(i + 1, j, k - 1)];
          }
          wbp = 0.5 * (wb + fabs(wb));
          wbn = 0.5 * (wb - fabs(wb));
          flux = (wbp * un[H(i, j, k - 1)] + wbn * un[H(i, j, k)]) * U->te[H(i, j, k - 1)] *
                 UDX(g, i) * UDY(g, j);
        }
        cftb[tb_ix(g, i, j, k, 1)] = flux;
        /* v momentum, convective w (:1711-1765) */
        wb = 0.5 * (wn[H(i, j + 1, k - 1)] + wn[H(i, j, k - 1)]);
        if (k >= K) {
          flux = wb * bcval(g, bcv, 6 + 5, i, j) * V->te[H(i, j, k - 1)] * VDX(g, i) * VDY(g, j);
        } else if (k == 1) {
          flux = wb * bcval(g, bcv, 6 + 4, i, j) * V->te[H(i, j, k)] * VDX(g, i) * VDY(g, j);
        } else {
          if (pte[H(i, j, k - 1)] <= 1.0e-5 || pte[H(i, j + 1, k - 1)] <= 1.0e-5) {
            const double d1 = f3(g, V->fct, i, j, k - 1, 1) * V->nx[NX(g, i, j, k - 1)] +
                              f3(g, V->fct, i, j, k - 1, 2) * V->ny[NX(g, i, j, k - 1)] +
                              f3(g, V->fct, i, j, k - 1, 3) * V->nz[NX(g, i, j, k - 1)] +
                              V->phi[NX(g, i, j, k - 1)];
            /* (FCT(:,1) - dz) * nx, as :1722-1723 write it */
            const double d2 = (f3(g, V->fct, i, j, k - 1, 1) - WDZ(g, k - 1)) * V->nx[NX(g, i, j, k)] +
                              f3(g, V->fct, i, j, k - 1, 2) * V->ny[NX(g, i, j, k)] +
                              f3(g, V->fct, i, j, k - 1, 3) * V->nz[NX(g, i, j, k)] +
                              V->phi[NX(g, i, j, k)];
            const double dhec = 0.5 * (d1 + d2);
            double dh = dhec;
            if (j < J && usable(g, W, i, j + 1, k - 1) &&
                vofc(g, W, i, j + 1, k - 1) >= vofc(g, W, i, j, k - 1)) {
              dh = plane_dist(g, W, i, j + 1, k - 1);
              wb = wn[H(i, j + 1, k - 1)];
            }
            /* the second test reads WCell%vof(i,j+1,k-1), which has no cell at
             * j = Jmax: read as unavailable there */
            if (j < J && usable(g, W, i, j, k - 1) &&
                vofc(g, W, i, j, k - 1) > vofc(g, W, i, j + 1, k - 1)) {
              dh = plane_dist(g, W, i, j, k - 1);
              wb = wn[H(i, j, k - 1)];
            }
            wb = (wb - wS) * dhec / (dh + kEps) + wS;
          } else {
            const double Sy = W->syn[H(i, j, k - 1)];
            eta = fabs(f3(g, V->fct, i, j, k - 1, 2) + 0.5 * VDY(g, j) - cent(g, W, i, j, k - 1, 2)) /
                  Sy;
            if (fabs(eta) >= 1.0) eta = 0.5;
            wb = (1.0 - eta) * wn[H(i, j, k - 1)] + eta * wn[H(i, j + 1, k - 1)];
          }
          wbp = 0.5 * (wb + fabs(wb));
          wbn = 0.5 * (wb - fabs(wb));
          flux = (wbp * vn[H(i, j, k - 1)] + wbn * vn[H(i, j, k)]) * V->te[H(i, j, k - 1)] *
                 VDX(g, i) * VDY(g, j);
        }
        cftb[tb_ix(g, i, j, k, 2)] = flux;
        /* w momentum, convective w (:1767-1785) */
        wb = 0.5 * (wn[H(i, j, k)] + wn[H(i, j, k - 1)]);
        wbp = 0.5 * (wb + fabs(wb));
        wbn = 0.5 * (wb - fabs(wb));
        if (k >= K) {
          flux = (wbp * wn[H(i, j, k - 1)] + wbn * wn[H(i, j, k)]) * W->te[H(i, j, k - 1)] *
                 WDX(g, i) * WDY(g, j);
        } else if (k == 1) {
          flux = (wbp * wn[H(i, j, k - 1)] + wbn * wn[H(i, j, k)]) * W->te[H(i, j, k)] * WDX(g, i) *
                 WDY(g, j);
        } else {
          eta = W->etat[H(i, j, k - 1)];
          wb = (1.0 - eta) * wn[H(i, j, k - 1)] + eta * wn[H(i, j, k)];
          wb = (wb - wS) * W->alt[H(i, j, k - 1)] + wS;
          wbp = 0.5 * (wb + fabs(wb));
          wbn = 0.5 * (wb - fabs(wb));
          flux = (wbp * wn[H(i, j, k - 1)] + wbn * wn[H(i, j, k)]) * W->te[H(i, j, k - 1)] *
                 WDX(g, i) * WDY(g, j);
        }
        cftb[tb_ix(g, i, j, k, 3)] = flux;
      }
#undef H
}

/* ---- SpaceOrder 2: ModifiedConvectiveFlux (PredictUVW.f90:2253-2644) ----
 *
 * Central interpolation with the cut-cell corrections.  The reference loops
 * i outermost and k innermost (:2277-2279) and keeps its local delh between
 * iterations: a cut-face branch whose two neighbour tests both fail
 * (:2326-2341 and its siblings) divides by the delh of the last branch that
 * set one in this call.  That carry is restated here as the loop does it;
 * before any branch of the call has set delh the reference reads an
 * undefined value, for which this restatement takes the face's own delhec
 * (the first-order routine's choice, :1415).  x**2.d0 is x*x (the exact
 * square GCC makes of pow(x, 2.0)). */

typedef struct delh_carry {
  int have;
  double v;
} delh_carry;

/* how often the last call read a carried delh, and how often it had none
 * to read (test instrumentation) */
static int64_t g_carried = 0, g_uncarried = 0;
void orc_conv_flux_second_carries(int64_t* carried, int64_t* uncarried) {
  *carried = g_carried;
  *uncarried = g_uncarried;
}

/* the two neighbour tests of a cut face: cell a first (>=), then cell b
 * (>), each setting delh and the convective velocity; `avail` false where
 * the cells lie beyond the arrays (read as no neighbour, as the first-order
 * routine's guards do) */
static void cut_tests(const orc_grid* g, const orc_pred_cell* C, int avail, int ai, int aj, int ak,
                      int bi, int bj, int bk, const double* vel, double* cv, double dhec,
                      delh_carry* dc, int init_to_dhec) {
  double delh;
  int set = 0;
  if (init_to_dhec) { delh = dhec; set = 1; }
  if (avail && usable(g, C, ai, aj, ak) && vofc(g, C, ai, aj, ak) >= vofc(g, C, bi, bj, bk)) {
    delh = plane_dist(g, C, ai, aj, ak);
    *cv = vel[HX(g, ai, aj, ak)];
    set = 1;
  }
  if (avail && usable(g, C, bi, bj, bk) && vofc(g, C, bi, bj, bk) > vofc(g, C, ai, aj, ak)) {
    delh = plane_dist(g, C, bi, bj, bk);
    *cv = vel[HX(g, bi, bj, bk)];
    set = 1;
  }
  if (set) {
    dc->have = 1;
    dc->v = delh;
  } else {
    delh = dc->have ? dc->v : dhec;
    if (dc->have) ++g_carried;
    else ++g_uncarried;
  }
  *cv = *cv * dhec / (delh + kEps);
}

/* |eta| >= 1 -> 0.5 */
static inline double clamp_eta(double eta) { return fabs(eta) >= 1.0 ? 0.5 : eta; }

void orc_conv_flux_second(const orc_grid* g, const double* pee, const double* pne,
                          const double* pte, const orc_pred_cell* U, const orc_pred_cell* V,
                          const orc_pred_cell* W, const double* un, const double* vn,
                          const double* wn, const double* const* bcv, double* cfew, double* cfns,
                          double* cftb) {
  const int I = g->I, J = g->J, K = g->K;
  g_carried = g_uncarried = 0;
#define H(i, j, k) HX(g, i, j, k)
  /* ---- iu = 1 (:2280-2408) ---- */
  {
    delh_carry dc = {0, 0.0};
    for (int i = 1; i <= I + 1; ++i)
      for (int j = 1; j <= J; ++j)
        for (int k = 1; k <= K; ++k) {
          double uw, eta, flux;
          /* u momentum, convective u (:2281-2298) */
          uw = 0.5 * (un[H(i, j, k)] + un[H(i - 1, j, k)]);
          if (i == 1) {
            flux = uw * uw * U->ee[H(i, j, k)] * UDY(g, j) * UDZ(g, k);
          } else if (i >= I) {
            flux = uw * uw * U->ee[H(i - 1, j, k)] * UDY(g, j) * UDZ(g, k);
          } else {
            eta = U->etae[H(i - 1, j, k)];
            uw = (1.0 - eta) * un[H(i - 1, j, k)] + eta * un[H(i, j, k)];
            const double a = uw * U->ale[H(i - 1, j, k)];
            flux = a * a * U->ee[H(i - 1, j, k)] * UDY(g, j) * UDZ(g, k);
          }
          cfew[ew_ix(g, i, j, k, 1)] = flux;
          /* v momentum, convective u (:2299-2357) */
          uw = 0.5 * (un[H(i - 1, j + 1, k)] + un[H(i - 1, j, k)]);
          if (i > I) {
            flux = uw * bcval(g, bcv, 6 + 1, j, k) * V->ee[H(i - 1, j, k)] * VDY(g, j) * VDZ(g, k);
          } else if (i == 1) {
            flux = uw * bcval(g, bcv, 6 + 0, j, k) * V->ee[H(i, j, k)] * VDY(g, j) * VDZ(g, k);
          } else {
            if (pee[H(i - 1, j, k)] <= 1.0e-5 || pee[H(i - 1, j + 1, k)] <= 1.0e-5) {
              const double d1 = f3(g, V->fce, i - 1, j, k, 1) * V->nx[NX(g, i - 1, j, k)] +
                                f3(g, V->fce, i - 1, j, k, 2) * V->ny[NX(g, i - 1, j, k)] +
                                f3(g, V->fce, i - 1, j, k, 3) * V->nz[NX(g, i - 1, j, k)] +
                                V->phi[NX(g, i - 1, j, k)];
              const double d2 = (f3(g, V->fce, i - 1, j, k, 1) - UDX(g, i - 1)) * V->nx[NX(g, i, j, k)] +
                                f3(g, V->fce, i - 1, j, k, 2) * V->ny[NX(g, i, j, k)] +
                                f3(g, V->fce, i - 1, j, k, 3) * V->nz[NX(g, i, j, k)] +
                                V->phi[NX(g, i, j, k)];
              cut_tests(g, U, j < J, i - 1, j + 1, k, i - 1, j, k, un, &uw, 0.5 * (d1 + d2), &dc, 0);
            } else {
              const double Sy = U->syn[H(i - 1, j, k)];
              eta = clamp_eta(fabs(f3(g, V->fce, i - 1, j, k, 2) + VDY(g, j) / 2.0 -
                                   cent(g, U, i - 1, j, k, 2)) / Sy);
              uw = (1.0 - eta) * un[H(i - 1, j, k)] + eta * un[H(i - 1, j + 1, k)];
            }
            eta = V->etae[H(i - 1, j, k)];
            const double vw = (1.0 - eta) * vn[H(i - 1, j, k)] + eta * vn[H(i, j, k)];
            flux = uw * vw * V->ale[H(i - 1, j, k)] * V->ee[H(i - 1, j, k)] * VDY(g, j) * VDZ(g, k);
          }
          cfew[ew_ix(g, i, j, k, 2)] = flux;
          /* w momentum, convective u (:2358-2407) */
          uw = 0.5 * (un[H(i - 1, j, k + 1)] + un[H(i - 1, j, k)]);
          if (i > I) {
            flux = uw * bcval(g, bcv, 12 + 1, j, k) * W->ee[H(i - 1, j, k)] * WDY(g, j) * WDZ(g, k);
          } else if (i == 1) {
            flux = uw * bcval(g, bcv, 12 + 0, j, k) * W->ee[H(i, j, k)] * WDY(g, j) * WDZ(g, k);
          } else {
            if (pee[H(i - 1, j, k)] <= 1.0e-5 || pee[H(i - 1, j, k + 1)] <= 1.0e-5) {
              const double d1 = f3(g, W->fce, i - 1, j, k, 1) * W->nx[NX(g, i - 1, j, k)] +
                                f3(g, W->fce, i - 1, j, k, 2) * W->ny[NX(g, i - 1, j, k)] +
                                f3(g, W->fce, i - 1, j, k, 3) * W->nz[NX(g, i - 1, j, k)] +
                                W->phi[NX(g, i - 1, j, k)];
              const double d2 = (f3(g, W->fce, i - 1, j, k, 1) - UDX(g, i - 1)) * W->nx[NX(g, i, j, k)] +
                                f3(g, W->fce, i - 1, j, k, 2) * W->ny[NX(g, i, j, k)] +
                                f3(g, W->fce, i - 1, j, k, 3) * W->nz[NX(g, i, j, k)] +
                                W->phi[NX(g, i, j, k)];
              cut_tests(g, U, k < K, i - 1, j, k + 1, i - 1, j, k, un, &uw, 0.5 * (d1 + d2), &dc, 0);
            } else {
              const double Sz = U->szt[H(i - 1, j, k)];
              eta = clamp_eta(fabs(f3(g, W->fce, i - 1, j, k, 3) + WDZ(g, k) / 2.0 -
                                   cent(g, U, i - 1, j, k, 3)) / Sz);
              uw = (1.0 - eta) * un[H(i - 1, j, k)] + eta * un[H(i - 1, j, k + 1)];
            }
            eta = W->etae[H(i - 1, j, k)];
            const double ww = (1.0 - eta) * wn[H(i - 1, j, k)] + eta * wn[H(i, j, k)];
            flux = uw * ww * W->ale[H(i - 1, j, k)] * W->ee[H(i - 1, j, k)] * WDY(g, j) * WDZ(g, k);
          }
          cfew[ew_ix(g, i, j, k, 3)] = flux;
        }
  }
  /* ---- iv = 1 (:2409-2520) ---- */
  {
    delh_carry dc = {0, 0.0};
    for (int i = 1; i <= I; ++i)
      for (int j = 1; j <= J + 1; ++j)
        for (int k = 1; k <= K; ++k) {
          double vs, eta, flux;
          /* u momentum, convective v (:2410-2457) */
          vs = 0.5 * (vn[H(i, j - 1, k)] + vn[H(i + 1, j - 1, k)]);
          if (j > J) {
            flux = vs * bcval(g, bcv, 3, i, k) * U->ne[H(i, j - 1, k)] * UDX(g, i) * UDZ(g, k);
          } else if (j == 1) {
            flux = vs * bcval(g, bcv, 2, i, k) * U->ne[H(i, j, k)] * UDX(g, i) * UDZ(g, k);
          } else {
            if (pne[H(i, j - 1, k)] <= 1.0e-5 || pne[H(i + 1, j - 1, k)] <= 1.0e-5) {
              const double d1 = f3(g, U->fcn, i, j - 1, k, 1) * U->nx[NX(g, i, j - 1, k)] +
                                f3(g, U->fcn, i, j - 1, k, 2) * U->ny[NX(g, i, j - 1, k)] +
                                f3(g, U->fcn, i, j - 1, k, 3) * U->nz[NX(g, i, j - 1, k)] +
                                U->phi[NX(g, i, j - 1, k)];
              const double d2 = f3(g, U->fcn, i, j - 1, k, 1) * U->nx[NX(g, i, j, k)] +
                                (f3(g, U->fcn, i, j - 1, k, 2) - VDY(g, j - 1)) * U->ny[NX(g, i, j, k)] +
                                f3(g, U->fcn, i, j - 1, k, 3) * U->nz[NX(g, i, j, k)] +
                                U->phi[NX(g, i, j, k)];
              cut_tests(g, V, i < I, i, j - 1, k, i + 1, j - 1, k, vn, &vs, 0.5 * (d1 + d2), &dc, 0);
            } else {
              const double Sx = V->sxe[H(i, j - 1, k)];
              eta = clamp_eta(fabs(f3(g, U->fcn, i, j - 1, k, 1) + 0.5 * UDX(g, i) -
                                   cent(g, V, i, j - 1, k, 1)) / Sx);
              vs = (1.0 - eta) * vn[H(i, j - 1, k)] + eta * vn[H(i + 1, j - 1, k)];
            }
            eta = U->etan[H(i, j - 1, k)];
            const double us = (1.0 - eta) * un[H(i, j - 1, k)] + eta * un[H(i, j, k)];
            flux = us * vs * U->aln[H(i, j - 1, k)] * U->ne[H(i, j - 1, k)] * UDX(g, i) * UDZ(g, k);
          }
          cfns[ns_ix(g, i, j, k, 1)] = flux;
          /* v momentum, convective v (:2458-2471) */
          vs = 0.5 * (vn[H(i, j, k)] + vn[H(i, j - 1, k)]);
          if (j >= J) {
            flux = vs * vs * V->ne[H(i, j - 1, k)] * VDX(g, i) * VDZ(g, k);
          } else if (j == 1) {
            flux = vs * vs * V->ne[H(i, j, k)] * VDX(g, i) * VDZ(g, k);
          } else {
            eta = V->etan[H(i, j - 1, k)];
            vs = (1.0 - eta) * vn[H(i, j - 1, k)] + eta * vn[H(i, j, k)];
            const double a = vs * V->aln[H(i, j - 1, k)];
            flux = a * a * V->ne[H(i, j - 1, k)] * VDX(g, i) * VDZ(g, k);
          }
          cfns[ns_ix(g, i, j, k, 2)] = flux;
          /* w momentum, convective v (:2472-2519) */
          vs = 0.5 * (vn[H(i, j - 1, k)] + vn[H(i, j - 1, k + 1)]);
          if (j > J) {
            flux = vs * bcval(g, bcv, 15, i, k) * W->ne[H(i, j - 1, k)] * WDX(g, i) * WDZ(g, k);
          } else if (j == 1) {
            flux = vs * bcval(g, bcv, 14, i, k) * W->ne[H(i, j, k)] * WDX(g, i) * WDZ(g, k);
          } else {
            if (pne[H(i, j - 1, k)] <= 1.0e-5 || pne[H(i, j - 1, k + 1)] <= 1.0e-5) {
              const double d1 = f3(g, W->fcn, i, j - 1, k, 1) * W->nx[NX(g, i, j - 1, k)] +
                                f3(g, W->fcn, i, j - 1, k, 2) * W->ny[NX(g, i, j - 1, k)] +
                                f3(g, W->fcn, i, j - 1, k, 3) * W->nz[NX(g, i, j - 1, k)] +
                                W->phi[NX(g, i, j - 1, k)];
              const double d2 = f3(g, W->fcn, i, j - 1, k, 1) * W->nx[NX(g, i, j, k)] +
                                (f3(g, W->fcn, i, j - 1, k, 2) - WDY(g, j - 1)) * W->ny[NX(g, i, j, k)] +
                                f3(g, W->fcn, i, j - 1, k, 3) * W->nz[NX(g, i, j, k)] +
                                W->phi[NX(g, i, j, k)];
              cut_tests(g, V, k < K, i, j - 1, k, i, j - 1, k + 1, vn, &vs, 0.5 * (d1 + d2), &dc, 0);
            } else {
              const double Sz = V->szt[H(i, j - 1, k)];
              eta = clamp_eta(fabs(f3(g, W->fcn, i, j - 1, k, 3) + 0.5 * WDZ(g, k) -
                                   cent(g, W, i, j - 1, k, 3)) / Sz);
              vs = (1.0 - eta) * vn[H(i, j - 1, k)] + eta * vn[H(i, j - 1, k + 1)];
            }
            eta = W->etan[H(i, j - 1, k)];
            const double ws = (1.0 - eta) * wn[H(i, j - 1, k)] + eta * wn[H(i, j, k)];
            flux = vs * ws * W->aln[H(i, j - 1, k)] * W->ne[H(i, j - 1, k)] * WDX(g, i) * WDZ(g, k);
          }
          cfns[ns_ix(g, i, j, k, 3)] = flux;
        }
  }
  /* ---- iw = 1 (:2521-2639) ---- */
  {
    delh_carry dc = {0, 0.0};
    for (int i = 1; i <= I; ++i)
      for (int j = 1; j <= J; ++j)
        for (int k = 1; k <= K + 1; ++k) {
          double wb, eta, flux;
          /* u momentum, convective w: the average takes vn12(i+1,j,k-1)
           * (:2523) */
          wb = 0.5 * (wn[H(i, j, k - 1)] + vn[H(i + 1, j, k - 1)]);
          if (k > K) {
            flux = wb * bcval(g, bcv, 5, i, j) * U->te[H(i, j, k - 1)] * UDX(g, i) * UDY(g, j);
          } else if (k == 1) {
            flux = wb * bcval(g, bcv, 4, i, j) * U->te[H(i, j, k)] * UDX(g, i) * UDY(g, j);
          } else {
            if (pte[H(i, j, k - 1)] <= 1.0e-5 || pte[H(i + 1, j, k - 1)] <= 1.0e-5) {
              /* (FCT(:,2) - dz) * ny, as :2537-2538 write it */
              const double d1 = f3(g, U->fct, i, j, k - 1, 1) * U->nx[NX(g, i, j, k - 1)] +
                                f3(g, U->fct, i, j, k - 1, 2) * U->ny[NX(g, i, j, k - 1)] +
                                f3(g, U->fct, i, j, k - 1, 3) * U->nz[NX(g, i, j, k - 1)] +
                                U->phi[NX(g, i, j, k - 1)];
              const double d2 = f3(g, U->fct, i, j, k - 1, 1) * U->nx[NX(g, i, j, k)] +
                                (f3(g, U->fct, i, j, k - 1, 2) - WDZ(g, k - 1)) * U->ny[NX(g, i, j, k)] +
                                f3(g, U->fct, i, j, k - 1, 3) * U->nz[NX(g, i, j, k)] +
                                U->phi[NX(g, i, j, k)];
              cut_tests(g, W, i < I, i, j, k - 1, i + 1, j, k - 1, wn, &wb, 0.5 * (d1 + d2), &dc, 0);
            } else {
              const double Sx = W->sxe[H(i, j, k - 1)];
              eta = clamp_eta(fabs(f3(g, U->fct, i, j, k - 1, 1) + 0.5 * UDX(g, i) -
                                   cent(g, W, i, j, k - 1, 1)) / Sx);
              wb = (1.0 - eta) * wn[H(i, j, k - 1)] + eta * wn[H(i + 1, j, k - 1)];
            }
            eta = U->etat[H(i, j, k - 1)];
            const double ub = (1.0 - eta) * un[H(i, j, k - 1)] + eta * un[H(i, j, k)];
            flux = wb * ub * U->alt[H(i, j, k - 1)] * U->te[H(i, j, k - 1)] * UDX(g, i) * UDY(g, j);
          }
          cftb[tb_ix(g, i, j, k, 1)] = flux;
          /* v momentum, convective w: delh = delhec before the tests
           * (:2596) */
          wb = 0.5 * (wn[H(i, j + 1, k - 1)] + wn[H(i, j, k - 1)]);
          if (k > K) {
            flux = wb * bcval(g, bcv, 11, i, j) * V->te[H(i, j, k - 1)] * VDX(g, i) * VDY(g, j);
          } else if (k == 1) {
            flux = wb * bcval(g, bcv, 10, i, j) * V->te[H(i, j, k)] * VDX(g, i) * VDY(g, j);
          } else {
            if (pte[H(i, j, k - 1)] <= 1.0e-5 || pte[H(i, j + 1, k - 1)] <= 1.0e-5) {
              /* (FCT(:,1) - dz) * nx, as :2592-2593 write it */
              const double d1 = f3(g, V->fct, i, j, k - 1, 1) * V->nx[NX(g, i, j, k - 1)] +
                                f3(g, V->fct, i, j, k - 1, 2) * V->ny[NX(g, i, j, k - 1)] +
                                f3(g, V->fct, i, j, k - 1, 3) * V->nz[NX(g, i, j, k - 1)] +
                                V->phi[NX(g, i, j, k - 1)];
              const double d2 = (f3(g, V->fct, i, j, k - 1, 1) - WDZ(g, k - 1)) * V->nx[NX(g, i, j, k)] +
                                f3(g, V->fct, i, j, k - 1, 2) * V->ny[NX(g, i, j, k)] +
                                f3(g, V->fct, i, j, k - 1, 3) * V->nz[NX(g, i, j, k)] +
                                V->phi[NX(g, i, j, k)];
              cut_tests(g, W, j < J, i, j + 1, k - 1, i, j, k - 1, wn, &wb, 0.5 * (d1 + d2), &dc, 1);
            } else {
              const double Sy = W->syn[H(i, j, k - 1)];
              eta = clamp_eta(fabs(f3(g, V->fct, i, j, k - 1, 2) + 0.5 * VDY(g, j) -
                                   cent(g, W, i, j, k - 1, 2)) / Sy);
              wb = (1.0 - eta) * wn[H(i, j, k - 1)] + eta * wn[H(i, j + 1, k - 1)];
            }
            eta = V->etat[H(i, j, k - 1)];
            const double vb = (1.0 - eta) * vn[H(i, j, k - 1)] + eta * vn[H(i, j, k)];
            flux = wb * vb * V->alt[H(i, j, k - 1)] * V->te[H(i, j, k - 1)] * VDX(g, i) * VDY(g, j);
          }
          cftb[tb_ix(g, i, j, k, 2)] = flux;
          /* w momentum, convective w (:2625-2638) */
          wb = 0.5 * (wn[H(i, j, k)] + wn[H(i, j, k - 1)]);
          if (k >= K) {
            flux = wb * wb * W->te[H(i, j, k - 1)] * WDX(g, i) * WDY(g, j);
          } else if (k == 1) {
            flux = wb * wb * W->te[H(i, j, k)] * WDX(g, i) * WDY(g, j);
          } else {
            eta = W->etat[H(i, j, k - 1)];
            wb = (1.0 - eta) * wn[H(i, j, k - 1)] + eta * wn[H(i, j, k)];
            const double a = wb * W->alt[H(i, j, k - 1)];
            flux = a * a * W->te[H(i, j, k - 1)] * WDX(g, i) * WDY(g, j);
          }
          cftb[tb_ix(g, i, j, k, 3)] = flux;
        }
  }
#undef H
}

/* Dp halo copies (:778-800): each statement copies a whole plane, in this
 * order, so the final value of a halo entry is the interior value at the
 * clamped index of every direction, with the extra plane Imax (U), Jmax (V)
 * or Kmax (W) copied from the one below. */
static void dp_halo(const orc_grid* g, double* dp, int iu, int iv, int iw) {
  const int I = g->I, J = g->J, K = g->K;
#define D(i, j, k) dp[HX(g, i, j, k)]
  for (int k = 0; k <= K + 1; ++k)
    for (int j = 0; j <= J + 1; ++j) D(0, j, k) = D(1, j, k);
  if (iu)
    for (int k = 0; k <= K + 1; ++k)
      for (int j = 0; j <= J + 1; ++j) D(I, j, k) = D(I - 1, j, k);
  for (int k = 0; k <= K + 1; ++k)
    for (int j = 0; j <= J + 1; ++j) D(I + 1, j, k) = D(I, j, k);
  for (int k = 0; k <= K + 1; ++k)
    for (int i = 0; i <= I + 1; ++i) D(i, 0, k) = D(i, 1, k);
  if (iv)
    for (int k = 0; k <= K + 1; ++k)
      for (int i = 0; i <= I + 1; ++i) D(i, J, k) = D(i, J - 1, k);
  for (int k = 0; k <= K + 1; ++k)
    for (int i = 0; i <= I + 1; ++i) D(i, J + 1, k) = D(i, J, k);
  for (int j = 0; j <= J + 1; ++j)
    for (int i = 0; i <= I + 1; ++i) D(i, j, 0) = D(i, j, 1);
  if (iw)
    for (int j = 0; j <= J + 1; ++j)
      for (int i = 0; i <= I + 1; ++i) D(i, j, K) = D(i, j, K - 1);
  for (int j = 0; j <= J + 1; ++j)
    for (int i = 0; i <= I + 1; ++i) D(i, j, K + 1) = D(i, j, K);
#undef D
}

int orc_predict(const orc_grid* g, const double* pee, const double* pne, const double* pte,
                const orc_pred_cell* const old[3], const orc_pred_cell* const cur[3],
                const double* const vel[3], const double* const vel_n[3], double* flux_div_old,
                const int32_t* bc_flag, const double* const* bcv, const orc_pred_params* p,
                double* const pred[3], double* const dp[3]) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K, Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  if (p->space_order != 1 && p->space_order != 2) return -1;
  /* un12 = TVar (TimeOrder 1, 3) or 1.5 TVar - 0.5 TVar_n (TimeOrder 2) (:149-159) */
  double* v12[3];
  for (int c = 0; c < 3; ++c) {
    v12[c] = (double*)malloc(sizeof(double) * Nh);
    for (int64_t q = 0; q < Nh; ++q)
      v12[c][q] = p->time_order == 2 ? 1.5 * vel[c][q] - 0.5 * vel_n[c][q] : vel[c][q];
  }
  double* cfew = (double*)calloc((size_t)(I + 2) * J * K * 3, sizeof(double));
  double* cfns = (double*)calloc((size_t)I * (J + 2) * K * 3, sizeof(double));
  double* cftb = (double*)calloc((size_t)I * J * (K + 2) * 3, sizeof(double));
  /* the first step always takes the first-order fluxes (:172-185) */
  if (p->itt == 1 || p->space_order == 1)
    orc_conv_flux_first(g, pee, pne, pte, old[0], old[1], old[2], v12[0], v12[1], v12[2], bcv,
                        cfew, cfns, cftb);
  else
    orc_conv_flux_second(g, pee, pne, pte, old[0], old[1], old[2], v12[0], v12[1], v12[2], bcv,
                         cfew, cfns, cftb);
  const double inv_fr2 = 1.0 / (p->fr * p->fr); /* 1.d0/Fr**2.d0 */
  for (int c = 0; c < 3; ++c) {
    const int iu = c == 0, iv = c == 1, iw = c == 2;
    const orc_pred_cell* C = cur[c];
    const orc_pred_cell* O = old[c];
    const double gc = p->gx * (double)iu + p->gy * (double)iv + p->gz * (double)iw;
    double* dpc = dp[c];
    double* pc = pred[c];
    for (int k = 1; k <= K - iw; ++k)
      for (int j = 1; j <= J - iv; ++j)
        for (int i = 1; i <= I - iu; ++i) {
          const int64_t n = NX(g, i, j, k), h = HX(g, i, j, k);
          /* flux divergence (:320-329) */
          const double Fe = cfew[ew_ix(g, i + 1, j, k, c + 1)];
          const double Fw = cfew[ew_ix(g, i, j, k, c + 1)];
          const double Fn = cfns[ns_ix(g, i, j + 1, k, c + 1)];
          const double Fs = cfns[ns_ix(g, i, j, k, c + 1)];
          const double Ft = cftb[tb_ix(g, i, j, k + 1, c + 1)];
          const double Fb = cftb[tb_ix(g, i, j, k, c + 1)];
          const double flux0 = Fe - Fw + Fn - Fs + Ft - Fb;
          double fdiv;
          if (p->itt == 1 || p->time_order == 1 || p->time_order == 2) fdiv = flux0;
          else fdiv = 1.5 * flux0 - 0.5 * flux_div_old[n + N * c];
          /* the diffusive part DFluxn0 is a sum of zero coefficients (:3672-3673) */
          flux_div_old[n + N * c] = flux0;
          const double dx = c == 0 ? UDX(g, i) : (c == 1 ? VDX(g, i) : WDX(g, i));
          const double dy = c == 0 ? UDY(g, j) : (c == 1 ? VDY(g, j) : WDY(g, j));
          const double dz = c == 0 ? UDZ(g, k) : (c == 1 ? VDZ(g, k) : WDZ(g, k));
          dpc[h] = 0.0;
          if (C->cell_type[n] == 2) {
            pc[h] = 0.0; /* DeltaGetValues: no unknown */
            continue;
          }
          /* wall friction (:404-410) */
          const double fr = (C->vof_l[n] / C->vof[n] * p->mu_w_ref +
                             (1.0 - C->vof_l[n] / C->vof[n]) * p->mu_a_ref) /
                            p->rey * C->wllh[n] / C->delh[n];
          /* aP with zero neighbour coefficients (:1025-1037) */
          const double aP = C->vof[n] * dx * dy * dz / p->dt + fr;
          dpc[h] = 1.0 / aP;
          /* right-hand side (:1206-1241) */
          const double s = iu * vel[0][h] + iv * vel[1][h] + iw * vel[2][h];
          double rhs = s * C->vof[n] * dx * dy * dz / p->dt -
                       s * (C->vof[n] - O->vof[n]) * dx * dy * dz / p->dt +
                       inv_fr2 * C->vof[n] * dx * dy * dz * gc / p->g;
          dpc[h] = C->vof[n] * dx * dy * dz * dpc[h];
          rhs = rhs - fdiv;
          pc[h] = rhs / aP;
        }
    dp_halo(g, dpc, iu, iv, iw);
  }
  /* PredictorVelocityBoundaryCondition (:3796-3894): flags of BCu (1..6 ->
   * 0..5), BCv (6..11), BCw (12..17) */
  double *pu = pred[0], *pv = pred[1], *pw = pred[2];
#define V3(a, i, j, k) (a)[HX(g, i, j, k)]
  for (int j = 1; j <= J; ++j)
    for (int k = 1; k <= K; ++k) {
      V3(pu, 0, j, k) = bc_flag[0] == 1 ? V3(pu, 1, j, k) : bcval(g, bcv, 0, j, k);
      if (bc_flag[1] == 1) V3(pu, I, j, k) = V3(pu, I - 1, j, k);
      else V3(pu, I, j, k) = bcval(g, bcv, 1, j, k);
      V3(pu, I + 1, j, k) = 2.0 * bcval(g, bcv, 1, j, k) - V3(pu, I - 1, j, k);
    }
  for (int i = 1; i <= I; ++i)
    for (int k = 1; k <= K; ++k) {
      V3(pv, i, 0, k) = bc_flag[8] == 1 ? V3(pv, i, 1, k) : bcval(g, bcv, 8, i, k);
      if (bc_flag[9] == 1) V3(pv, i, J, k) = V3(pv, i, J - 1, k);
      else V3(pv, i, J, k) = bcval(g, bcv, 9, i, k);
      V3(pv, i, J + 1, k) = 2.0 * bcval(g, bcv, 9, i, k) - V3(pv, i, J - 1, k);
    }
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j) {
      V3(pw, i, j, 0) = bc_flag[16] == 1 ? V3(pw, i, j, 1) : bcval(g, bcv, 16, i, j);
      if (bc_flag[17] == 1) V3(pw, i, j, K) = V3(pw, i, j, K - 1);
      else V3(pw, i, j, K) = bcval(g, bcv, 17, i, j);
      V3(pw, i, j, K + 1) = 2.0 * bcval(g, bcv, 17, i, j) - V3(pw, i, j, K - 1);
    }
  /* PredictorVelocityInternalCellCondition (:3896-3916) */
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k), h = HX(g, i, j, k);
        if (cur[0]->cell_type[n] == 2) pu[h] = 0.0;
        if (cur[1]->cell_type[n] == 2) pv[h] = 0.0;
        if (cur[2]->cell_type[n] == 2) pw[h] = 0.0;
      }
#undef V3
  free(cfew);
  free(cfns);
  free(cftb);
  for (int c = 0; c < 3; ++c) free(v12[c]);
  return 0;
}
