// The momentum predictor PredictingUVW (PredictUVW.f90:49-803) for
// SpaceOrder 1 and 2: u*, v*, w* and the d-coefficients Dp that the
// projection reads.
//
// DiffusiveFlux computes its coefficients and then zeroes them
// (PredictUVW.f90:3672-3673), so every neighbour coefficient of the three
// momentum systems vanishes and each HYPRE solve (:692-759) is the division
// rhs / aP of a diagonal system.  What remains is:
//
//   k_pred_faces  the first-order upwind convective fluxes with the cut-cell
//                 corrections, ModifiedConvectiveFluxFirstOrder
//                 (:1338-1794): one thread per (i,j,k) writes the face
//                 fluxes CFEW(i,j,k,1:3), CFNS(i,j,k,1:3), CFTB(i,j,k,1:3)
//                 it owns (x, y and z faces at index i, j, k);
//   k_pred_faces2 the SpaceOrder 2 fluxes, ModifiedConvectiveFlux
//                 (:2253-2644), one thread per k-column (below);
//   k_pred_cells  per U, V, W cell: the flux divergence, Adams-Bashforth
//                 for TimeOrder 3 (:320-337), FluxDivOld, the wall friction
//                 Sp (:404-410), aP and Dp (:1025-1037), the right-hand
//                 side and Dp scaling (:1206-1241) and u* = rhs / aP;
//   k_pred_dp_halo  the Dp halo copies (:778-800): the sequence of plane
//                 copies leaves every halo entry equal to the interior
//                 value at the clamped index (the component's own last
//                 plane copied from the one before it);
//   k_pred_bc     PredictorVelocityBoundaryCondition (:3796-3894) and then
//                 PredictorVelocityInternalCellCondition (:3896-3916) on the
//                 last plane of each component (the cells kernel zeroes the
//                 solid cells of the others).
//
// Every expression keeps the reference's operation order with contraction
// off, so the results equal a sequential evaluation bit for bit (the
// oracle's).  Two cut-cell tests of the reference read a cell array at
// index Jmax + 1 (:1416-1430, :1744-1751), outside the array; they are
// read as "no such neighbour", as the sibling branches that guard the index
// do.

#include <hip/hip_runtime.h>

#include "sp_device.h"
#include "sp_internal.h"

namespace sp {
namespace {

// The members of one U/V/W cell set the predictor reads (device pointers).
struct PredCell {
  const double *phi, *nx, *ny, *nz, *vof, *vof_l;
  const int32_t *ct, *moex;
  const double *ee, *ne, *te, *fce, *fcn, *fct, *cent;
  const double *sxe, *syn, *szt, *etae, *etan, *etat, *ale, *aln, *alt;
  const double *wllh, *delh;
};

struct PredArgs {
  int I, J, K, SX, SY;
  int64_t N, Nh;
  const double *dxp, *dyp, *dzp, *dxu, *dyv, *dzw;
  const double *pee, *pne, *pte;       // PCellO apertures (halo)
  PredCell old[3], cur[3];
  const double *vel[3], *veln[3];      // TVar, TVar_n (halo)
  int extrapolate;                     // TimeOrder 2: un12 = 1.5 TVar - 0.5 TVar_n
  const double* bcv[18];
  int32_t flag[18];
  double* fdo;                          // FluxDivOld (I,J,K,3)
  double *cfew, *cfns, *cftb;           // face fluxes
  double *pred[3], *dp[3];
  double dt, inv_fr2, gx, gy, gz, g, mu_w, mu_a, rey;
  int ab2;                              // Adams-Bashforth source (TimeOrder 3, itt > 1)
};

constexpr double kEps = 1.0e-20;  // local epsi of the flux routine (:1362)

struct Ix {
  int I, J, K, SX, SY;
  int64_t N, Nh;
  __device__ int64_t h(int i, int j, int k) const { return i + (int64_t)SX * (j + (int64_t)SY * k); }
  __device__ int64_t c(int i, int j, int k) const {
    return (i - 1) + (int64_t)I * ((j - 1) + (int64_t)J * (k - 1));
  }
  __device__ int64_t ew(int i, int j, int k, int q) const {
    return i + (int64_t)(I + 2) * ((j - 1) + (int64_t)J * (k - 1)) + (int64_t)(I + 2) * J * K * q;
  }
  __device__ int64_t ns(int i, int j, int k, int q) const {
    return (i - 1) + (int64_t)I * (j + (int64_t)(J + 2) * (k - 1)) + (int64_t)I * (J + 2) * K * q;
  }
  __device__ int64_t tb(int i, int j, int k, int q) const {
    return (i - 1) + (int64_t)I * ((j - 1) + (int64_t)J * k) + (int64_t)I * J * (K + 2) * q;
  }
};

__device__ inline Ix ix_of(const PredArgs& A) { return Ix{A.I, A.J, A.K, A.SX, A.SY, A.N, A.Nh}; }

// un12 of component q at halo index h
__device__ inline double v12(const PredArgs& A, int q, int64_t h) {
#pragma clang fp contract(off)
  return A.extrapolate ? 1.5 * A.vel[q][h] - 0.5 * A.veln[q][h] : A.vel[q][h];
}

__device__ inline double bcval(const PredArgs& A, int q, int a, int b) {
  const int f = q % 6;
  const int n1 = f < 2 ? A.J : A.I;
  return A.bcv[q][(a - 1) + (int64_t)n1 * (b - 1)];
}

__device__ inline double plane_dist(const Ix& X, const PredCell& C, int i, int j, int k) {
#pragma clang fp contract(off)
  const int64_t n = X.c(i, j, k);
  return fabs(C.cent[n] * C.nx[n] + C.cent[n + X.N] * C.ny[n] + C.cent[n + 2 * X.N] * C.nz[n] +
              C.phi[n]);
}

__device__ inline bool usable(const Ix& X, const PredCell& C, int i, int j, int k) {
  const int64_t n = X.c(i, j, k);
  return C.moex[n] != 1 && C.ct[n] != 2;
}

// |FC . n + phi| halves of a cut face: the face centre (halo x3 array fc at
// (fi,fj,fk)) against cell a's plane and, moved by `shift` along axis `ax`,
// against cell b's plane
__device__ inline double cut_dhec(const Ix& X, const PredCell& C, const double* fc, int fi, int fj,
                                  int fk, int ai, int aj, int ak, int bi, int bj, int bk, int ax,
                                  double shift) {
#pragma clang fp contract(off)
  const int64_t f = X.h(fi, fj, fk);
  const double f1 = fc[f], f2 = fc[f + X.Nh], f3 = fc[f + 2 * X.Nh];
  const int64_t na = X.c(ai, aj, ak), nb = X.c(bi, bj, bk);
  const double d1 = f1 * C.nx[na] + f2 * C.ny[na] + f3 * C.nz[na] + C.phi[na];
  const double g1 = ax == 0 ? f1 - shift : f1;
  const double g2 = ax == 1 ? f2 - shift : f2;
  const double d2 = g1 * C.nx[nb] + g2 * C.ny[nb] + f3 * C.nz[nb] + C.phi[nb];
  return 0.5 * (d1 + d2);
}

__device__ inline double up_p(double a) {
#pragma clang fp contract(off)
  return 0.5 * (a + fabs(a));
}
__device__ inline double up_n(double a) {
#pragma clang fp contract(off)
  return 0.5 * (a - fabs(a));
}

// ---- x faces: CFEW(i,j,k,1:3), i = 1..I+1 (:1366-1525) ----
__device__ void flux_x(const PredArgs& A, const Ix& X, int i, int j, int k) {
#pragma clang fp contract(off)
  const PredCell &U = A.old[0], &V = A.old[1], &W = A.old[2];
  const int I = A.I, J = A.J, K = A.K;
  const double dyu = A.dyp[j - 1], dzu = A.dzp[k - 1];
  double uw, uwp, uwn, eta, flux;
  // u momentum, convective u
  {
    const double ui = v12(A, 0, X.h(i, j, k)), um = v12(A, 0, X.h(i - 1, j, k));
    uw = 0.5 * (ui + um);
    uwp = up_p(uw);
    uwn = up_n(uw);
    if (i == 1) {
      flux = (uwn * ui + uwp * um) * U.ee[X.h(i, j, k)] * dyu * dzu;
    } else if (i >= I) {
      flux = (uwn * ui + uwp * um) * U.ee[X.h(i - 1, j, k)] * dyu * dzu;
    } else {
      eta = U.etae[X.h(i - 1, j, k)];
      uw = (1.0 - eta) * um + eta * ui;
      uw = (uw - 0.0) * U.ale[X.h(i - 1, j, k)] + 0.0;
      uwp = up_p(uw);
      uwn = up_n(uw);
      flux = (uwn * ui + uwp * um) * U.ee[X.h(i - 1, j, k)] * dyu * dzu;
    }
    A.cfew[X.ew(i, j, k, 0)] = flux;
  }
  // v momentum, convective u (V grid: dy = dyv, dz = dzp)
  {
    const double dyv = A.dyv[j - 1], dzv = A.dzp[k - 1];
    uw = 0.5 * (v12(A, 0, X.h(i - 1, j + 1, k)) + v12(A, 0, X.h(i - 1, j, k)));
    if (i > I) {
      flux = uw * bcval(A, 7, j, k) * V.ee[X.h(i - 1, j, k)] * dyv * dzv;
    } else if (i == 1) {
      flux = uw * bcval(A, 6, j, k) * V.ee[X.h(i, j, k)] * dyv * dzv;
    } else {
      if (A.pee[X.h(i - 1, j, k)] <= 1.0e-5 || A.pee[X.h(i - 1, j + 1, k)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, V, V.fce, i - 1, j, k, i - 1, j, k, i, j, k, 0,
                                     A.dxu[i - 2]);
        double dh;
        if (j < J && usable(X, U, i - 1, j + 1, k) &&
            U.vof[X.c(i - 1, j + 1, k)] >= U.vof[X.c(i - 1, j, k)]) {
          dh = plane_dist(X, U, i - 1, j + 1, k);
          uw = v12(A, 0, X.h(i - 1, j + 1, k));
        } else if (j < J && usable(X, U, i - 1, j, k) &&
                   U.vof[X.c(i - 1, j, k)] > U.vof[X.c(i - 1, j + 1, k)]) {
          dh = plane_dist(X, U, i - 1, j, k);
          uw = v12(A, 0, X.h(i - 1, j, k));
        } else {
          dh = dhec;
        }
        uw = (uw - 0.0) * dhec / (dh + kEps) + 0.0;
      } else {
        const double Sy = U.syn[X.h(i - 1, j, k)];
        eta = fabs(V.fce[X.h(i - 1, j, k) + X.Nh] + dyv / 2.0 - U.cent[X.c(i - 1, j, k) + X.N]) / Sy;
        if (fabs(eta) >= 1.0) eta = 0.5;
        uw = (1.0 - eta) * v12(A, 0, X.h(i - 1, j, k)) + eta * v12(A, 0, X.h(i - 1, j + 1, k));
      }
      uwp = up_p(uw);
      uwn = up_n(uw);
      flux = (uwp * v12(A, 1, X.h(i - 1, j, k)) + uwn * v12(A, 1, X.h(i, j, k))) *
             V.ee[X.h(i - 1, j, k)] * dyv * dzv;
    }
    A.cfew[X.ew(i, j, k, 1)] = flux;
  }
  // w momentum, convective u (W grid: dy = dyp, dz = dzw)
  {
    const double dyw = A.dyp[j - 1], dzw = A.dzw[k - 1];
    uw = 0.5 * (v12(A, 0, X.h(i - 1, j, k + 1)) + v12(A, 0, X.h(i - 1, j, k)));
    if (i >= I) {
      flux = uw * bcval(A, 13, j, k) * W.ee[X.h(i - 1, j, k)] * dyw * dzw;
    } else if (i == 1) {
      flux = uw * bcval(A, 12, j, k) * W.ee[X.h(i, j, k)] * dyw * dzw;
    } else {
      if (A.pee[X.h(i - 1, j, k)] <= 1.0e-5 || A.pee[X.h(i - 1, j, k + 1)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, W, W.fce, i - 1, j, k, i - 1, j, k, i, j, k, 0,
                                     A.dxu[i - 2]);
        double dh;
        if (k < K) {
          if (usable(X, U, i - 1, j, k + 1) &&
              U.vof[X.c(i - 1, j, k + 1)] >= U.vof[X.c(i - 1, j, k)]) {
            dh = plane_dist(X, U, i - 1, j, k + 1);
            uw = v12(A, 0, X.h(i - 1, j, k + 1));
          } else if (usable(X, U, i - 1, j, k) &&
                     U.vof[X.c(i - 1, j, k)] > U.vof[X.c(i - 1, j, k + 1)]) {
            dh = plane_dist(X, U, i - 1, j, k);
            uw = v12(A, 0, X.h(i - 1, j, k));
          } else {
            dh = dhec;
          }
        } else {
          dh = dhec;
        }
        // the scaling is applied twice (:1492-1494)
        uw = uw * dhec / (dh + kEps);
        uw = (uw - 0.0) * dhec / (dh + kEps) + 0.0;
      } else {
        const double Sz = U.szt[X.h(i - 1, j, k)];
        eta = fabs(W.fce[X.h(i - 1, j, k) + 2 * X.Nh] + dzw / 2.0 -
                   U.cent[X.c(i - 1, j, k) + 2 * X.N]) /
              Sz;
        if (fabs(eta) >= 1.0) eta = 0.5;
        uw = (1.0 - eta) * v12(A, 0, X.h(i - 1, j, k)) + eta * v12(A, 0, X.h(i - 1, j, k + 1));
      }
      uwp = up_p(uw);
      uwn = up_n(uw);
      flux = (uwp * v12(A, 2, X.h(i - 1, j, k)) + uwn * v12(A, 2, X.h(i, j, k))) *
             W.ee[X.h(i - 1, j, k)] * dyw * dzw;
    }
    A.cfew[X.ew(i, j, k, 2)] = flux;
  }
}

// ---- y faces: CFNS(i,j,k,1:3), j = 1..J+1 (:1527-1643) ----
__device__ void flux_y(const PredArgs& A, const Ix& X, int i, int j, int k) {
#pragma clang fp contract(off)
  const PredCell &U = A.old[0], &V = A.old[1], &W = A.old[2];
  const int I = A.I, J = A.J, K = A.K;
  double vs, vsp, vsn, eta, flux;
  // u momentum, convective v (U grid: dx = dxu, dz = dzp)
  {
    const double dxu = A.dxu[i - 1], dzu = A.dzp[k - 1];
    vs = 0.5 * (v12(A, 1, X.h(i, j - 1, k)) + v12(A, 1, X.h(i + 1, j - 1, k)));
    if (j >= J) {
      flux = vs * bcval(A, 3, i, k) * U.ne[X.h(i, j - 1, k)] * dxu * dzu;
    } else if (j == 1) {
      flux = vs * bcval(A, 2, i, k) * U.ne[X.h(i, j, k)] * dxu * dzu;
    } else {
      if (A.pne[X.h(i, j - 1, k)] <= 1.0e-5 || A.pne[X.h(i + 1, j - 1, k)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, U, U.fcn, i, j - 1, k, i, j - 1, k, i, j, k, 1,
                                     A.dyv[j - 2]);
        double dh;
        if (i < I) {
          if (usable(X, V, i, j - 1, k) && V.vof[X.c(i, j - 1, k)] >= V.vof[X.c(i + 1, j - 1, k)]) {
            dh = plane_dist(X, V, i, j - 1, k);
            vs = v12(A, 1, X.h(i, j - 1, k));
          } else if (usable(X, V, i + 1, j - 1, k) &&
                     V.vof[X.c(i + 1, j - 1, k)] > V.vof[X.c(i, j - 1, k)]) {
            dh = plane_dist(X, V, i + 1, j - 1, k);
            vs = v12(A, 1, X.h(i + 1, j - 1, k));
          } else {
            dh = dhec;
          }
        } else {
          dh = dhec;
        }
        vs = (vs - 0.0) * dhec / (dh + kEps) + 0.0;
      } else {
        const double Sx = V.sxe[X.h(i, j - 1, k)];
        eta = fabs(U.fcn[X.h(i, j - 1, k)] + 0.5 * dxu - V.cent[X.c(i, j - 1, k)]) / Sx;
        if (fabs(eta) >= 1.0) eta = 0.5;
        vs = (1.0 - eta) * v12(A, 1, X.h(i, j - 1, k)) + eta * v12(A, 1, X.h(i + 1, j - 1, k));
      }
      vsp = up_p(vs);
      vsn = up_n(vs);
      flux = (vsn * v12(A, 0, X.h(i, j, k)) + vsp * v12(A, 0, X.h(i, j - 1, k))) *
             U.ne[X.h(i, j - 1, k)] * dxu * dzu;
    }
    A.cfns[X.ns(i, j, k, 0)] = flux;
  }
  // v momentum, convective v: the interior branch's interpolated vs is not
  // used, the flux keeps the plain average's parts (:1569-1585)
  {
    const double dxv = A.dxp[i - 1], dzv = A.dzp[k - 1];
    const double vj = v12(A, 1, X.h(i, j, k)), vm = v12(A, 1, X.h(i, j - 1, k));
    vs = 0.5 * (vj + vm);
    vsp = up_p(vs);
    vsn = up_n(vs);
    if (j >= J) {
      flux = (vsn * vj + vsp * vm) * V.ne[X.h(i, j - 1, k)] * dxv * dzv;
    } else if (j == 1) {
      flux = (vsn * vj + vsp * vm) * V.ne[X.h(i, j, k)] * dxv * dzv;
    } else {
      flux = (vsn * vj + vsp * vm) * V.ne[X.h(i, j - 1, k)] * dxv * dzv;
    }
    A.cfns[X.ns(i, j, k, 1)] = flux;
  }
  // w momentum, convective v (W grid: dx = dxp, dz = dzw)
  {
    const double dxw = A.dxp[i - 1], dzw = A.dzw[k - 1];
    vs = 0.5 * (v12(A, 1, X.h(i, j - 1, k)) + v12(A, 1, X.h(i, j - 1, k + 1)));
    if (j >= J) {
      flux = vs * bcval(A, 15, i, k) * W.ne[X.h(i, j - 1, k)] * dxw * dzw;
    } else if (j == 1) {
      flux = vs * bcval(A, 14, i, k) * W.ne[X.h(i, j, k)] * dxw * dzw;
    } else {
      if (A.pne[X.h(i, j - 1, k)] <= 1.0e-5 || A.pne[X.h(i, j - 1, k + 1)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, W, W.fcn, i, j - 1, k, i, j - 1, k, i, j, k, 1,
                                     A.dyp[j - 2]);
        double dh;
        if (k < K) {
          if (usable(X, V, i, j - 1, k) && V.vof[X.c(i, j - 1, k)] >= V.vof[X.c(i, j - 1, k + 1)]) {
            dh = plane_dist(X, V, i, j - 1, k);
            vs = v12(A, 1, X.h(i, j - 1, k));
          } else if (usable(X, V, i, j - 1, k + 1) &&
                     V.vof[X.c(i, j - 1, k + 1)] > V.vof[X.c(i, j - 1, k)]) {
            dh = plane_dist(X, V, i, j - 1, k + 1);
            vs = v12(A, 1, X.h(i, j - 1, k + 1));
          } else {
            dh = dhec;
          }
        } else {
          dh = dhec;
        }
        vs = (vs - 0.0) * dhec / (dh + kEps) + 0.0;
      } else {
        const double Sz = V.szt[X.h(i, j - 1, k)];
        eta = fabs(W.fcn[X.h(i, j - 1, k) + 2 * X.Nh] + 0.5 * dzw -
                   W.cent[X.c(i, j - 1, k) + 2 * X.N]) /
              Sz;
        if (fabs(eta) >= 1.0) eta = 0.5;
        vs = (1.0 - eta) * v12(A, 1, X.h(i, j - 1, k)) + eta * v12(A, 1, X.h(i, j - 1, k + 1));
      }
      vsp = up_p(vs);
      vsn = up_n(vs);
      flux = (vsp * v12(A, 2, X.h(i, j - 1, k)) + vsn * v12(A, 2, X.h(i, j, k))) *
             W.ne[X.h(i, j - 1, k)] * dxw * dzw;
    }
    A.cfns[X.ns(i, j, k, 2)] = flux;
  }
}

// ---- z faces: CFTB(i,j,k,1:3), k = 1..K+1 (:1644-1790) ----
__device__ void flux_z(const PredArgs& A, const Ix& X, int i, int j, int k) {
#pragma clang fp contract(off)
  const PredCell &U = A.old[0], &V = A.old[1], &W = A.old[2];
  const int I = A.I, J = A.J, K = A.K;
  double wb, wbp, wbn, eta, flux;
  // u momentum, convective w: the average takes vn12(i+1,j,k-1) (:1647),
  // and the bottom boundary BCu%VarT (:1654)
  {
    const double dxu = A.dxu[i - 1], dyu = A.dyp[j - 1];
    wb = 0.5 * (v12(A, 2, X.h(i, j, k - 1)) + v12(A, 1, X.h(i + 1, j, k - 1)));
    if (k >= K) {
      flux = wb * bcval(A, 5, i, j) * U.te[X.h(i, j, k - 1)] * dxu * dyu;
    } else if (k == 1) {
      flux = wb * bcval(A, 5, i, j) * U.te[X.h(i, j, k)] * dxu * dyu;
    } else {
      if (A.pte[X.h(i, j, k - 1)] <= 1.0e-5 || A.pte[X.h(i + 1, j, k - 1)] <= 1.0e-5) {
        // (FCT(:,2) - dz) * ny, as :1662-1663 write it
        const double dhec = cut_dhec(X, U, U.fct, i, j, k - 1, i, j, k - 1, i, j, k, 1,
                                     A.dzw[k - 2]);
        double dh;
        if (i < I) {
          if (usable(X, W, i, j, k - 1) && W.vof[X.c(i, j, k - 1)] >= W.vof[X.c(i + 1, j, k - 1)]) {
            dh = plane_dist(X, W, i, j, k - 1);
            wb = v12(A, 2, X.h(i, j, k - 1));
          } else if (usable(X, W, i + 1, j, k - 1) &&
                     W.vof[X.c(i + 1, j, k - 1)] > W.vof[X.c(i, j, k - 1)]) {
            dh = plane_dist(X, W, i + 1, j, k - 1);
            wb = v12(A, 2, X.h(i + 1, j, k - 1));
          } else {
            dh = dhec;
          }
        } else {
          dh = dhec;
        }
        wb = (wb - 0.0) * dhec / (dh + kEps) + 0.0;
      } else {
        const double Sx = W.sxe[X.h(i, j, k - 1)];
        eta = fabs(U.fct[X.h(i, j, k - 1)] + 0.5 * dxu - W.cent[X.c(i, j, k - 1)]) / Sx;
        if (fabs(eta) >= 1.0) eta = 0.5;
        wb = (1.0 - eta) * v12(A, 2, X.h(i, j, k - 1)) + eta * v12(A, 2, X.h(i + 1, j, k - 1));
      }
      wbp = up_p(wb);
      wbn = up_n(wb);
      flux = (wbp * v12(A, 0, X.h(i, j, k - 1)) + wbn * v12(A, 0, X.h(i, j, k))) *
             U.te[X.h(i, j, k - 1)] * dxu * dyu;
    }
    A.cftb[X.tb(i, j, k, 0)] = flux;
  }
  // v momentum, convective w (V grid: dx = dxp, dy = dyv)
  {
    const double dxv = A.dxp[i - 1], dyv = A.dyv[j - 1];
    wb = 0.5 * (v12(A, 2, X.h(i, j + 1, k - 1)) + v12(A, 2, X.h(i, j, k - 1)));
    if (k >= K) {
      flux = wb * bcval(A, 11, i, j) * V.te[X.h(i, j, k - 1)] * dxv * dyv;
    } else if (k == 1) {
      flux = wb * bcval(A, 10, i, j) * V.te[X.h(i, j, k)] * dxv * dyv;
    } else {
      if (A.pte[X.h(i, j, k - 1)] <= 1.0e-5 || A.pte[X.h(i, j + 1, k - 1)] <= 1.0e-5) {
        // (FCT(:,1) - dz) * nx, as :1722-1723 write it
        const double dhec = cut_dhec(X, V, V.fct, i, j, k - 1, i, j, k - 1, i, j, k, 0,
                                     A.dzw[k - 2]);
        double dh = dhec;
        if (j < J && usable(X, W, i, j + 1, k - 1) &&
            W.vof[X.c(i, j + 1, k - 1)] >= W.vof[X.c(i, j, k - 1)]) {
          dh = plane_dist(X, W, i, j + 1, k - 1);
          wb = v12(A, 2, X.h(i, j + 1, k - 1));
        }
        if (j < J && usable(X, W, i, j, k - 1) &&
            W.vof[X.c(i, j, k - 1)] > W.vof[X.c(i, j + 1, k - 1)]) {
          dh = plane_dist(X, W, i, j, k - 1);
          wb = v12(A, 2, X.h(i, j, k - 1));
        }
        wb = (wb - 0.0) * dhec / (dh + kEps) + 0.0;
      } else {
        const double Sy = W.syn[X.h(i, j, k - 1)];
        eta = fabs(V.fct[X.h(i, j, k - 1) + X.Nh] + 0.5 * dyv - W.cent[X.c(i, j, k - 1) + X.N]) /
              Sy;
        if (fabs(eta) >= 1.0) eta = 0.5;
        wb = (1.0 - eta) * v12(A, 2, X.h(i, j, k - 1)) + eta * v12(A, 2, X.h(i, j + 1, k - 1));
      }
      wbp = up_p(wb);
      wbn = up_n(wb);
      flux = (wbp * v12(A, 1, X.h(i, j, k - 1)) + wbn * v12(A, 1, X.h(i, j, k))) *
             V.te[X.h(i, j, k - 1)] * dxv * dyv;
    }
    A.cftb[X.tb(i, j, k, 1)] = flux;
  }
  // w momentum, convective w (W grid: dx = dxp, dy = dyp)
  {
    const double dxw = A.dxp[i - 1], dyw = A.dyp[j - 1];
    const double wk = v12(A, 2, X.h(i, j, k)), wm = v12(A, 2, X.h(i, j, k - 1));
    wb = 0.5 * (wk + wm);
    wbp = up_p(wb);
    wbn = up_n(wb);
    if (k >= K) {
      flux = (wbp * wm + wbn * wk) * W.te[X.h(i, j, k - 1)] * dxw * dyw;
    } else if (k == 1) {
      flux = (wbp * wm + wbn * wk) * W.te[X.h(i, j, k)] * dxw * dyw;
    } else {
      eta = W.etat[X.h(i, j, k - 1)];
      wb = (1.0 - eta) * wm + eta * wk;
      wb = (wb - 0.0) * W.alt[X.h(i, j, k - 1)] + 0.0;
      wbp = up_p(wb);
      wbn = up_n(wb);
      flux = (wbp * wm + wbn * wk) * W.te[X.h(i, j, k - 1)] * dxw * dyw;
    }
    A.cftb[X.tb(i, j, k, 2)] = flux;
  }
}

// One thread per (i, j, k) in [1, I+1] x [1, J+1] x [1, K+1]: the x face
// fluxes at i (j <= J, k <= K), the y faces at j (i <= I, k <= K), the z
// faces at k (i <= I, j <= J).
__global__ void __launch_bounds__(256) k_pred_faces(PredArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > A.I + 1 || j > A.J + 1 || k > A.K + 1) return;
  const Ix X = ix_of(A);
  if (j <= A.J && k <= A.K) flux_x(A, X, i, j, k);
  if (i <= A.I && k <= A.K) flux_y(A, X, i, j, k);
  if (i <= A.I && j <= A.J) flux_z(A, X, i, j, k);
}

// ---- SpaceOrder 2: ModifiedConvectiveFlux (:2253-2644) ----
//
// Central interpolation with the cut-cell corrections.  The reference loops
// i outermost and k innermost (:2277-2279), and a cut face whose two
// neighbour tests both fail divides by the local delh of the last cut face
// of the same call that set one (the variable keeps its value between
// iterations).  The GPU resolves that carry in passes per direction: one
// thread walks each piece of kPieceK planes of a k-column (i, j) -- the
// innermost loop -- and records the piece's last delh; one thread per column
// chains its pieces; one workgroup scans the columns in loop order.  The
// piece walk also records the carry into every iteration within its piece,
// so the flux pass runs one thread per face set, as the first-order kernel
// does, each starting from its iteration's carry.  Before any cut face of the call has set delh the reference
// reads an undefined value; the face's own delhec is taken there (the
// first-order routine's choice, :1415).  x**2.d0 is x*x, the exact square.

struct Carry {
  int have;
  double v;
};

// The neighbour tests of a cut face (cells a then b), the delh carry and the
// scaled convective velocity cv * delhec / (delh + epsi).  `avail` false
// where the cells lie beyond the arrays (no neighbour, as the first-order
// routine's guards read them); `init`: delh = delhec before the tests
// (:2596).
__device__ inline double cut2(const PredArgs& A, const Ix& X, const PredCell& C, bool avail,
                              int ai, int aj, int ak, int bi, int bj, int bk, int q, double cv,
                              double dhec, Carry& cr, bool init) {
#pragma clang fp contract(off)
  double delh = 0.0;
  bool set = init;
  if (init) delh = dhec;
  if (avail && usable(X, C, ai, aj, ak) && C.vof[X.c(ai, aj, ak)] >= C.vof[X.c(bi, bj, bk)]) {
    delh = plane_dist(X, C, ai, aj, ak);
    cv = v12(A, q, X.h(ai, aj, ak));
    set = true;
  }
  if (avail && usable(X, C, bi, bj, bk) && C.vof[X.c(bi, bj, bk)] > C.vof[X.c(ai, aj, ak)]) {
    delh = plane_dist(X, C, bi, bj, bk);
    cv = v12(A, q, X.h(bi, bj, bk));
    set = true;
  }
  if (set) {
    cr.have = 1;
    cr.v = delh;
  } else {
    delh = cr.have ? cr.v : dhec;
  }
  return cv * dhec / (delh + kEps);
}

__device__ inline double clamp_eta(double eta) { return fabs(eta) >= 1.0 ? 0.5 : eta; }

// x faces, iteration (i, j, k) of the iu = 1 call (:2280-2408)
template <bool WRITE>
__device__ void flux2_x(const PredArgs& A, const Ix& X, int i, int j, int k, Carry& cr) {
#pragma clang fp contract(off)
  const PredCell &U = A.old[0], &V = A.old[1], &W = A.old[2];
  const int I = A.I, J = A.J, K = A.K;
  double uw, eta, flux;
  if (WRITE) {  // u momentum, convective u
    const double dyu = A.dyp[j - 1], dzu = A.dzp[k - 1];
    const double ui = v12(A, 0, X.h(i, j, k)), um = v12(A, 0, X.h(i - 1, j, k));
    uw = 0.5 * (ui + um);
    if (i == 1) {
      flux = uw * uw * U.ee[X.h(i, j, k)] * dyu * dzu;
    } else if (i >= I) {
      flux = uw * uw * U.ee[X.h(i - 1, j, k)] * dyu * dzu;
    } else {
      eta = U.etae[X.h(i - 1, j, k)];
      uw = (1.0 - eta) * um + eta * ui;
      const double a = uw * U.ale[X.h(i - 1, j, k)];
      flux = a * a * U.ee[X.h(i - 1, j, k)] * dyu * dzu;
    }
    A.cfew[X.ew(i, j, k, 0)] = flux;
  }
  {  // v momentum, convective u
    const double dyv = A.dyv[j - 1], dzv = A.dzp[k - 1];
    uw = 0.5 * (v12(A, 0, X.h(i - 1, j + 1, k)) + v12(A, 0, X.h(i - 1, j, k)));
    if (i > I) {
      flux = uw * bcval(A, 7, j, k) * V.ee[X.h(i - 1, j, k)] * dyv * dzv;
    } else if (i == 1) {
      flux = uw * bcval(A, 6, j, k) * V.ee[X.h(i, j, k)] * dyv * dzv;
    } else {
      if (A.pee[X.h(i - 1, j, k)] <= 1.0e-5 || A.pee[X.h(i - 1, j + 1, k)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, V, V.fce, i - 1, j, k, i - 1, j, k, i, j, k, 0,
                                     A.dxu[i - 2]);
        uw = cut2(A, X, U, j < J, i - 1, j + 1, k, i - 1, j, k, 0, uw, dhec, cr, false);
      } else {
        const double Sy = U.syn[X.h(i - 1, j, k)];
        eta = clamp_eta(fabs(V.fce[X.h(i - 1, j, k) + X.Nh] + dyv / 2.0 -
                             U.cent[X.c(i - 1, j, k) + X.N]) / Sy);
        uw = (1.0 - eta) * v12(A, 0, X.h(i - 1, j, k)) + eta * v12(A, 0, X.h(i - 1, j + 1, k));
      }
      eta = V.etae[X.h(i - 1, j, k)];
      const double vw = (1.0 - eta) * v12(A, 1, X.h(i - 1, j, k)) + eta * v12(A, 1, X.h(i, j, k));
      flux = uw * vw * V.ale[X.h(i - 1, j, k)] * V.ee[X.h(i - 1, j, k)] * dyv * dzv;
    }
    if (WRITE) A.cfew[X.ew(i, j, k, 1)] = flux;
  }
  {  // w momentum, convective u
    const double dyw = A.dyp[j - 1], dzw = A.dzw[k - 1];
    uw = 0.5 * (v12(A, 0, X.h(i - 1, j, k + 1)) + v12(A, 0, X.h(i - 1, j, k)));
    if (i > I) {
      flux = uw * bcval(A, 13, j, k) * W.ee[X.h(i - 1, j, k)] * dyw * dzw;
    } else if (i == 1) {
      flux = uw * bcval(A, 12, j, k) * W.ee[X.h(i, j, k)] * dyw * dzw;
    } else {
      if (A.pee[X.h(i - 1, j, k)] <= 1.0e-5 || A.pee[X.h(i - 1, j, k + 1)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, W, W.fce, i - 1, j, k, i - 1, j, k, i, j, k, 0,
                                     A.dxu[i - 2]);
        uw = cut2(A, X, U, k < K, i - 1, j, k + 1, i - 1, j, k, 0, uw, dhec, cr, false);
      } else {
        const double Sz = U.szt[X.h(i - 1, j, k)];
        eta = clamp_eta(fabs(W.fce[X.h(i - 1, j, k) + 2 * X.Nh] + dzw / 2.0 -
                             U.cent[X.c(i - 1, j, k) + 2 * X.N]) / Sz);
        uw = (1.0 - eta) * v12(A, 0, X.h(i - 1, j, k)) + eta * v12(A, 0, X.h(i - 1, j, k + 1));
      }
      eta = W.etae[X.h(i - 1, j, k)];
      const double ww = (1.0 - eta) * v12(A, 2, X.h(i - 1, j, k)) + eta * v12(A, 2, X.h(i, j, k));
      flux = uw * ww * W.ale[X.h(i - 1, j, k)] * W.ee[X.h(i - 1, j, k)] * dyw * dzw;
    }
    if (WRITE) A.cfew[X.ew(i, j, k, 2)] = flux;
  }
}

// y faces, iteration (i, j, k) of the iv = 1 call (:2409-2520)
template <bool WRITE>
__device__ void flux2_y(const PredArgs& A, const Ix& X, int i, int j, int k, Carry& cr) {
#pragma clang fp contract(off)
  const PredCell &U = A.old[0], &V = A.old[1], &W = A.old[2];
  const int I = A.I, J = A.J, K = A.K;
  double vs, eta, flux;
  {  // u momentum, convective v
    const double dxu = A.dxu[i - 1], dzu = A.dzp[k - 1];
    vs = 0.5 * (v12(A, 1, X.h(i, j - 1, k)) + v12(A, 1, X.h(i + 1, j - 1, k)));
    if (j > J) {
      flux = vs * bcval(A, 3, i, k) * U.ne[X.h(i, j - 1, k)] * dxu * dzu;
    } else if (j == 1) {
      flux = vs * bcval(A, 2, i, k) * U.ne[X.h(i, j, k)] * dxu * dzu;
    } else {
      if (A.pne[X.h(i, j - 1, k)] <= 1.0e-5 || A.pne[X.h(i + 1, j - 1, k)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, U, U.fcn, i, j - 1, k, i, j - 1, k, i, j, k, 1,
                                     A.dyv[j - 2]);
        vs = cut2(A, X, V, i < I, i, j - 1, k, i + 1, j - 1, k, 1, vs, dhec, cr, false);
      } else {
        const double Sx = V.sxe[X.h(i, j - 1, k)];
        eta = clamp_eta(fabs(U.fcn[X.h(i, j - 1, k)] + 0.5 * dxu - V.cent[X.c(i, j - 1, k)]) / Sx);
        vs = (1.0 - eta) * v12(A, 1, X.h(i, j - 1, k)) + eta * v12(A, 1, X.h(i + 1, j - 1, k));
      }
      eta = U.etan[X.h(i, j - 1, k)];
      const double us = (1.0 - eta) * v12(A, 0, X.h(i, j - 1, k)) + eta * v12(A, 0, X.h(i, j, k));
      flux = us * vs * U.aln[X.h(i, j - 1, k)] * U.ne[X.h(i, j - 1, k)] * dxu * dzu;
    }
    if (WRITE) A.cfns[X.ns(i, j, k, 0)] = flux;
  }
  if (WRITE) {  // v momentum, convective v
    const double dxv = A.dxp[i - 1], dzv = A.dzp[k - 1];
    const double vj = v12(A, 1, X.h(i, j, k)), vm = v12(A, 1, X.h(i, j - 1, k));
    vs = 0.5 * (vj + vm);
    if (j >= J) {
      flux = vs * vs * V.ne[X.h(i, j - 1, k)] * dxv * dzv;
    } else if (j == 1) {
      flux = vs * vs * V.ne[X.h(i, j, k)] * dxv * dzv;
    } else {
      eta = V.etan[X.h(i, j - 1, k)];
      vs = (1.0 - eta) * vm + eta * vj;
      const double a = vs * V.aln[X.h(i, j - 1, k)];
      flux = a * a * V.ne[X.h(i, j - 1, k)] * dxv * dzv;
    }
    A.cfns[X.ns(i, j, k, 1)] = flux;
  }
  {  // w momentum, convective v
    const double dxw = A.dxp[i - 1], dzw = A.dzw[k - 1];
    vs = 0.5 * (v12(A, 1, X.h(i, j - 1, k)) + v12(A, 1, X.h(i, j - 1, k + 1)));
    if (j > J) {
      flux = vs * bcval(A, 15, i, k) * W.ne[X.h(i, j - 1, k)] * dxw * dzw;
    } else if (j == 1) {
      flux = vs * bcval(A, 14, i, k) * W.ne[X.h(i, j, k)] * dxw * dzw;
    } else {
      if (A.pne[X.h(i, j - 1, k)] <= 1.0e-5 || A.pne[X.h(i, j - 1, k + 1)] <= 1.0e-5) {
        const double dhec = cut_dhec(X, W, W.fcn, i, j - 1, k, i, j - 1, k, i, j, k, 1,
                                     A.dyp[j - 2]);
        vs = cut2(A, X, V, k < K, i, j - 1, k, i, j - 1, k + 1, 1, vs, dhec, cr, false);
      } else {
        const double Sz = V.szt[X.h(i, j - 1, k)];
        eta = clamp_eta(fabs(W.fcn[X.h(i, j - 1, k) + 2 * X.Nh] + 0.5 * dzw -
                             W.cent[X.c(i, j - 1, k) + 2 * X.N]) / Sz);
        vs = (1.0 - eta) * v12(A, 1, X.h(i, j - 1, k)) + eta * v12(A, 1, X.h(i, j - 1, k + 1));
      }
      eta = W.etan[X.h(i, j - 1, k)];
      const double ws = (1.0 - eta) * v12(A, 2, X.h(i, j - 1, k)) + eta * v12(A, 2, X.h(i, j, k));
      flux = vs * ws * W.aln[X.h(i, j - 1, k)] * W.ne[X.h(i, j - 1, k)] * dxw * dzw;
    }
    if (WRITE) A.cfns[X.ns(i, j, k, 2)] = flux;
  }
}

// z faces, iteration (i, j, k) of the iw = 1 call (:2521-2639)
template <bool WRITE>
__device__ void flux2_z(const PredArgs& A, const Ix& X, int i, int j, int k, Carry& cr) {
#pragma clang fp contract(off)
  const PredCell &U = A.old[0], &V = A.old[1], &W = A.old[2];
  const int I = A.I, J = A.J, K = A.K;
  double wb, eta, flux;
  {  // u momentum, convective w: the average takes vn12(i+1,j,k-1) (:2523)
    const double dxu = A.dxu[i - 1], dyu = A.dyp[j - 1];
    wb = 0.5 * (v12(A, 2, X.h(i, j, k - 1)) + v12(A, 1, X.h(i + 1, j, k - 1)));
    if (k > K) {
      flux = wb * bcval(A, 5, i, j) * U.te[X.h(i, j, k - 1)] * dxu * dyu;
    } else if (k == 1) {
      flux = wb * bcval(A, 4, i, j) * U.te[X.h(i, j, k)] * dxu * dyu;
    } else {
      if (A.pte[X.h(i, j, k - 1)] <= 1.0e-5 || A.pte[X.h(i + 1, j, k - 1)] <= 1.0e-5) {
        // (FCT(:,2) - dz) * ny, as :2537-2538 write it
        const double dhec = cut_dhec(X, U, U.fct, i, j, k - 1, i, j, k - 1, i, j, k, 1,
                                     A.dzw[k - 2]);
        wb = cut2(A, X, W, i < I, i, j, k - 1, i + 1, j, k - 1, 2, wb, dhec, cr, false);
      } else {
        const double Sx = W.sxe[X.h(i, j, k - 1)];
        eta = clamp_eta(fabs(U.fct[X.h(i, j, k - 1)] + 0.5 * dxu - W.cent[X.c(i, j, k - 1)]) / Sx);
        wb = (1.0 - eta) * v12(A, 2, X.h(i, j, k - 1)) + eta * v12(A, 2, X.h(i + 1, j, k - 1));
      }
      eta = U.etat[X.h(i, j, k - 1)];
      const double ub = (1.0 - eta) * v12(A, 0, X.h(i, j, k - 1)) + eta * v12(A, 0, X.h(i, j, k));
      flux = wb * ub * U.alt[X.h(i, j, k - 1)] * U.te[X.h(i, j, k - 1)] * dxu * dyu;
    }
    if (WRITE) A.cftb[X.tb(i, j, k, 0)] = flux;
  }
  {  // v momentum, convective w: delh = delhec before the tests (:2596)
    const double dxv = A.dxp[i - 1], dyv = A.dyv[j - 1];
    wb = 0.5 * (v12(A, 2, X.h(i, j + 1, k - 1)) + v12(A, 2, X.h(i, j, k - 1)));
    if (k > K) {
      flux = wb * bcval(A, 11, i, j) * V.te[X.h(i, j, k - 1)] * dxv * dyv;
    } else if (k == 1) {
      flux = wb * bcval(A, 10, i, j) * V.te[X.h(i, j, k)] * dxv * dyv;
    } else {
      if (A.pte[X.h(i, j, k - 1)] <= 1.0e-5 || A.pte[X.h(i, j + 1, k - 1)] <= 1.0e-5) {
        // (FCT(:,1) - dz) * nx, as :2592-2593 write it
        const double dhec = cut_dhec(X, V, V.fct, i, j, k - 1, i, j, k - 1, i, j, k, 0,
                                     A.dzw[k - 2]);
        wb = cut2(A, X, W, j < J, i, j + 1, k - 1, i, j, k - 1, 2, wb, dhec, cr, true);
      } else {
        const double Sy = W.syn[X.h(i, j, k - 1)];
        eta = clamp_eta(fabs(V.fct[X.h(i, j, k - 1) + X.Nh] + 0.5 * dyv -
                             W.cent[X.c(i, j, k - 1) + X.N]) / Sy);
        wb = (1.0 - eta) * v12(A, 2, X.h(i, j, k - 1)) + eta * v12(A, 2, X.h(i, j + 1, k - 1));
      }
      eta = V.etat[X.h(i, j, k - 1)];
      const double vb = (1.0 - eta) * v12(A, 1, X.h(i, j, k - 1)) + eta * v12(A, 1, X.h(i, j, k));
      flux = wb * vb * V.alt[X.h(i, j, k - 1)] * V.te[X.h(i, j, k - 1)] * dxv * dyv;
    }
    if (WRITE) A.cftb[X.tb(i, j, k, 1)] = flux;
  }
  if (WRITE) {  // w momentum, convective w
    const double dxw = A.dxp[i - 1], dyw = A.dyp[j - 1];
    const double wk = v12(A, 2, X.h(i, j, k)), wm = v12(A, 2, X.h(i, j, k - 1));
    wb = 0.5 * (wk + wm);
    if (k >= K) {
      flux = wb * wb * W.te[X.h(i, j, k - 1)] * dxw * dyw;
    } else if (k == 1) {
      flux = wb * wb * W.te[X.h(i, j, k)] * dxw * dyw;
    } else {
      eta = W.etat[X.h(i, j, k - 1)];
      wb = (1.0 - eta) * wm + eta * wk;
      const double a = wb * W.alt[X.h(i, j, k - 1)];
      flux = a * a * W.te[X.h(i, j, k - 1)] * dxw * dyw;
    }
    A.cftb[X.tb(i, j, k, 2)] = flux;
  }
}

// k-planes of one column piece: a thread walks one piece
constexpr int kPieceK = 16;

__host__ __device__ inline int pieces_of(int nk) { return (nk + kPieceK - 1) / kPieceK; }

// The carries of one direction: per iteration (i, j, k) the carry within
// its piece (index (i-1) + NI*((j-1) + NJ*(k-1))), per piece (stored
// piece-major, piece p of column c at p * ncol + c, columns numbered in loop
// order (i-1)*NJ + j-1) and per column.
struct CarryBufs {
  double* iv;       // iteration: last delh set by an earlier iteration of its piece
  uint8_t* ih;      //   and whether there was one
  double *pv, *ph;  // piece: last delh set in the piece, then the carry into it
  double *cv, *ch;  // column: last delh set in the column, then the carry into it
};

template <int D, bool WRITE>
__device__ inline void flux2(const PredArgs& A, const Ix& X, int i, int j, int k, Carry& cr) {
  if (D == 0) flux2_x<WRITE>(A, X, i, j, k, cr);
  else if (D == 1) flux2_y<WRITE>(A, X, i, j, k, cr);
  else flux2_z<WRITE>(A, X, i, j, k, cr);
}

// One thread per piece (i, j, p): planes k = p*kPieceK+1 .. of the k-column
// (i, j) of direction D's loop.  Evaluates the cut-face tests of each
// iteration in loop order and records the carry into every iteration and
// the piece's last delh.
template <int D>
__global__ void __launch_bounds__(256) k_pred_carry_pieces(PredArgs A, CarryBufs B) {
  const int NI = A.I + (D == 0), NJ = A.J + (D == 1), NK = A.K + (D == 2);
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int p = blockIdx.z;
  if (i > NI || j > NJ) return;
  const int64_t ncol = (int64_t)NI * NJ, col = (int64_t)(i - 1) * NJ + (j - 1);
  const Ix X = ix_of(A);
  Carry cr{0, 0.0};
  const int k0 = p * kPieceK + 1, k1 = min(NK, k0 + kPieceK - 1);
  int64_t n = (i - 1) + (int64_t)NI * ((j - 1) + (int64_t)NJ * (k0 - 1));
  for (int k = k0; k <= k1; ++k, n += ncol) {
    B.ih[n] = (uint8_t)cr.have;
    B.iv[n] = cr.v;
    flux2<D, false>(A, X, i, j, k, cr);
  }
  const int64_t x = p * ncol + col;
  B.ph[x] = cr.have ? 1.0 : 0.0;
  B.pv[x] = cr.v;
}

// The fluxes of iteration (i, j, k) of direction D's loop from the carry
// into it: within its piece, else into the piece, else into the column.
template <int D>
__device__ void face2(const PredArgs& A, const Ix& X, const CarryBufs& B, int i, int j, int k) {
  const int NI = A.I + (D == 0), NJ = A.J + (D == 1);
  const int64_t ncol = (int64_t)NI * NJ, col = (int64_t)(i - 1) * NJ + (j - 1);
  const int64_t n = (i - 1) + (int64_t)NI * ((j - 1) + (int64_t)NJ * (k - 1));
  Carry cr{0, 0.0};
  if (B.ih[n]) {
    cr.have = 1;
    cr.v = B.iv[n];
  } else {
    const int64_t x = (int64_t)((k - 1) / kPieceK) * ncol + col;
    if (B.ph[x] != 0.0) {
      cr.have = 1;
      cr.v = B.pv[x];
    } else if (B.ch[col] != 0.0) {
      cr.have = 1;
      cr.v = B.cv[col];
    }
  }
  flux2<D, true>(A, X, i, j, k, cr);
}

// One thread per (i, j, k) writes the x, y and z face fluxes it owns, as the
// first-order kernel does.
__global__ void __launch_bounds__(256) k_pred_faces2(PredArgs A, CarryBufs Bx, CarryBufs By,
                                                     CarryBufs Bz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > A.I + 1 || j > A.J + 1 || k > A.K + 1) return;
  const Ix X = ix_of(A);
  if (j <= A.J && k <= A.K) face2<0>(A, X, Bx, i, j, k);
  if (i <= A.I && k <= A.K) face2<1>(A, X, By, i, j, k);
  if (i <= A.I && j <= A.J) face2<2>(A, X, Bz, i, j, k);
}

// One thread per column: the pieces' exclusive carries within the column
// (in place) and the column's last delh.
__global__ void __launch_bounds__(256) k_pred_piece_carry(CarryBufs B, int64_t ncol, int np) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  int h = 0;
  double v = 0.0;
  for (int p = 0; p < np; ++p) {
    const int64_t x = p * ncol + col;
    const bool xh = B.ph[x] != 0.0;
    const double xv = B.pv[x];
    B.ph[x] = h ? 1.0 : 0.0;
    B.pv[x] = v;
    if (xh) { h = 1; v = xv; }
  }
  B.ch[col] = h ? 1.0 : 0.0;
  B.cv[col] = v;
}

// Exclusive scan of the column carries in loop order, in place: afterwards
// column c holds the last delh set by a column before it (or none).  One
// workgroup per direction walks the columns in tiles of 4096: coalesced
// loads into LDS, four consecutive columns per thread, a scan over the
// threads, the carry of the earlier tiles.
constexpr int kScanTile = 4096;

__global__ void __launch_bounds__(1024) k_pred_carry_scan(CarryBufs B0, CarryBufs B1, CarryBufs B2,
                                                          int64_t n0, int64_t n1, int64_t n2) {
  __shared__ double sv[kScanTile];
  __shared__ int sh[kScanTile];
  __shared__ double pv[1024];
  __shared__ int ph[1024];
  const int d = blockIdx.x, t = threadIdx.x;
  const CarryBufs& B = d == 0 ? B0 : (d == 1 ? B1 : B2);
  const int64_t n = d == 0 ? n0 : (d == 1 ? n1 : n2);
  double* v = B.cv;
  double* h = B.ch;
  int carry_h = 0;
  double carry_v = 0.0;
  for (int64_t base = 0; base < n; base += kScanTile) {
    for (int e = 0; e < 4; ++e) {
      const int l = e * 1024 + t;
      const int64_t x = base + l;
      sh[l] = x < n ? (h[x] != 0.0) : 0;
      sv[l] = x < n ? v[x] : 0.0;
    }
    __syncthreads();
    int lh = 0;
    double lv = 0.0;
    for (int e = 0; e < 4; ++e)
      if (sh[4 * t + e]) { lh = 1; lv = sv[4 * t + e]; }
    ph[t] = lh;
    pv[t] = lv;
    __syncthreads();
    // inclusive scan over the threads: no delh in a run takes the last before it
    for (int o = 1; o < 1024; o <<= 1) {
      const int qh = t >= o ? ph[t - o] : 0;
      const double qv = t >= o ? pv[t - o] : 0.0;
      __syncthreads();
      if (!ph[t] && qh) { ph[t] = 1; pv[t] = qv; }
      __syncthreads();
    }
    int eh = t > 0 ? ph[t - 1] : 0;
    double ev = t > 0 ? pv[t - 1] : 0.0;
    if (!eh) { eh = carry_h; ev = carry_v; }
    for (int e = 0; e < 4; ++e) {
      const int l = 4 * t + e;
      const int xh = sh[l];
      const double xv = sv[l];
      sh[l] = eh;
      sv[l] = ev;
      if (xh) { eh = 1; ev = xv; }
    }
    __syncthreads();
    if (ph[1023]) { carry_h = 1; carry_v = pv[1023]; }
    for (int e = 0; e < 4; ++e) {
      const int l = e * 1024 + t;
      const int64_t x = base + l;
      if (x < n) {
        h[x] = sh[l] ? 1.0 : 0.0;
        v[x] = sv[l];
      }
    }
    __syncthreads();
  }
}

// One thread per cell (i, j, k): the U, V, W cells it owns.
__global__ void __launch_bounds__(256) k_pred_cells(PredArgs A) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > A.I || j > A.J || k > A.K) return;
  const Ix X = ix_of(A);
  const int64_t n = X.c(i, j, k), h = X.h(i, j, k);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if ((c == 0 && i == A.I) || (c == 1 && j == A.J) || (c == 2 && k == A.K)) continue;
    const double Fe = A.cfew[X.ew(i + 1, j, k, c)], Fw = A.cfew[X.ew(i, j, k, c)];
    const double Fn = A.cfns[X.ns(i, j + 1, k, c)], Fs = A.cfns[X.ns(i, j, k, c)];
    const double Ft = A.cftb[X.tb(i, j, k + 1, c)], Fb = A.cftb[X.tb(i, j, k, c)];
    const double flux0 = Fe - Fw + Fn - Fs + Ft - Fb;
    double* fdo = A.fdo + n + X.N * c;
    const double fdiv = A.ab2 ? 1.5 * flux0 - 0.5 * *fdo : flux0;
    *fdo = flux0;
    const PredCell& C = A.cur[c];
    double* dp = A.dp[c];
    double* pr = A.pred[c];
    if (C.ct[n] == 2) {
      dp[h] = 0.0;
      pr[h] = 0.0;
      continue;
    }
    const double dx = c == 0 ? A.dxu[i - 1] : A.dxp[i - 1];
    const double dy = c == 1 ? A.dyv[j - 1] : A.dyp[j - 1];
    const double dz = c == 2 ? A.dzw[k - 1] : A.dzp[k - 1];
    const double vof = C.vof[n], vofl = C.vof_l[n];
    // wall friction (:404-410)
    const double fr =
        (vofl / vof * A.mu_w + (1.0 - vofl / vof) * A.mu_a) / A.rey * C.wllh[n] / C.delh[n];
    // aP with zero neighbour coefficients (:1025-1037)
    const double aP = vof * dx * dy * dz / A.dt + fr;
    double dpv = 1.0 / aP;
    // right-hand side (:1206-1241)
    const double s = (double)(c == 0) * A.vel[0][h] + (double)(c == 1) * A.vel[1][h] +
                     (double)(c == 2) * A.vel[2][h];
    const double gc = A.gx * (double)(c == 0) + A.gy * (double)(c == 1) + A.gz * (double)(c == 2);
    double rhs = s * vof * dx * dy * dz / A.dt - s * (vof - A.old[c].vof[n]) * dx * dy * dz / A.dt +
                 A.inv_fr2 * vof * dx * dy * dz * gc / A.g;
    dpv = vof * dx * dy * dz * dpv;
    rhs = rhs - fdiv;
    dp[h] = dpv;
    pr[h] = rhs / aP;
  }
}

// Dp halo: every entry outside the component's computed range takes the
// value at the clamped index (the component's last plane from the one
// before it).
__global__ void __launch_bounds__(256) k_pred_dp_halo(PredArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > A.I + 1 || j > A.J + 1 || k > A.K + 1) return;
  const Ix X = ix_of(A);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int ti = c == 0 ? A.I - 1 : A.I, tj = c == 1 ? A.J - 1 : A.J, tk = c == 2 ? A.K - 1 : A.K;
    if (i >= 1 && i <= ti && j >= 1 && j <= tj && k >= 1 && k <= tk) continue;
    const int si = i < 1 ? 1 : (i > ti ? ti : i);
    const int sj = j < 1 ? 1 : (j > tj ? tj : j);
    const int sk = k < 1 ? 1 : (k > tk ? tk : k);
    A.dp[c][X.h(i, j, k)] = A.dp[c][X.h(si, sj, sk)];
  }
}

// Predictor boundary conditions and the solid-cell condition of the
// component's last plane (:3796-3916).  Thread (a, b) of face set q:
// q = 0: u at i = 0, I, I+1 for (j, k) = (a, b); q = 1: v at j = 0, J, J+1
// for (i, k); q = 2: w at k = 0, K, K+1 for (i, j).
__global__ void __launch_bounds__(256) k_pred_bc(PredArgs A) {
#pragma clang fp contract(off)
  const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int b = blockIdx.y + 1;
  const int q = blockIdx.z;
  const Ix X = ix_of(A);
  const int na = q == 0 ? A.J : A.I, nb = q == 2 ? A.J : A.K;
  if (a > na || b > nb) return;
  const int L = q == 0 ? A.I : (q == 1 ? A.J : A.K);
  auto at = [&](int m) -> int64_t {
    return q == 0 ? X.h(m, a, b) : (q == 1 ? X.h(a, m, b) : X.h(a, b, m));
  };
  double* p = A.pred[q];
  const int fw = 6 * q + 2 * q, fe = fw + 1;  // W/E of u, S/N of v, B/T of w
  const double ve = bcval(A, fe, a, b);
  p[at(0)] = A.flag[fw] == 1 ? p[at(1)] : bcval(A, fw, a, b);
  const double before = p[at(L - 1)];
  p[at(L)] = A.flag[fe] == 1 ? before : ve;
  p[at(L + 1)] = 2.0 * ve - before;
  const int64_t n = q == 0 ? X.c(L, a, b) : (q == 1 ? X.c(a, L, b) : X.c(a, b, L));
  if (A.cur[q].ct[n] == 2) p[at(L)] = 0.0;
}

}  // namespace

static PredCell pred_cell(const sp_cell_geom& g, const sp_cell_interp& x) {
  PredCell c{};
  c.phi = g.phi; c.nx = g.nx; c.ny = g.ny; c.nz = g.nz; c.vof = g.vof; c.vof_l = g.vof_l;
  c.ct = g.cell_type; c.moex = x.moex;
  c.ee = g.ee_area; c.ne = g.ne_area; c.te = g.te_area;
  c.fce = g.fce; c.fcn = g.fcn; c.fct = g.fct; c.cent = g.cell_cent;
  c.sxe = x.sxe; c.syn = x.syn; c.szt = x.szt;
  c.etae = x.etae; c.etan = x.etan; c.etat = x.etat;
  c.ale = x.ale; c.aln = x.aln; c.alt = x.alt;
  c.wllh = g.wllh; c.delh = g.delh;
  return c;
}

sp_status predict_uvw(sp_ctx* c, const sp_pred_inputs& in, const sp_pred_outputs& out,
                      double* faces) {
  const Dims& D = c->dm;
  PredArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K; A.SX = D.SX; A.SY = D.SY; A.N = D.N; A.Nh = D.Nh;
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp; A.dxu = c->dxu; A.dyv = c->dyv; A.dzw = c->dzw;
  A.pee = in.p_ee_area; A.pne = in.p_ne_area; A.pte = in.p_te_area;
  const sp_cell_geom* og[3] = {in.u_old, in.v_old, in.w_old};
  const sp_cell_interp* ox[3] = {in.u_old_x, in.v_old_x, in.w_old_x};
  const sp_cell_geom* cg[3] = {in.u_cur, in.v_cur, in.w_cur};
  for (int q = 0; q < 3; ++q) {
    A.old[q] = pred_cell(*og[q], *ox[q]);
    sp_cell_interp none{};
    A.cur[q] = pred_cell(*cg[q], none);
  }
  A.vel[0] = in.u; A.vel[1] = in.v; A.vel[2] = in.w;
  A.veln[0] = in.u_n; A.veln[1] = in.v_n; A.veln[2] = in.w_n;
  A.extrapolate = in.time_order == 2;
  A.ab2 = in.itt != 1 && in.time_order != 1 && in.time_order != 2;
  for (int q = 0; q < 18; ++q) { A.bcv[q] = in.bc_var[q]; A.flag[q] = in.bc_flag[q]; }
  A.fdo = in.flux_div_old;
  const int64_t few = (int64_t)(D.I + 2) * D.J * D.K * 3, fns = (int64_t)D.I * (D.J + 2) * D.K * 3;
  A.cfew = faces; A.cfns = faces + few; A.cftb = faces + few + fns;
  // SpaceOrder 2 from the second step on (the first step takes the
  // first-order fluxes, :172-185)
  const bool second = in.space_order == 2 && in.itt != 1;
  A.pred[0] = out.u; A.pred[1] = out.v; A.pred[2] = out.w;
  A.dp[0] = out.dp_u; A.dp[1] = out.dp_v; A.dp[2] = out.dp_w;
  A.dt = in.dt;
  A.inv_fr2 = 1.0 / (in.fr * in.fr);  // 1.d0/Fr**2.d0
  A.gx = in.gx; A.gy = in.gy; A.gz = in.gz; A.g = in.g;
  A.mu_w = in.mu_w_ref; A.mu_a = in.mu_a_ref; A.rey = in.rey;
  const dim3 blk(64, 4, 1);
  int slot;
  const double NB = (double)D.N;
  // per cell: old velocities (x3, x2 for TimeOrder 2), P apertures, the
  // old cells' apertures, interpolation and centre data read by the
  // non-cut branches -> 9 face fluxes
  if (!second) {
    timer_begin(c, "pred_faces", NB * (8.0 * (3 + 3 + 9 + 4 + 6 + 6 + 6 + (A.extrapolate ? 3 : 0)) + 72.0), &slot);
    hipLaunchKernelGGL(k_pred_faces, dim3((D.I + 1 + 63) / 64, (D.J + 1 + 3) / 4, D.K + 1), blk,
                       0, c->stream, A);
    timer_end(c, slot);
  } else {
    // the carries follow the face fluxes in the same buffer
    const int np[3] = {pieces_of(D.K), pieces_of(D.K), pieces_of(D.K + 1)};
    const int64_t nc[3] = {(int64_t)(D.I + 1) * D.J, (int64_t)D.I * (D.J + 1), (int64_t)D.I * D.J};
    const int64_t nk[3] = {D.K, D.K, D.K + 1};
    CarryBufs cb[3];
    double* q = faces + predict_flux_doubles(D);
    for (int d = 0; d < 3; ++d) {
      cb[d].iv = q; q += nc[d] * nk[d];
      cb[d].pv = q; q += nc[d] * np[d];
      cb[d].ph = q; q += nc[d] * np[d];
      cb[d].cv = q; q += nc[d];
      cb[d].ch = q; q += nc[d];
    }
    uint8_t* qb = reinterpret_cast<uint8_t*>(q);
    for (int d = 0; d < 3; ++d) {
      cb[d].ih = qb;
      qb += nc[d] * nk[d];
    }
    const dim3 g0((D.I + 1 + 63) / 64, (D.J + 3) / 4, np[0]),
        g1((D.I + 63) / 64, (D.J + 1 + 3) / 4, np[1]), g2((D.I + 63) / 64, (D.J + 3) / 4, np[2]);
    // latency-bound passes over the cut-face tests: no byte count is claimed
    timer_begin(c, "pred_faces2_carry", 0.0, &slot);
    hipLaunchKernelGGL(k_pred_carry_pieces<0>, g0, blk, 0, c->stream, A, cb[0]);
    hipLaunchKernelGGL(k_pred_carry_pieces<1>, g1, blk, 0, c->stream, A, cb[1]);
    hipLaunchKernelGGL(k_pred_carry_pieces<2>, g2, blk, 0, c->stream, A, cb[2]);
    for (int d = 0; d < 3; ++d)
      hipLaunchKernelGGL(k_pred_piece_carry, dim3((unsigned)((nc[d] + 255) / 256)), dim3(256), 0,
                         c->stream, cb[d], nc[d], np[d]);
    hipLaunchKernelGGL(k_pred_carry_scan, dim3(3), dim3(1024), 0, c->stream, cb[0], cb[1], cb[2],
                       nc[0], nc[1], nc[2]);
    timer_end(c, slot);
    // the first-order kernel's bytes and the carry reads (9 B per face set)
    timer_begin(c, "pred_faces2",
                NB * (8.0 * (3 + 3 + 9 + 4 + 6 + 6 + 6 + (A.extrapolate ? 3 : 0)) + 72.0 + 27.0),
                &slot);
    hipLaunchKernelGGL(k_pred_faces2, dim3((D.I + 1 + 63) / 64, (D.J + 1 + 3) / 4, D.K + 1), blk,
                       0, c->stream, A, cb[0], cb[1], cb[2]);
    timer_end(c, slot);
  }
  // per cell: 9 face fluxes, FluxDivOld, current vof/vofL/type/WlLh/delh,
  // old vof, velocities -> FluxDivOld, Dp, u*
  timer_begin(c, "pred_cells", NB * (72.0 + 48.0 + 3 * (8 + 8 + 4 + 8 + 8) + 24 + 24 + 48), &slot);
  hipLaunchKernelGGL(k_pred_cells, dim3((D.I + 63) / 64, (D.J + 3) / 4, D.K), blk, 0, c->stream, A);
  timer_end(c, slot);
  hipLaunchKernelGGL(k_pred_dp_halo, dim3((D.I + 2 + 63) / 64, (D.J + 2 + 3) / 4, D.K + 2), blk, 0,
                     c->stream, A);
  const int amax = std::max(D.I, D.J), bmax = std::max(D.J, D.K);
  hipLaunchKernelGGL(k_pred_bc, dim3((amax + 255) / 256, bmax, 3), dim3(256), 0, c->stream, A);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "predictor launch");
}

size_t predict_flux_doubles(const Dims& D) {
  return (size_t)(D.I + 2) * D.J * D.K * 3 + (size_t)D.I * (D.J + 2) * D.K * 3 +
         (size_t)D.I * D.J * (D.K + 2) * 3;
}

size_t predict_face_doubles(const Dims& D, bool second) {
  if (!second) return predict_flux_doubles(D);
  // the face fluxes, then the SpaceOrder 2 carries of each direction's
  // loop: per iteration, per 16-plane piece and per k-column
  const size_t p0 = (size_t)pieces_of(D.K), p2 = (size_t)pieces_of(D.K + 1);
  const size_t c0 = (size_t)(D.I + 1) * D.J, c1 = (size_t)D.I * (D.J + 1), c2 = (size_t)D.I * D.J;
  const size_t it = c0 * D.K + c1 * D.K + c2 * (D.K + 1);  // iterations of the three loops
  // values, then the flags as bytes
  return predict_flux_doubles(D) + it + 2 * (c0 * p0 + c1 * p0 + c2 * p2) + 2 * (c0 + c1 + c2) +
         (it + 7) / 8;
}

}  // namespace sp
