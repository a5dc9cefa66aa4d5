/*
 * CPU FP64 restatement of SIMCOFlow's pressure projection (see orc.h).
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the timed CPU
 * baseline.  Built with -ffp-contract=off so the GFM coefficients and the
 * assembled rows follow the reference's operation order exactly.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define HX(g, i, j, k) ((int64_t)(i) + (int64_t)((g)->I + 2) * ((int64_t)(j) + (int64_t)((g)->J + 2) * (int64_t)(k)))
#define NX(g, i, j, k) ((int64_t)((i) - 1) + (int64_t)(g)->I * ((int64_t)((j) - 1) + (int64_t)(g)->J * (int64_t)((k) - 1)))

static inline int imax_(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* GFM face coefficients: ProjectionP.f90:123-265, called for the six  */
/* faces W,S,B,E,N,T at ProjectionP.f90:53-64.                          */
/* ------------------------------------------------------------------ */
void orc_gfm_coef(const orc_grid* g, const double* phiL, const double* vofL, const double* vof,
                  const double* dpu, const double* dpv, const double* dpw, double rho_l,
                  double rho_g, double rho_ref, double epsi, double* pocoef) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K;
  const double tol = 1.0e-30;                       /* ProjectionP.f90:18 */
  /* (ium,jvm,kwm,iup,jvp,kwp) per face, ProjectionP.f90:53-64 */
  static const int sel[6][6] = {{1, 0, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0}, {0, 0, 1, 0, 0, 0},
                                {0, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1}};
  for (int f = 0; f < 6; ++f) {
    const int ium = sel[f][0], jvm = sel[f][1], kwm = sel[f][2];
    const int iup = sel[f][3], jvp = sel[f][4], kwp = sel[f][5];
    const double* dp = (f % 3 == 0) ? dpu : ((f % 3 == 1) ? dpv : dpw);
    double* T = pocoef + (int64_t)f * N;
    for (int64_t q = 0; q < N; ++q) T[q] = 0.0;    /* :138 */
    const double BetaP = 1.0 / (rho_l / rho_ref);   /* :140 */
    const double BetaM = 1.0 / (rho_g / rho_ref);   /* :141 */
    /* BetaW/BetaD are locals that persist across loop iterations; when a
     * neighbour is neither wet nor dry (solid) the previous value is reused
     * (:170-248).  Their value before the first assignment is undefined in
     * the reference; the restatement starts from the P cell's own phase. */
    double BetaW = BetaM, BetaD = BetaP;
    for (int i = 1; i <= I; ++i)
      for (int j = 1; j <= J; ++j)
        for (int k = 1; k <= K; ++k) {
          const int ii = i - ium + iup, jj = j - jvm + jvp, kk = k - kwm + kwp;
          const int im = i - ium, jm = j - jvm, km = k - kwm;
          /* :151-153 (only one term survives: 1*d + 0*.. + 0*..) */
          double DGrid;
          if (ium + iup) DGrid = g->dxu[imax_(im, 1) - 1];
          else if (jvm + jvp) DGrid = g->dyv[imax_(jm, 1) - 1];
          else DGrid = g->dzw[imax_(km, 1) - 1];
          const int64_t P = HX(g, i, j, k), Q = HX(g, ii, jj, kk);
          const double Lamda = fabs(phiL[P]) / (fabs(phiL[P]) + fabs(phiL[Q]) + tol); /* :155 */
          const double one_m = 1.0 - epsi;
          const int wetP = (vofL[P] >= 0.5 && vof[P] > one_m) ||
                           (phiL[P] < 0.0 && vof[P] <= one_m && vof[P] > epsi);
          const int dryP = (vofL[P] < 0.5 && vof[P] > one_m) ||
                           (phiL[P] >= 0.0 && vof[P] <= one_m && vof[P] > epsi);
          const int wetQ = (vofL[Q] >= 0.5 && vof[Q] > one_m) ||
                           (phiL[Q] < 0.0 && vof[Q] <= one_m && vof[Q] > epsi);
          const int dryQ = (vofL[Q] < 0.5 && vof[Q] > one_m) ||
                           (phiL[Q] >= 0.0 && vof[Q] <= one_m && vof[Q] > epsi);
          const double Dp = dp[HX(g, im, jm, km)];
          const int64_t n = NX(g, i, j, k);
          if (wetP) {                                          /* :170 */
            if (dryQ) BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;  /* :176 */
            else if (wetQ) BetaW = BetaM;                             /* :181 */
            T[n] = Dp / DGrid * BetaP * BetaM / BetaW;               /* :194 */
          } else if (dryP) {                                   /* :212 */
            if (wetQ) BetaD = Lamda * BetaP + (1.0 - Lamda) * BetaM;  /* :218 */
            else if (dryQ) BetaD = BetaP;                             /* :223 */
            T[n] = Dp / DGrid * BetaP * BetaM / BetaD;               /* :235 */
          }
        }
    /* boundary faces copy the first interior value, :256-264 */
    if (ium == 1 || iup == 1) {
      const int dst = ium * 1 + iup * I, src = ium * 2 + iup * (I - 1);
      for (int k = 1; k <= K; ++k)
        for (int j = 1; j <= J; ++j) T[NX(g, dst, j, k)] = T[NX(g, src, j, k)];
    }
    if (jvm == 1 || jvp == 1) {
      const int dst = jvm * 1 + jvp * J, src = jvm * 2 + jvp * (J - 1);
      for (int k = 1; k <= K; ++k)
        for (int i = 1; i <= I; ++i) T[NX(g, i, dst, k)] = T[NX(g, i, src, k)];
    }
    if (kwm == 1 || kwp == 1) {
      const int dst = kwm * 1 + kwp * K, src = kwm * 2 + kwp * (K - 1);
      for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) T[NX(g, i, j, dst)] = T[NX(g, i, j, src)];
    }
  }
}

/* Same result as orc_gfm_coef, computed in parallel: every face whose
 * neighbour is classified is independent; a face next to an unclassified cell
 * takes the value assigned by the last earlier cell of the reference's i-j-k
 * loop, found by a look-back within its k column and a prefix over columns. */
static int gfm_face_kind(const orc_grid* g, const double* phiL, const double* vofL,
                         const double* vof, double epsi, double BetaP, double BetaM, int f, int i,
                         int j, int k, double* beta) {
  static const int sel[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int ii = i + sel[f][0], jj = j + sel[f][1], kk = k + sel[f][2];
  const int64_t P = HX(g, i, j, k), Q = HX(g, ii, jj, kk);
  const double one_m = 1.0 - epsi;
  const int wetP = (vofL[P] >= 0.5 && vof[P] > one_m) || (phiL[P] < 0.0 && vof[P] <= one_m && vof[P] > epsi);
  const int dryP = !wetP && ((vofL[P] < 0.5 && vof[P] > one_m) || (phiL[P] >= 0.0 && vof[P] <= one_m && vof[P] > epsi));
  if (!wetP && !dryP) return 0;
  const int wetQ = (vofL[Q] >= 0.5 && vof[Q] > one_m) || (phiL[Q] < 0.0 && vof[Q] <= one_m && vof[Q] > epsi);
  const int dryQ = !wetQ && ((vofL[Q] < 0.5 && vof[Q] > one_m) || (phiL[Q] >= 0.0 && vof[Q] <= one_m && vof[Q] > epsi));
  const double Lamda = fabs(phiL[P]) / (fabs(phiL[P]) + fabs(phiL[Q]) + 1.0e-30);
  if (wetP) {
    if (dryQ) { *beta = Lamda * BetaM + (1.0 - Lamda) * BetaP; return 1; }
    if (wetQ) { *beta = BetaM; return 1; }
    return -1;
  }
  if (wetQ) { *beta = Lamda * BetaP + (1.0 - Lamda) * BetaM; return 2; }
  if (dryQ) { *beta = BetaP; return 2; }
  return -2;
}

void orc_gfm_coef_fast(const orc_grid* g, const double* phiL, const double* vofL, const double* vof,
                       const double* dpu, const double* dpv, const double* dpw, double rho_l,
                       double rho_g, double rho_ref, double epsi, double* pocoef) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K, cols = (int64_t)I * J;
  const double BetaP = 1.0 / (rho_l / rho_ref), BetaM = 1.0 / (rho_g / rho_ref);
  int64_t* last = (int64_t*)malloc(sizeof(int64_t) * 2 * cols);
  int64_t* pre = (int64_t*)malloc(sizeof(int64_t) * 2 * cols);
  for (int f = 0; f < 6; ++f) {
    const int dir = f % 3, minus = f < 3;
    const double* dp = dir == 0 ? dpu : (dir == 1 ? dpv : dpw);
    const double* dg = dir == 0 ? g->dxu : (dir == 1 ? g->dyv : g->dzw);
    double* T = pocoef + (int64_t)f * N;
    int64_t nstale = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : nstale)
    for (int k = 1; k <= K; ++k)
      for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) {
          double beta = 0;
          const int r = gfm_face_kind(g, phiL, vofL, vof, epsi, BetaP, BetaM, f, i, j, k, &beta);
          const int c = dir == 0 ? i : (dir == 1 ? j : k);
          const double DGrid = dg[(minus ? (c - 1 > 1 ? c - 1 : 1) : c) - 1];
          const int64_t hd = minus ? HX(g, i - (dir == 0), j - (dir == 1), k - (dir == 2)) : HX(g, i, j, k);
          const double pref = dp[hd] / DGrid * BetaP * BetaM;
          T[NX(g, i, j, k)] = r > 0 ? pref / beta : (r < 0 ? NAN : 0.0);
          nstale += r < 0;
        }
    if (nstale) {
#pragma omp parallel for schedule(static)
      for (int64_t col = 0; col < cols; ++col) {
        const int i = (int)(col / J) + 1, j = (int)(col % J) + 1;
        int64_t lw = -1, ld = -1;
        for (int k = K; k >= 1 && (lw < 0 || ld < 0); --k) {
          double beta;
          const int r = gfm_face_kind(g, phiL, vofL, vof, epsi, BetaP, BetaM, f, i, j, k, &beta);
          if (r == 1 && lw < 0) lw = col * K + (k - 1);
          if (r == 2 && ld < 0) ld = col * K + (k - 1);
        }
        last[col] = lw;
        last[cols + col] = ld;
      }
      int64_t rw = -1, rd = -1;
      for (int64_t col = 0; col < cols; ++col) {
        pre[col] = rw; pre[cols + col] = rd;
        if (last[col] > rw) rw = last[col];
        if (last[cols + col] > rd) rd = last[cols + col];
      }
#pragma omp parallel for collapse(2) schedule(dynamic, 64)
      for (int k = 1; k <= K; ++k)
        for (int j = 1; j <= J; ++j)
          for (int i = 1; i <= I; ++i) {
            const int64_t n = NX(g, i, j, k);
            if (!(T[n] != T[n])) continue;
            double beta0;
            const int kind = -gfm_face_kind(g, phiL, vofL, vof, epsi, BetaP, BetaM, f, i, j, k, &beta0);
            const int64_t col = (int64_t)(i - 1) * J + (j - 1);
            int64_t pos = -1;
            for (int kk = k - 1; kk >= 1; --kk) {
              double b2;
              if (gfm_face_kind(g, phiL, vofL, vof, epsi, BetaP, BetaM, f, i, j, kk, &b2) == kind) {
                pos = col * K + (kk - 1);
                break;
              }
            }
            if (pos < 0) pos = pre[(kind - 1) * cols + col];
            double beta = kind == 1 ? BetaM : BetaP;
            if (pos >= 0) {
              const int64_t pc = pos / K;
              gfm_face_kind(g, phiL, vofL, vof, epsi, BetaP, BetaM, f, (int)(pc / J) + 1,
                            (int)(pc % J) + 1, (int)(pos % K) + 1, &beta);
            }
            const int c = dir == 0 ? i : (dir == 1 ? j : k);
            const double DGrid = dg[(minus ? (c - 1 > 1 ? c - 1 : 1) : c) - 1];
            const int64_t hd = minus ? HX(g, i - (dir == 0), j - (dir == 1), k - (dir == 2)) : HX(g, i, j, k);
            T[n] = dp[hd] / DGrid * BetaP * BetaM / beta;
          }
    }
    if (dir == 0) {
      const int dst = minus ? 1 : I, src = minus ? 2 : I - 1;
      for (int k = 1; k <= K; ++k)
        for (int j = 1; j <= J; ++j) T[NX(g, dst, j, k)] = T[NX(g, src, j, k)];
    } else if (dir == 1) {
      const int dst = minus ? 1 : J, src = minus ? 2 : J - 1;
      for (int k = 1; k <= K; ++k)
        for (int i = 1; i <= I; ++i) T[NX(g, i, dst, k)] = T[NX(g, i, src, k)];
    } else {
      const int dst = minus ? 1 : K, src = minus ? 2 : K - 1;
      for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) T[NX(g, i, j, dst)] = T[NX(g, i, j, src)];
    }
  }
  free(last);
  free(pre);
}

/* ------------------------------------------------------------------ */
/* Assembled rows: ProjectionP.f90:330-478                              */
/* ------------------------------------------------------------------ */
void orc_matrix_rows(const orc_grid* g, const double* pc, const double* ee, const double* ne,
                     const double* te, const int32_t* ct, const int32_t* fl, double* vals,
                     int64_t* cols) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K;
  const double *P1 = pc, *P2 = pc + N, *P3 = pc + 2 * N, *P4 = pc + 3 * N, *P5 = pc + 4 * N,
               *P6 = pc + 5 * N;
  const int WB = fl[0], EB = fl[1], SB = fl[2], NB = fl[3], BB = fl[4], TB = fl[5];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k);
        double* v = vals + 7 * n;
        int64_t* c = cols + 7 * n;
        for (int q = 0; q < 7; ++q) { v[q] = 0.0; c[q] = -1; }
        if (ct[n] == 2) continue;                                   /* :333 */
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        if (i > 1 && ct[NX(g, i - 1, j, k)] != 2) {                /* :348-356 */
          c[0] = NX(g, i - 1, j, k);
          v[0] = -P1[n] * ee[HX(g, i - 1, j, k)] * dy * dz;
        }
        if (j > 1 && ct[NX(g, i, j - 1, k)] != 2) {                /* :358-366 */
          c[1] = NX(g, i, j - 1, k);
          v[1] = -P2[n] * ne[HX(g, i, j - 1, k)] * dx * dz;
        }
        if (k > 1 && ct[NX(g, i, j, k - 1)] != 2) {                /* :368-376 */
          c[2] = NX(g, i, j, k - 1);
          v[2] = -P3[n] * te[HX(g, i, j, k - 1)] * dx * dy;
        }
        double d = P1[n] * ee[HX(g, i - 1, j, k)] * dy * dz +        /* :379-390 */
                   P2[n] * ne[HX(g, i, j - 1, k)] * dx * dz +
                   P3[n] * te[HX(g, i, j, k - 1)] * dx * dy +
                   P4[n] * ee[HX(g, i, j, k)] * dy * dz +
                   P5[n] * ne[HX(g, i, j, k)] * dx * dz +
                   P6[n] * te[HX(g, i, j, k)] * dx * dy;
        if (WB == 1 && i == 1) d = d - P1[n] * ee[HX(g, i - 1, j, k)] * dy * dz; /* :407-430 */
        if (SB == 1 && j == 1) d = d - P2[n] * ne[HX(g, i, j - 1, k)] * dx * dz;
        if (BB == 1 && k == 1) d = d - P3[n] * te[HX(g, i, j, k - 1)] * dx * dy;
        if (EB == 1 && i == I) d = d - P4[n] * ee[HX(g, i, j, k)] * dy * dz;
        if (NB == 1 && j == J) d = d - P5[n] * ne[HX(g, i, j, k)] * dx * dz;
        if (TB == 1 && k == K) d = d - P6[n] * te[HX(g, i, j, k)] * dx * dy;
        c[3] = n;
        v[3] = d;
        if (i < I && ct[NX(g, i + 1, j, k)] != 2) {                /* :433-441 */
          c[4] = NX(g, i + 1, j, k);
          v[4] = -P4[n] * ee[HX(g, i, j, k)] * dy * dz;
        }
        if (j < J && ct[NX(g, i, j + 1, k)] != 2) {                /* :443-451 */
          c[5] = NX(g, i, j + 1, k);
          v[5] = -P5[n] * ne[HX(g, i, j, k)] * dx * dz;
        }
        if (k < K && ct[NX(g, i, j, k + 1)] != 2) {                /* :453-461 */
          c[6] = NX(g, i, j, k + 1);
          v[6] = -P6[n] * te[HX(g, i, j, k)] * dx * dy;
        }
      }
}

/* RHS: ProjectionP.f90:513-564 */
void orc_rhs(const orc_grid* g, const double* pc, const double* ee, const double* ne,
             const double* te, const int32_t* ct, const int32_t* fl, const double* const* bv,
             const double* u, const double* v, const double* w, const double* vof,
             const double* vofO, double dt, double* b) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K;
  const double *P1 = pc, *P2 = pc + N, *P3 = pc + 2 * N, *P4 = pc + 3 * N, *P5 = pc + 4 * N,
               *P6 = pc + 5 * N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k);
        if (ct[n] == 2) { b[n] = 0.0; continue; }
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        const int64_t h = HX(g, i, j, k);
        double r = -dy * dz * (ee[h] * u[h] - ee[HX(g, i - 1, j, k)] * u[HX(g, i - 1, j, k)]) -
                   dx * dz * (ne[h] * v[h] - ne[HX(g, i, j - 1, k)] * v[HX(g, i, j - 1, k)]) -
                   dx * dy * (te[h] * w[h] - te[HX(g, i, j, k - 1)] * w[HX(g, i, j, k - 1)]);
        r = r + (vof[h] - vofO[h]) * dx * dy * dz / dt;                  /* :527 */
        if (fl[0] == 0 && i == 1) r = r + P1[n] * bv[0][(j - 1) + (int64_t)J * (k - 1)]; /* :529-546 */
        if (fl[1] == 0 && i == I) r = r + P4[n] * bv[1][(j - 1) + (int64_t)J * (k - 1)];
        if (fl[2] == 0 && j == 1) r = r + P2[n] * bv[2][(i - 1) + (int64_t)I * (k - 1)];
        if (fl[3] == 0 && j == J) r = r + P5[n] * bv[3][(i - 1) + (int64_t)I * (k - 1)];
        if (fl[4] == 0 && k == 1) r = r + P3[n] * bv[4][(i - 1) + (int64_t)I * (j - 1)];
        if (fl[5] == 0 && k == K) r = r + P6[n] * bv[5][(i - 1) + (int64_t)I * (j - 1)];
        b[n] = r;
      }
}

void orc_rows_apply(int64_t n, const double* vals, const int64_t* cols, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (int q = 0; q < 7; ++q)
      if (cols[7 * r + q] >= 0) s += vals[7 * r + q] * x[cols[7 * r + q]];
    y[r] = s;
  }
}

/* ------------------------------------------------------------------ */
/* Symmetric face operator of the engine                               */
/* ------------------------------------------------------------------ */
void orc_sym_operator(const orc_grid* g, const double* pc, const double* ee, const double* ne,
                      const double* te, const int32_t* ct, const int32_t* fl, double* cx,
                      double* cy, double* cz, double* e) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K;
  const double *P1 = pc, *P2 = pc + N, *P3 = pc + 2 * N, *P4 = pc + 3 * N, *P5 = pc + 4 * N,
               *P6 = pc + 5 * N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k);
        cx[n] = cy[n] = cz[n] = e[n] = 0.0;
        if (ct[n] == 2) continue;
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        const int64_t h = HX(g, i, j, k);
        if (i < I && ct[NX(g, i + 1, j, k)] != 2) cx[n] = P4[n] * ee[h] * dy * dz;
        if (j < J && ct[NX(g, i, j + 1, k)] != 2) cy[n] = P5[n] * ne[h] * dx * dz;
        if (k < K && ct[NX(g, i, j, k + 1)] != 2) cz[n] = P6[n] * te[h] * dx * dy;
        double ex = 0.0;
        if (i == 1 ? fl[0] == 0 : ct[NX(g, i - 1, j, k)] == 2)
          ex = ex + P1[n] * ee[HX(g, i - 1, j, k)] * dy * dz;
        if (j == 1 ? fl[2] == 0 : ct[NX(g, i, j - 1, k)] == 2)
          ex = ex + P2[n] * ne[HX(g, i, j - 1, k)] * dx * dz;
        if (k == 1 ? fl[4] == 0 : ct[NX(g, i, j, k - 1)] == 2)
          ex = ex + P3[n] * te[HX(g, i, j, k - 1)] * dx * dy;
        if (i == I ? fl[1] == 0 : ct[NX(g, i + 1, j, k)] == 2) ex = ex + P4[n] * ee[h] * dy * dz;
        if (j == J ? fl[3] == 0 : ct[NX(g, i, j + 1, k)] == 2) ex = ex + P5[n] * ne[h] * dx * dz;
        if (k == K ? fl[5] == 0 : ct[NX(g, i, j, k + 1)] == 2) ex = ex + P6[n] * te[h] * dx * dy;
        e[n] = ex;
      }
}

/* d = ((((cW + cS) + cB) + cE) + cN) + cT + e; inactive cells get d = 0 */
static int64_t sym_diag_level(int nx, int ny, int nz, const double* cx, const double* cy,
                              const double* cz, const double* e, double* d) {
  int64_t active = 0;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : active)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t n = i + sy * j + sz * k;
        double s = (i > 0 ? cx[n - 1] : 0.0);
        s = s + (j > 0 ? cy[n - sy] : 0.0);
        s = s + (k > 0 ? cz[n - sz] : 0.0);
        s = s + cx[n];
        s = s + cy[n];
        s = s + cz[n];
        s = s + e[n];
        d[n] = s > 0.0 ? s : 0.0;
        active += s > 0.0;
      }
  return active;
}

int64_t orc_sym_diag(int I, int J, int K, const double* cx, const double* cy, const double* cz,
                     const double* e, double* d) {
  return sym_diag_level(I, J, K, cx, cy, cz, e, d);
}

static void sym_apply_level(int nx, int ny, int nz, const double* cx, const double* cy,
                            const double* cz, const double* d, const double* x, double* y) {
  const int64_t sy = nx, sz = (int64_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t n = i + sy * j + sz * k;
        if (d[n] == 0.0) { y[n] = 0.0; continue; }
        double s = d[n] * x[n];
        if (i > 0) s -= cx[n - 1] * x[n - 1];
        if (j > 0) s -= cy[n - sy] * x[n - sy];
        if (k > 0) s -= cz[n - sz] * x[n - sz];
        if (i < nx - 1) s -= cx[n] * x[n + 1];
        if (j < ny - 1) s -= cy[n] * x[n + sy];
        if (k < nz - 1) s -= cz[n] * x[n + sz];
        y[n] = s;
      }
}

void orc_sym_apply(int I, int J, int K, const double* cx, const double* cy, const double* cz,
                   const double* d, const double* x, double* y) {
  sym_apply_level(I, J, K, cx, cy, cz, d, x, y);
}

/* ------------------------------------------------------------------ */
/* Cell-centred multigrid (piecewise-constant transfer, summed faces)  */
/* ------------------------------------------------------------------ */
typedef struct mg_level {
  int nx, ny, nz;
  int64_t n, active;
  double *cx, *cy, *cz, *e, *d, *x, *b, *r;
} mg_level;

typedef struct mg_hier {
  int L, singular;
  mg_level lv[24];
} mg_hier;

static double* dalloc(int64_t n) {
  double* p = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  return p;
}

int orc_auto_levels(int I, int J, int K) {
  int L = 1, nx = I, ny = J, nz = K;
  while (nx >= 4 && ny >= 4 && nz >= 4 && L < 24) {
    nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
    ++L;
  }
  return L;
}

static void mg_build(mg_hier* H, int I, int J, int K, const double* cx, const double* cy,
                     const double* cz, const double* e, int levels, double scale) {
  memset(H, 0, sizeof(*H));
  int L = levels > 0 ? levels : orc_auto_levels(I, J, K);
  int maxL = orc_auto_levels(I, J, K);
  if (L > maxL) L = maxL;
  H->L = L;
  for (int l = 0; l < L; ++l) {
    mg_level* A = &H->lv[l];
    if (l == 0) { A->nx = I; A->ny = J; A->nz = K; }
    else {
      A->nx = (H->lv[l - 1].nx + 1) / 2;
      A->ny = (H->lv[l - 1].ny + 1) / 2;
      A->nz = (H->lv[l - 1].nz + 1) / 2;
    }
    A->n = (int64_t)A->nx * A->ny * A->nz;
    A->cx = dalloc(A->n); A->cy = dalloc(A->n); A->cz = dalloc(A->n); A->e = dalloc(A->n);
    A->d = dalloc(A->n); A->x = dalloc(A->n); A->b = dalloc(A->n); A->r = dalloc(A->n);
    if (l == 0) {
      memcpy(A->cx, cx, sizeof(double) * A->n);
      memcpy(A->cy, cy, sizeof(double) * A->n);
      memcpy(A->cz, cz, sizeof(double) * A->n);
      memcpy(A->e, e, sizeof(double) * A->n);
    } else {
      const mg_level* F = &H->lv[l - 1];
      const int64_t fsy = F->nx, fsz = (int64_t)F->nx * F->ny;
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 0; k < A->nz; ++k)
        for (int j = 0; j < A->ny; ++j)
          for (int i = 0; i < A->nx; ++i) {
            const int64_t n = i + (int64_t)A->nx * (j + (int64_t)A->ny * k);
            double sx = 0, sy = 0, sz = 0, se = 0;
            for (int c = 0; c < 2; ++c)
              for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                  const int fi = 2 * i + a, fj = 2 * j + b, fk = 2 * k + c;
                  if (fi >= F->nx || fj >= F->ny || fk >= F->nz) continue;
                  const int64_t m = fi + fsy * fj + fsz * fk;
                  se += F->e[m];
                  if (a == 1) sx += F->cx[m];
                  if (b == 1) sy += F->cy[m];
                  if (c == 1) sz += F->cz[m];
                }
            A->cx[n] = scale * sx; A->cy[n] = scale * sy; A->cz[n] = scale * sz;
            A->e[n] = scale * se;
          }
    }
    A->active = sym_diag_level(A->nx, A->ny, A->nz, A->cx, A->cy, A->cz, A->e, A->d);
  }
  double esum = 0;
  for (int64_t q = 0; q < H->lv[0].n; ++q) esum += H->lv[0].d[q] > 0 ? fabs(H->lv[0].e[q]) : 0.0;
  H->singular = esum == 0.0;
}

static void mg_free(mg_hier* H) {
  for (int l = 0; l < H->L; ++l) {
    mg_level* A = &H->lv[l];
    free(A->cx); free(A->cy); free(A->cz); free(A->e); free(A->d); free(A->x); free(A->b);
    free(A->r);
  }
}

/* one colour of Gauss-Seidel, over-relaxed: x_i += omega ((b_i + sum c x_nb)/d_i - x_i) */
static void gs_color(const mg_level* A, int color, double omega) {
  const int nx = A->nx, ny = A->ny, nz = A->nz;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const double *cx = A->cx, *cy = A->cy, *cz = A->cz, *d = A->d, *b = A->b;
  double* x = A->x;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = ((j + k + color) & 1); i < nx; i += 2) {
        const int64_t n = i + sy * j + sz * k;
        if (d[n] == 0.0) continue;
        double s = b[n];
        if (i > 0) s += cx[n - 1] * x[n - 1];
        if (j > 0) s += cy[n - sy] * x[n - sy];
        if (k > 0) s += cz[n - sz] * x[n - sz];
        if (i < nx - 1) s += cx[n] * x[n + 1];
        if (j < ny - 1) s += cy[n] * x[n + sy];
        if (k < nz - 1) s += cz[n] * x[n + sz];
        x[n] = x[n] + omega * (s / d[n] - x[n]);
      }
}

static void project_mean(const mg_level* A, double* v) {
  double s = 0;
#pragma omp parallel for reduction(+ : s)
  for (int64_t q = 0; q < A->n; ++q) s += A->d[q] > 0 ? v[q] : 0.0;
  if (A->active == 0) return;
  const double m = s / (double)A->active;
#pragma omp parallel for
  for (int64_t q = 0; q < A->n; ++q) if (A->d[q] > 0) v[q] -= m;
}

static void vcycle(mg_hier* H, int l, const orc_solver_opts* o) {
  mg_level* A = &H->lv[l];
  const double om = l == 0 ? o->omega : o->omega_coarse;
  memset(A->x, 0, sizeof(double) * A->n);
  if (l == H->L - 1) {
    if (H->singular) project_mean(A, A->b);
    for (int s = 0; s < o->coarse_sweeps; ++s) { gs_color(A, 0, om); gs_color(A, 1, om); }
    for (int s = 0; s < o->coarse_sweeps; ++s) { gs_color(A, 1, om); gs_color(A, 0, om); }
    if (H->singular) project_mean(A, A->x);
    return;
  }
  for (int s = 0; s < o->pre; ++s) { gs_color(A, 0, om); gs_color(A, 1, om); }
  sym_apply_level(A->nx, A->ny, A->nz, A->cx, A->cy, A->cz, A->d, A->x, A->r);
  mg_level* C = &H->lv[l + 1];
  const int64_t fsy = A->nx, fsz = (int64_t)A->nx * A->ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < C->nz; ++k)
    for (int j = 0; j < C->ny; ++j)
      for (int i = 0; i < C->nx; ++i) {
        double s = 0;
        for (int c = 0; c < 2; ++c)
          for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
              const int fi = 2 * i + a, fj = 2 * j + b, fk = 2 * k + c;
              if (fi >= A->nx || fj >= A->ny || fk >= A->nz) continue;
              const int64_t m = fi + fsy * fj + fsz * fk;
              if (A->d[m] > 0) s += A->b[m] - A->r[m];
            }
        C->b[i + (int64_t)C->nx * (j + (int64_t)C->ny * k)] = s;
      }
  vcycle(H, l + 1, o);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < A->nz; ++k)
    for (int j = 0; j < A->ny; ++j)
      for (int i = 0; i < A->nx; ++i) {
        const int64_t m = i + fsy * j + fsz * k;
        if (A->d[m] > 0) A->x[m] += C->x[(i / 2) + (int64_t)C->nx * ((j / 2) + (int64_t)C->ny * (k / 2))];
      }
  for (int s = 0; s < o->post; ++s) { gs_color(A, 1, om); gs_color(A, 0, om); }
}

static double dot_active(int64_t n, const double* d, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t q = 0; q < n; ++q) s += d[q] > 0 ? a[q] * b[q] : 0.0;
  return s;
}

static double max_div(const orc_grid* g, const double* d, const double* r) {
  const int I = g->I, J = g->J, K = g->K;
  double m = 0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < J; ++j)
      for (int i = 0; i < I; ++i) {
        const int64_t n = i + (int64_t)I * (j + (int64_t)J * k);
        if (d[n] > 0) {
          const double v = fabs(r[n]) / (g->dxp[i] * g->dyp[j] * g->dzp[k]);
          if (v > m) m = v;
        }
      }
  return m;
}

int orc_pcg(const orc_grid* g, const double* cx, const double* cy, const double* cz,
            const double* e, const double* vals, const int64_t* cols, const double* b0,
            const orc_solver_opts* o, double* x, orc_solver_stats* st) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t n = (int64_t)I * J * K;
  double t0 = omp_get_wtime();
  mg_hier H;
  mg_build(&H, I, J, K, cx, cy, cz, e, o->levels, o->coarse_scale);
  const mg_level* F = &H.lv[0];
  const double* d = F->d;
  double *r = dalloc(n), *p = dalloc(n), *q = dalloc(n), *z = dalloc(n);
  for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = d[i] > 0 ? b0[i] : 0.0; }
  if (H.singular) project_mean(F, r);
  int it = 0, conv = 0;
  double rmax = max_div(g, d, r);
  st->initial_div = rmax;
  if (rmax <= o->div_tol) conv = 1;
  double rz = 0;
  mg_level* M = &H.lv[0];
  /* preconditioner */
#define APPLY_PREC(src, dst)                                                   \
  do {                                                                         \
    if (o->precond == 2) {                                                     \
      memcpy(M->b, src, sizeof(double) * n);                                   \
      vcycle(&H, 0, o);                                                        \
      memcpy(dst, M->x, sizeof(double) * n);                                   \
    } else if (o->precond == 1) {                                              \
      for (int64_t i_ = 0; i_ < n; ++i_) dst[i_] = d[i_] > 0 ? src[i_] / d[i_] : 0.0; \
    } else {                                                                   \
      memcpy(dst, src, sizeof(double) * n);                                    \
    }                                                                          \
  } while (0)
  /* singular systems: shift z by its diagonal-weighted mean so the iterate
   * keeps the stiff phase's pressure near zero (representation floor) */
  double dsum = 0;
  for (int64_t i = 0; i < n; ++i) dsum += d[i] > 0 ? d[i] : 0.0;
#define SHIFT_Z()                                                              \
  do {                                                                         \
    if (H.singular && dsum > 0) {                                              \
      const double m_ = dot_active(n, d, d, z) / dsum;                         \
      for (int64_t i_ = 0; i_ < n; ++i_) if (d[i_] > 0) z[i_] -= m_;           \
    }                                                                          \
  } while (0)
  if (!conv) {
    APPLY_PREC(r, z);
    SHIFT_Z();
    memcpy(p, z, sizeof(double) * n);
    rz = dot_active(n, d, r, z);
  }
  while (!conv && it < o->max_iter) {
    if (o->use_rows) orc_rows_apply(n, vals, cols, p, q);
    else sym_apply_level(I, J, K, F->cx, F->cy, F->cz, d, p, q);
    const double pq = dot_active(n, d, p, q);
    const double alpha = rz / pq;
    /* breakdown below the FP64 floor (solves run to a tolerance they cannot
     * reach): keep the last finite iterate */
    if (!(pq > 0.0) || !isfinite(alpha)) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      if (d[i] > 0) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    ++it;
    rmax = max_div(g, d, r);
    if (rmax <= o->div_tol) { conv = 1; break; }
    APPLY_PREC(r, z);
    SHIFT_Z();
    const double rz_new = dot_active(n, d, r, z);
    const double zq = dot_active(n, d, z, q);
    const double beta = -alpha * zq / rz;
    if (!(rz_new != 0.0) || !isfinite(beta)) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = d[i] > 0 ? z[i] + beta * p[i] : 0.0;
    rz = rz_new;
  }
#undef APPLY_PREC
#undef SHIFT_Z
  st->iterations = it;
  st->converged = conv;
  st->final_div = rmax;
  st->levels = H.L;
  st->seconds = omp_get_wtime() - t0;
  free(r); free(p); free(q); free(z);
  const int singular = H.singular;
  mg_free(&H);
  return singular;
}

void orc_gauge(int I, int J, int K, const double* d, int singular, int pi, int pj, int pk,
               double* x) {
  if (!singular) return;
  const int64_t n = (int64_t)I * J * K;
  double shift;
  if (pi >= 1 && pj >= 1 && pk >= 1 && pi <= I && pj <= J && pk <= K &&
      d[(pi - 1) + (int64_t)I * ((pj - 1) + (int64_t)J * (pk - 1))] > 0) {
    shift = x[(pi - 1) + (int64_t)I * ((pj - 1) + (int64_t)J * (pk - 1))];
  } else {
    double s = 0; int64_t c = 0;
    for (int64_t q = 0; q < n; ++q) if (d[q] > 0) { s += x[q]; ++c; }
    shift = c ? s / (double)c : 0.0;
  }
  for (int64_t q = 0; q < n; ++q) if (d[q] > 0) x[q] -= shift;
}

/* ProjectionP.f90:591-616 gather (solid cells 0) and :620-643 ghosts */
void orc_pressure_field(const orc_grid* g, const int32_t* ct, const int32_t* fl, const double* x,
                        double* Pp) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  for (int64_t q = 0; q < Nh; ++q) Pp[q] = 0.0;
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k);
        Pp[HX(g, i, j, k)] = ct[n] != 2 ? x[n] : 0.0;
      }
  for (int j = 1; j <= J; ++j)
    for (int k = 1; k <= K; ++k) {
      Pp[HX(g, 0, j, k)] = (double)fl[0] * Pp[HX(g, 1, j, k)];
      Pp[HX(g, I + 1, j, k)] = (double)fl[1] * Pp[HX(g, I, j, k)];
    }
  for (int i = 1; i <= I; ++i)
    for (int k = 1; k <= K; ++k) {
      Pp[HX(g, i, 0, k)] = (double)fl[2] * Pp[HX(g, i, 1, k)];
      Pp[HX(g, i, J + 1, k)] = (double)fl[3] * Pp[HX(g, i, J, k)];
    }
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j) {
      Pp[HX(g, i, j, 0)] = (double)fl[4] * Pp[HX(g, i, j, 1)];
      Pp[HX(g, i, j, K + 1)] = (double)fl[5] * Pp[HX(g, i, j, K)];
    }
}

/* ComputePUVW.f90:181-232 */
void orc_correct(const orc_grid* g, const double* pc, const double* Pp, const int32_t* ctu,
                 const int32_t* ctv, const int32_t* ctw, const int32_t* mu, const int32_t* mv,
                 const int32_t* mw, const double* us, const double* vs, const double* ws,
                 double* u, double* v, double* w) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K;
  const double *P4 = pc + 3 * N, *P5 = pc + 4 * N, *P6 = pc + 5 * N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k), h = HX(g, i, j, k);
        u[h] = (mu[n] != 1 && ctu[n] != 2) ? us[h] - (Pp[HX(g, i + 1, j, k)] - Pp[h]) * P4[n] : 0.0;
        v[h] = (mv[n] != 1 && ctv[n] != 2) ? vs[h] - (Pp[HX(g, i, j + 1, k)] - Pp[h]) * P5[n] : 0.0;
        w[h] = (mw[n] != 1 && ctw[n] != 2) ? ws[h] - (Pp[HX(g, i, j, k + 1)] - Pp[h]) * P6[n] : 0.0;
      }
}

/* Solver.f90:236-250 */
double orc_mres(const orc_grid* g, const double* ee, const double* ne, const double* te,
                const double* u, const double* v, const double* w, double* mres) {
  const int I = g->I, J = g->J, K = g->K;
  double m = 0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        const int64_t h = HX(g, i, j, k);
        const double r = dy * dz * (ee[h] * u[h] - ee[HX(g, i - 1, j, k)] * u[HX(g, i - 1, j, k)]) +
                         dx * dz * (ne[h] * v[h] - ne[HX(g, i, j - 1, k)] * v[HX(g, i, j - 1, k)]) +
                         dx * dy * (te[h] * w[h] - te[HX(g, i, j, k - 1)] * w[HX(g, i, j, k - 1)]);
        if (mres) mres[NX(g, i, j, k)] = r;
        const double a = fabs(r) / (dx * dy * dz);
        if (a > m) m = a;
      }
  return m;
}

/* ------------------------------------------------------------------ */
/* Adjoint (derivation in the header of tests/test_oracle_adjoint.py)  */
/* ------------------------------------------------------------------ */
int orc_adjoint(const orc_grid* g, const double* pc, const double* ee, const double* ne,
                const double* te, const int32_t* ct, const int32_t* fl, const int32_t* ctu,
                const int32_t* ctv, const int32_t* ctw, const int32_t* mu, const int32_t* mv,
                const int32_t* mw, const double* Pp, const double* gu, const double* gv,
                const double* gw, const double* gp, int pin_i, int pin_j, int pin_k,
                const orc_solver_opts* o, double* gus, double* gvs, double* gws, double* gkx,
                double* gky, double* gkz, double* lambda_out) {
  const int I = g->I, J = g->J, K = g->K;
  const int64_t N = (int64_t)I * J * K, Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  const double *P4 = pc + 3 * N, *P5 = pc + 4 * N, *P6 = pc + 5 * N;
  double *cx = dalloc(N), *cy = dalloc(N), *cz = dalloc(N), *e = dalloc(N), *d = dalloc(N);
  orc_sym_operator(g, pc, ee, ne, te, ct, fl, cx, cy, cz, e);
  sym_diag_level(I, J, K, cx, cy, cz, e, d);
  double* pbar = dalloc(N);
  for (int64_t q = 0; q < Nh; ++q) { gus[q] = 0; gvs[q] = 0; gws[q] = 0; }
  /* correction part: u = m (u* - (P(i+1) - P(i)) K) */
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k), h = HX(g, i, j, k);
        const int m_u = mu[n] != 1 && ctu[n] != 2, m_v = mv[n] != 1 && ctv[n] != 2,
                  m_w = mw[n] != 1 && ctw[n] != 2;
        const double tu = m_u ? gu[h] : 0.0, tv = m_v ? gv[h] : 0.0, tw = m_w ? gw[h] : 0.0;
        gus[h] = tu; gvs[h] = tv; gws[h] = tw;
        gkx[n] = -tu * (Pp[HX(g, i + 1, j, k)] - Pp[h]);
        gky[n] = -tv * (Pp[HX(g, i, j + 1, k)] - Pp[h]);
        gkz[n] = -tw * (Pp[HX(g, i, j, k + 1)] - Pp[h]);
        /* dP(i+1) gets -t*K, dP(i) gets +t*K; the ghost P(I+1) = flag*P(I) */
        if (d[n] > 0) pbar[n] += tu * P4[n] + tv * P5[n] + tw * P6[n];
        if (i < I) { if (d[NX(g, i + 1, j, k)] > 0) pbar[NX(g, i + 1, j, k)] -= tu * P4[n]; }
        else if (d[n] > 0) pbar[n] -= (double)fl[1] * tu * P4[n];
        if (j < J) { if (d[NX(g, i, j + 1, k)] > 0) pbar[NX(g, i, j + 1, k)] -= tv * P5[n]; }
        else if (d[n] > 0) pbar[n] -= (double)fl[3] * tv * P5[n];
        if (k < K) { if (d[NX(g, i, j, k + 1)] > 0) pbar[NX(g, i, j, k + 1)] -= tw * P6[n]; }
        else if (d[n] > 0) pbar[n] -= (double)fl[5] * tw * P6[n];
      }
  /* pressure cotangent: Pp = x - shift on active cells, the face ghosts
   * flag * Pp of the inner cell (orc_pressure_field), the shift x(pin) or
   * the active-cell mean of x (orc_gauge) in singular systems */
  if (gp) {
    double esum = 0.0;
    for (int64_t q = 0; q < N; ++q) esum += e[q];
    double qsum = 0.0;
    int64_t nact = 0;
    for (int k = 1; k <= K; ++k)
      for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) {
          const int64_t n = NX(g, i, j, k);
          if (!(d[n] > 0)) continue;
          double q = gp[HX(g, i, j, k)];
          if (i == 1) q += (double)fl[0] * gp[HX(g, 0, j, k)];
          if (i == I) q += (double)fl[1] * gp[HX(g, I + 1, j, k)];
          if (j == 1) q += (double)fl[2] * gp[HX(g, i, 0, k)];
          if (j == J) q += (double)fl[3] * gp[HX(g, i, J + 1, k)];
          if (k == 1) q += (double)fl[4] * gp[HX(g, i, j, 0)];
          if (k == K) q += (double)fl[5] * gp[HX(g, i, j, K + 1)];
          pbar[n] += q;
          qsum += q;
          ++nact;
        }
    if (esum == 0.0 && nact > 0) {
      const int pinned = pin_i >= 1 && pin_j >= 1 && pin_k >= 1 && pin_i <= I && pin_j <= J &&
                         pin_k <= K && d[NX(g, pin_i, pin_j, pin_k)] > 0;
      if (pinned) {
        pbar[NX(g, pin_i, pin_j, pin_k)] -= qsum;
      } else {
        for (int64_t n = 0; n < N; ++n)
          if (d[n] > 0) pbar[n] -= qsum / (double)nact;
      }
    }
  }
  /* lambda = A^{-1} pbar */
  double* lam = dalloc(N);
  orc_solver_stats st;
  orc_solver_opts oo = *o;
  oo.use_rows = 0;
  const int singular = orc_pcg(g, cx, cy, cz, e, NULL, NULL, pbar, &oo, lam, &st);
  orc_gauge(I, J, K, d, singular, 0, 0, 0, lam);
  /* RHS part: b(i) = -dy dz (EE(i) u*(i) - EE(i-1) u*(i-1)) - ... */
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 0; i <= I; ++i) {
        const double lp = (i >= 1 && d[NX(g, i, j, k)] > 0) ? lam[NX(g, i, j, k)] : 0.0;
        const double ln = (i + 1 <= I && d[NX(g, i + 1, j, k)] > 0) ? lam[NX(g, i + 1, j, k)] : 0.0;
        const double area = g->dyp[j - 1] * g->dzp[k - 1];
        gus[HX(g, i, j, k)] += area * ee[HX(g, i, j, k)] * (ln - lp);
      }
  for (int k = 1; k <= K; ++k)
    for (int j = 0; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const double lp = (j >= 1 && d[NX(g, i, j, k)] > 0) ? lam[NX(g, i, j, k)] : 0.0;
        const double ln = (j + 1 <= J && d[NX(g, i, j + 1, k)] > 0) ? lam[NX(g, i, j + 1, k)] : 0.0;
        const double area = g->dxp[i - 1] * g->dzp[k - 1];
        gvs[HX(g, i, j, k)] += area * ne[HX(g, i, j, k)] * (ln - lp);
      }
  for (int k = 0; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const double lp = (k >= 1 && d[NX(g, i, j, k)] > 0) ? lam[NX(g, i, j, k)] : 0.0;
        const double ln = (k + 1 <= K && d[NX(g, i, j, k + 1)] > 0) ? lam[NX(g, i, j, k + 1)] : 0.0;
        const double area = g->dxp[i - 1] * g->dyp[j - 1];
        gws[HX(g, i, j, k)] += area * te[HX(g, i, j, k)] * (ln - lp);
      }
  /* operator part: dL/dc_f = -(lam_N - lam_P)(P_N - P_P) on interior faces between active cells */
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k), h = HX(g, i, j, k);
        if (d[n] <= 0) continue;
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        if (i < I && ct[NX(g, i + 1, j, k)] != 2 && d[NX(g, i + 1, j, k)] > 0) {
          const int64_t m = NX(g, i + 1, j, k);
          gkx[n] -= ee[h] * dy * dz * (lam[m] - lam[n]) * (Pp[HX(g, i + 1, j, k)] - Pp[h]);
        }
        if (j < J && ct[NX(g, i, j + 1, k)] != 2 && d[NX(g, i, j + 1, k)] > 0) {
          const int64_t m = NX(g, i, j + 1, k);
          gky[n] -= ne[h] * dx * dz * (lam[m] - lam[n]) * (Pp[HX(g, i, j + 1, k)] - Pp[h]);
        }
        if (k < K && ct[NX(g, i, j, k + 1)] != 2 && d[NX(g, i, j, k + 1)] > 0) {
          const int64_t m = NX(g, i, j, k + 1);
          gkz[n] -= te[h] * dx * dy * (lam[m] - lam[n]) * (Pp[HX(g, i, j, k + 1)] - Pp[h]);
        }
      }
  if (lambda_out) memcpy(lambda_out, lam, sizeof(double) * N);
  free(cx); free(cy); free(cz); free(e); free(d); free(pbar); free(lam);
  return st.converged ? 0 : -3;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ */
/* Post-correction wall step of UpdatePUVW                             */
/* ------------------------------------------------------------------ */
/* TypicalBC (BoundaryFunction.f90:760-781) as BCx%West..Top call it from
 * BoundaryConditionVarNew: the "inner" argument is a full halo slice such as
 * Vari%u(1,:,:), passed to an assumed-shape dummy whose element (a,b) is
 * therefore the halo entry (a-1,b-1) of the slice (ComputePUVW.f90:431-435
 * passes Vari%u(1,:,:) with the halo, PGrid%dx(1,:,:) without).  The face
 * value of element (a,b), 1-based, is
 *   Dirichlet: const;  Neumann: X(src, a-1, b-1) - const*dxyz(a,b)/2
 * with dxyz = mult * spacing of the boundary cell (BoundaryFunction.f90:83-513
 * for the multipliers). */
static void wall_face_values(const orc_grid* g, const double* X, int f, int src, double mult,
                             int flag, double cst, double* var) {
  const int I = g->I, J = g->J, K = g->K;
  const int n1 = f < 2 ? J : I, n2 = f < 4 ? K : J;
  for (int b = 1; b <= n2; ++b)
    for (int a = 1; a <= n1; ++a) {
      double* o = var + (a - 1) + (int64_t)n1 * (b - 1);
      if (flag == 0) { *o = cst; continue; }
      int64_t h;
      double d;
      if (f < 2) { h = HX(g, src, a - 1, b - 1); d = f == 0 ? g->dxp[0] : g->dxp[I - 1]; }
      else if (f < 4) { h = HX(g, a - 1, src, b - 1); d = f == 2 ? g->dyp[0] : g->dyp[J - 1]; }
      else { h = HX(g, a - 1, b - 1, src); d = f == 4 ? g->dzp[0] : g->dzp[K - 1]; }
      const double dxyz = mult * d;
      *o = X[h] - cst * dxyz / 2.0;
    }
}

/* halo index of element (a,b) (1-based) of the plane `plane` of face f */
static int64_t wall_plane_index(const orc_grid* g, int f, int plane, int a, int b) {
  if (f < 2) return HX(g, plane, a, b);
  if (f < 4) return HX(g, a, plane, b);
  return HX(g, a, b, plane);
}

/* BoundaryConditionVarNew (ComputePUVW.f90:418-694) for one variable q
 * (0 u, 1 v, 2 w, 3 p), faces in the reference's order W,E,S,N,B,T. */
static void wall_reset_var(const orc_grid* g, int q, double* X, const int32_t* fl,
                           const double* cst, double* const* var) {
  const int I = g->I, J = g->J, K = g->K;
  const int last[6] = {1, I, 1, J, 1, K};       /* boundary cell plane */
  const int ghost[6] = {0, I + 1, 0, J + 1, 0, K + 1};
  for (int f = 0; f < 6; ++f) {
    /* slice the BC routine reads and the spacing multiplier: u E reads
     * u(Imax-1,:,:) with -2 dx (ComputePUVW.f90:440-444,
     * BoundaryFunction.f90:102), v N v(:,Jmax-1,:) with -2 dy (:532-537,
     * :251), w T w(:,:,Kmax-1) with -2 dz (:618-622, :399); every other face
     * reads the boundary cell's own plane with +dx/dy/dz on W/S/B and
     * -dx/dy/dz on E/N/T (BoundaryFunction.f90:83-513). */
    int src = last[f];
    double mult = (f % 2 == 0) ? 1.0 : -1.0;
    if ((q == 0 && f == 1) || (q == 1 && f == 3) || (q == 2 && f == 5)) {
      src = last[f] - 1;
      mult = -2.0;
    }
    const int n1 = f < 2 ? J : I, n2 = f < 4 ? K : J;
    double* V = var[f];
    wall_face_values(g, X, f, src, mult, fl[f], cst[f], V);
    /* direct assignment of the wall-normal face: u(0) and u(Imax) (:437,446),
     * v(:,0,:) and v(:,Jmax,:) (:529,538), w(:,:,0) and w(:,:,Kmax) (:615,624);
     * for p the top cell layer p(:,:,Kmax) (:689) */
    const int assign = (q == 0 && f < 2) || (q == 1 && (f == 2 || f == 3)) ||
                       (q == 2 && f >= 4) || (q == 3 && f == 5);
    const int target = f % 2 == 0 ? 0 : last[f];
    if (assign)
      for (int b = 1; b <= n2; ++b)
        for (int a = 1; a <= n1; ++a)
          X[wall_plane_index(g, f, target, a, b)] = V[(a - 1) + (int64_t)n1 * (b - 1)];
    /* ComputeGhostVarBoundary (:696-708) on every face except u W, v S, w B */
    const int no_ghost = (q == 0 && f == 0) || (q == 1 && f == 2) || (q == 2 && f == 4);
    if (no_ghost) continue;
    for (int b = 1; b <= n2; ++b)
      for (int a = 1; a <= n1; ++a) {
        const double bc = V[(a - 1) + (int64_t)n1 * (b - 1)];
        const double in = X[wall_plane_index(g, f, last[f], a, b)];
        X[wall_plane_index(g, f, ghost[f], a, b)] = fl[f] == 1 ? bc - in : 2.0 * bc - in;
      }
  }
}

/* Global mass fix of the first Dirichlet pressure face (ComputePUVW.f90:256-384,
 * an if/elseif chain W,E,S,N,B,T).  The north branch of the reference loops
 * j over 1..Jmax with k left at its value after the preceding loops (past
 * Kmax), so it reads outside the arrays; the restatement takes the branch's
 * evident intent, the mirror of the south branch at j = Jmax. */
static void wall_mass_fix(const orc_grid* g, const int32_t* flp, const double* ee, const double* ne,
                          const double* te, double* u, double* v, double* w) {
  const int I = g->I, J = g->J, K = g->K;
  int f = 0;
  while (f < 6 && flp[f] != 0) ++f;
  if (f == 6) return;
  if (f < 2) {
    const int i = f == 0 ? 1 : I;
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t h = HX(g, i, j, k);
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        if (f == 0)
          u[HX(g, i - 1, j, k)] = (dx * dz * (ne[h] * v[h] - ne[h] * v[HX(g, i, j - 1, k)]) +
                                   dx * dy * (te[h] * w[h] - te[h] * w[HX(g, i, j, k - 1)]) +
                                   dy * dz * ee[h] * u[h]) / ee[h] / dy / dz;
        else
          u[h] = (dx * dz * (-ne[h] * v[h] + ne[h] * v[HX(g, i, j - 1, k)]) +
                  dx * dy * (-te[h] * w[h] + te[h] * w[HX(g, i, j, k - 1)]) +
                  dy * dz * ee[h] * u[HX(g, i - 1, j, k)]) / ee[h] / dy / dz;
      }
  } else if (f < 4) {
    const int j = f == 2 ? 1 : J;
    for (int i = 1; i <= I; ++i)
      for (int k = 1; k <= K; ++k) {
        const int64_t h = HX(g, i, j, k);
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        if (f == 2)
          v[HX(g, i, j - 1, k)] = (dy * dz * (ee[h] * u[h] - ee[h] * u[HX(g, i - 1, j, k)]) +
                                   dx * dy * (te[h] * w[h] - te[h] * w[HX(g, i, j, k - 1)]) +
                                   dx * dz * ne[h] * v[h]) / ne[h] / dx / dz;
        else
          v[h] = (-dy * dz * (ee[h] * u[h] - ee[h] * u[HX(g, i - 1, j, k)]) -
                  dx * dy * (te[h] * w[h] - te[h] * w[HX(g, i, j, k - 1)]) +
                  dx * dz * ne[h] * v[HX(g, i, j - 1, k)]) / ne[h] / dx / dz;
      }
  } else {
    const int k = f == 4 ? 1 : K;
    for (int i = 1; i <= I; ++i)
      for (int j = 1; j <= J; ++j) {
        const int64_t h = HX(g, i, j, k);
        const double dx = g->dxp[i - 1], dy = g->dyp[j - 1], dz = g->dzp[k - 1];
        if (f == 4)
          w[HX(g, i, j, k - 1)] = (dy * dz * (ee[h] * u[h] - ee[h] * u[HX(g, i - 1, j, k)]) +
                                   dx * dz * (ne[h] * v[h] - ne[h] * v[HX(g, i, j - 1, k)]) +
                                   dx * dy * te[h] * w[h]) / te[h] / dx / dy;
        else
          w[h] = (-dy * dz * (ee[h] * u[h] - ee[h] * u[HX(g, i - 1, j, k)]) -
                  dx * dz * (ne[h] * v[h] - ne[h] * v[HX(g, i, j - 1, k)]) +
                  dx * dy * te[h] * w[HX(g, i, j, k - 1)]) / te[h] / dx / dy;
      }
  }
}

void orc_wall_step(const orc_grid* g, const double* ee, const double* ne, const double* te,
                   const int32_t* ctp, const int32_t* ctu, const int32_t* ctv, const int32_t* ctw,
                   const int32_t* fl, const double* cst, double* u, double* v, double* w,
                   double* p, double* const* var) {
  const int I = g->I, J = g->J, K = g->K;
  double* X[4] = {u, v, w, p};
  const int64_t plane[6] = {(int64_t)J * K, (int64_t)J * K, (int64_t)I * K,
                            (int64_t)I * K, (int64_t)I * J, (int64_t)I * J};
  for (int q = 0; q < 4; ++q) {
    double* V[6];
    for (int f = 0; f < 6; ++f) V[f] = (double*)malloc(sizeof(double) * plane[f]);
    wall_reset_var(g, q, X[q], fl + 6 * q, cst + 6 * q, V);
    for (int f = 0; f < 6; ++f) {
      if (var && var[6 * q + f]) memcpy(var[6 * q + f], V[f], sizeof(double) * plane[f]);
      free(V[f]);
    }
  }
  wall_mass_fix(g, fl + 18, ee, ne, te, u, v, w);
  /* VariablesInternalCellCondition (Solver.f90:467-501) */
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t n = NX(g, i, j, k), h = HX(g, i, j, k);
        if (ctp[n] == 2) p[h] = 0.0;
        if (ctu[n] == 2) u[h] = 0.0;
        if (ctv[n] == 2) v[h] = 0.0;
        if (ctw[n] == 2) w[h] = 0.0;
      }
}

/* ------------------------------------------------------------------ */
/* Cut-cell geometry: GridPreprocess (CutCell.f90:62-325) with         */
/* DefineCell (:329-365), CellGeoCal (:369-671), FaceGeoCal (:883-981). */
/* ------------------------------------------------------------------ */

/* FaceGeoCal (CutCell.f90:883-981): the part of the rectangle
 * [-d1/2,d1/2] x [-d2/2,d2/2] (at offset d3 along the third axis) where
 * n1*x + n2*y + n3*d3 + phi > 0, as an area fraction and a centroid. */
void orc_face_geo(double d1, double d2, double n1, double n2, double n3, double d3, double phi,
                  double* area, double cent[3]) {
  const double d11 = d1 / 2.0, d22 = d2 / 2.0;
  /* node(6,2) in the reference; 8 + 1 here because separately rounded corner
   * values can alternate in sign when the normal is ~1e-17 (the reference
   * then writes past node's end) */
  double nx_[9] = {0}, ny_[9] = {0};
  int t = 0;
  double dpt[4];
  dpt[0] = d11 * n1 - d22 * n2 + d3 * n3 + phi;   /* :900-903 */
  dpt[1] = -d11 * n1 - d22 * n2 + d3 * n3 + phi;
  dpt[2] = -d11 * n1 + d22 * n2 + d3 * n3 + phi;
  dpt[3] = d11 * n1 + d22 * n2 + d3 * n3 + phi;
  if (dpt[0] > 0.0) { nx_[t] = d11; ny_[t] = -d22; ++t; }                                  /* :905 */
  if (dpt[0] * dpt[1] < 0.0) { nx_[t] = (d22 * n2 - d3 * n3 - phi) / n1; ny_[t] = -d22; ++t; }
  if (dpt[1] > 0.0) { nx_[t] = -d11; ny_[t] = -d22; ++t; }
  if (dpt[1] * dpt[2] < 0.0) { nx_[t] = -d11; ny_[t] = (d11 * n1 - d3 * n3 - phi) / n2; ++t; }
  if (dpt[2] > 0.0) { nx_[t] = -d11; ny_[t] = d22; ++t; }
  if (dpt[2] * dpt[3] < 0.0) { nx_[t] = (-d22 * n2 - d3 * n3 - phi) / n1; ny_[t] = d22; ++t; }
  if (dpt[3] > 0.0) { nx_[t] = d11; ny_[t] = d22; ++t; }
  if (dpt[3] * dpt[0] < 0.0) { nx_[t] = d11; ny_[t] = (-d11 * n1 - d3 * n3 - phi) / n2; ++t; }
  nx_[t] = nx_[0]; ny_[t] = ny_[0];                                                           /* :958 */
  double vol = 0.0;
  for (int k = 0; k < t; ++k) vol = vol + 0.5 * (nx_[k] * ny_[k + 1] - nx_[k + 1] * ny_[k]);
  *area = fabs(vol / (d1 * d2));
  cent[0] = 0.0; cent[1] = 0.0;
  if (*area > 0.0) {
    for (int k = 0; k < t; ++k) {
      const double cr = nx_[k] * ny_[k + 1] - nx_[k + 1] * ny_[k];
      cent[0] = cent[0] + 1.0 / 6.0 / vol * (nx_[k] + nx_[k + 1]) * cr;
      cent[1] = cent[1] + 1.0 / 6.0 / vol * (ny_[k] + ny_[k + 1]) * cr;
    }
  }
  cent[2] = d3;
}

#define CG_EPSI 1.0e-3  /* CutCell.f90:11 */

/* DefineCell (CutCell.f90:329-365) */
static void define_cell(int I, int J, int K, orc_cell* c) {
  const int64_t N = (int64_t)I * J * K;
  for (int64_t q = 0; q < N; ++q) {
    if (c->vof[q] >= 1.0 - CG_EPSI) {
      c->type[q] = 0;
      c->vofl[q] = c->vofl[q] / c->vof[q];
      c->vof[q] = 1.0;
    }
    if (c->vof[q] <= CG_EPSI) {
      c->type[q] = 2;
      c->vof[q] = 0.0;
      c->vofl[q] = 0.0;
    }
    if (c->vof[q] >= CG_EPSI && c->vof[q] <= 1.0 - CG_EPSI) c->type[q] = 1;
  }
}

/* CellGeoCal (CutCell.f90:369-671).  `demote`: the MaxFace < 1e-2 rule of
 * :577-582, which applies where TCell%MoExCell is not allocated (the P grid:
 * Main.f90:377-399 allocates it for U, V, W only). */
static void cell_geo_cal(int I, int J, int K, orc_cell* c, int demote) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  const int64_t N = (int64_t)I * J * K, Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  const double tol = 1.0e-20;
  double* WE = dalloc(N); double* SE = dalloc(N); double* BE = dalloc(N);
  double* Ce[6];  /* E, W, N, S, T, B; 3 components each */
  for (int f = 0; f < 6; ++f) Ce[f] = dalloc(3 * N);
  for (int64_t q = 0; q < N; ++q) { WE[q] = 1.0; SE[q] = 1.0; BE[q] = 1.0; c->wllh[q] = 0.0; }
  for (int64_t q = 0; q < Nh; ++q) { c->ee[q] = 1.0; c->ne[q] = 1.0; c->te[q] = 1.0; }
  double* fce[3] = {c->fce, c->fce + Nh, c->fce + 2 * Nh};
  double* fcn[3] = {c->fcn, c->fcn + Nh, c->fcn + 2 * Nh};
  double* fct[3] = {c->fct, c->fct + Nh, c->fct + 2 * Nh};
#define CE(f, q, m) Ce[f][(int64_t)(m) * N + (q)]
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k), h = HX(g, i, j, k);
        const double dx = c->dx[q], dy = c->dy[q], dz = c->dz[q], ph = c->phi[q];
        const double nx = c->nx[q], ny = c->ny[q], nz = c->nz[q];
        double F[3];
        for (int m = 0; m < 3; ++m) c->cent[(int64_t)m * N + q] = 0.0;
        orc_face_geo(dx, dy, nx, ny, nz, dz / 2.0, ph, &c->te[h], F);          /* :414 */
        CE(4, q, 0) = F[0]; CE(4, q, 1) = F[1]; CE(4, q, 2) = F[2];
        orc_face_geo(dx, dy, nx, ny, nz, -dz / 2.0, ph, &BE[q], F);            /* :424 */
        CE(5, q, 0) = F[0]; CE(5, q, 1) = F[1]; CE(5, q, 2) = F[2];
        orc_face_geo(dy, dz, ny, nz, nx, dx / 2.0, ph, &c->ee[h], F);          /* :434 */
        CE(0, q, 0) = F[2]; CE(0, q, 1) = F[0]; CE(0, q, 2) = F[1];
        orc_face_geo(dy, dz, ny, nz, nx, -dx / 2.0, ph, &WE[q], F);            /* :444 */
        CE(1, q, 0) = F[2]; CE(1, q, 1) = F[0]; CE(1, q, 2) = F[1];
        orc_face_geo(dz, dx, nz, nx, ny, dy / 2.0, ph, &c->ne[h], F);          /* :454 */
        CE(2, q, 0) = F[1]; CE(2, q, 1) = F[2]; CE(2, q, 2) = F[0];
        orc_face_geo(dz, dx, nz, nx, ny, -dy / 2.0, ph, &SE[q], F);            /* :464 */
        CE(3, q, 0) = F[1]; CE(3, q, 1) = F[2]; CE(3, q, 2) = F[0];
        const double a = (c->ee[h] - WE[q]) * dy * dz;                         /* :477-482 */
        const double b = (c->ne[h] - SE[q]) * dx * dz;
        const double e = (c->te[h] - BE[q]) * dx * dy;
        c->wllh[q] = sqrt(a * a + b * b + e * e);
      }
  /* shared faces (:487-555) */
  for (int j = 1; j <= J; ++j)
    for (int k = 1; k <= K; ++k)
      for (int i = 2; i <= I; ++i) {
        const int64_t q = NX(g, i, j, k), qm = NX(g, i - 1, j, k), hm = HX(g, i - 1, j, k);
        const double w = WE[q], e = c->ee[hm];
        const double avg = (w * w + e * e) / (w + e + tol);
        fce[1][hm] = (CE(1, q, 1) * w + CE(0, qm, 1) * e) / (w + e + tol);
        fce[2][hm] = (CE(1, q, 2) * w + CE(0, qm, 2) * e) / (w + e + tol);
        fce[0][hm] = CE(0, qm, 0);
        c->ee[hm] = avg;
      }
  for (int i = 1; i <= I; ++i)
    for (int k = 1; k <= K; ++k)
      for (int j = 2; j <= J; ++j) {
        const int64_t q = NX(g, i, j, k), qm = NX(g, i, j - 1, k), hm = HX(g, i, j - 1, k);
        const double s = SE[q], n = c->ne[hm];
        const double avg = (s * s + n * n) / (s + n + tol);
        fcn[0][hm] = (CE(3, q, 0) * s + CE(2, qm, 0) * n) / (s + n + tol);
        fcn[2][hm] = (CE(3, q, 2) * s + CE(2, qm, 2) * n) / (s + n + tol);
        fcn[1][hm] = CE(2, qm, 1);
        c->ne[hm] = avg;
      }
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 2; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k), qm = NX(g, i, j, k - 1), hm = HX(g, i, j, k - 1);
        const double b = BE[q], t = c->te[hm];
        const double avg = (b * b + t * t) / (b + t + tol);
        fct[0][hm] = (CE(5, q, 0) * b + CE(4, qm, 0) * t) / (b + t + tol);
        /* :541-543 weighs the bottom face's first component here too */
        fct[1][hm] = (CE(5, q, 0) * b + CE(4, qm, 1) * t) / (b + t + tol);
        fct[2][hm] = CE(4, qm, 2);
        c->te[hm] = avg;
      }
#undef CE
  if (demote)                                                                  /* :557-569 */
    for (int i = 2; i <= I; ++i)
      for (int j = 2; j <= J; ++j)
        for (int k = 2; k <= K; ++k) {
          double mf = c->ee[HX(g, i - 1, j, k)];
          const double o[5] = {c->ee[HX(g, i, j, k)], c->ne[HX(g, i, j - 1, k)],
                               c->ne[HX(g, i, j, k)], c->te[HX(g, i, j, k - 1)],
                               c->te[HX(g, i, j, k)]};
          for (int m = 0; m < 5; ++m) mf = o[m] > mf ? o[m] : mf;
          if (mf < 1.0e-2) c->type[NX(g, i, j, k)] = 2;
        }
  for (int i = 1; i <= I; ++i)                                                 /* :571-599 */
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int t = c->type[NX(g, i, j, k)];
        if (t == 2) {
          c->ee[HX(g, i - 1, j, k)] = 0.0; c->ee[HX(g, i, j, k)] = 0.0;
          c->ne[HX(g, i, j - 1, k)] = 0.0; c->ne[HX(g, i, j, k)] = 0.0;
          c->te[HX(g, i, j, k - 1)] = 0.0; c->te[HX(g, i, j, k)] = 0.0;
        }
        if (t == 1) {
          if (i + 1 <= I && c->type[NX(g, i + 1, j, k)] == 2) c->ee[HX(g, i, j, k)] = 0.0;
          if (j + 1 <= J && c->type[NX(g, i, j + 1, k)] == 2) c->ne[HX(g, i, j, k)] = 0.0;
          if (k + 1 <= K && c->type[NX(g, i, j, k + 1)] == 2) c->te[HX(g, i, j, k)] = 0.0;
        }
      }
  /* boundary copies of the face centres, in the order of :601-640 */
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k <= K + 1; ++k)
      for (int j = 0; j <= J + 1; ++j) {
        fce[m][HX(g, 0, j, k)] = fce[m][HX(g, 1, j, k)];
        fcn[m][HX(g, 0, j, k)] = fcn[m][HX(g, 1, j, k)];
        fct[m][HX(g, 0, j, k)] = fct[m][HX(g, 1, j, k)];
      }
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k <= K + 1; ++k)
      for (int j = 0; j <= J + 1; ++j) fce[m][HX(g, I, j, k)] = fce[m][HX(g, I - 1, j, k)];
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k <= K + 1; ++k)
      for (int i = 0; i <= I + 1; ++i) {
        fce[m][HX(g, i, 0, k)] = fce[m][HX(g, i, 1, k)];
        fcn[m][HX(g, i, 0, k)] = fcn[m][HX(g, i, 1, k)];
        fct[m][HX(g, i, 0, k)] = fct[m][HX(g, i, 1, k)];
      }
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k <= K + 1; ++k)
      for (int i = 0; i <= I + 1; ++i) fcn[m][HX(g, i, J, k)] = fcn[m][HX(g, i, J - 1, k)];
  for (int m = 0; m < 3; ++m)
    for (int j = 0; j <= J + 1; ++j)
      for (int i = 0; i <= I + 1; ++i) {
        fce[m][HX(g, i, j, 0)] = fce[m][HX(g, i, j, 1)];
        fcn[m][HX(g, i, j, 0)] = fcn[m][HX(g, i, j, 1)];
        fct[m][HX(g, i, j, 0)] = fct[m][HX(g, i, j, 1)];
      }
  for (int m = 0; m < 3; ++m)
    for (int j = 0; j <= J + 1; ++j)
      for (int i = 0; i <= I + 1; ++i) fct[m][HX(g, i, j, K)] = fct[m][HX(g, i, j, K - 1)];
  for (int i = 2; i <= I; ++i)                                                 /* :642-666 */
    for (int j = 2; j <= J; ++j)
      for (int k = 2; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k);
        if (!(fabs(c->vof[q] - 1.0) < CG_EPSI)) continue;
        const double dx = c->dx[q], dy = c->dy[q], dz = c->dz[q];
        if (c->ee[HX(g, i - 1, j, k)] == 0.0 || c->ee[HX(g, i, j, k)] == 0.0) {
          c->wllh[q] = dy * dz; c->delh[q] = dx / 2.0;
        }
        if (c->ne[HX(g, i, j - 1, k)] == 0.0 || c->ne[HX(g, i, j, k)] == 0.0) {
          c->wllh[q] = dx * dz; c->delh[q] = dy / 2.0;
        }
        if (c->te[HX(g, i, j, k - 1)] == 0.0 || c->te[HX(g, i, j, k)] == 0.0) {
          c->wllh[q] = dx * dy; c->delh[q] = dz / 2.0;
        }
      }
  free(WE); free(SE); free(BE);
  for (int f = 0; f < 6; ++f) free(Ce[f]);
}

void orc_grid_preprocess(int I, int J, int K, orc_cell* P, orc_cell* U, orc_cell* V,
                         orc_cell* W) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  const int64_t N = (int64_t)I * J * K, Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  orc_cell* cs[4] = {P, U, V, W};
  for (int m = 0; m < 4; ++m) define_cell(I, J, K, cs[m]);                     /* :73-76 */
  for (int m = 0; m < 4; ++m) cell_geo_cal(I, J, K, cs[m], m == 0);            /* :80-83 */
#define F3(a, c_, h) ((a)[(int64_t)(c_) * Nh + (h)])
  /* face centres of U, V, W on stretched spacing (:89-159) */
  for (int j = 1; j <= J; ++j)
    for (int k = 1; k <= K; ++k) {
      for (int i = 1; i <= I - 1; ++i)
        if (U->type[NX(g, i, j, k)] != 1 && U->type[NX(g, i + 1, j, k)] != 1 &&
            fabs(U->dx[NX(g, i, j, k)] - U->dx[NX(g, i + 1, j, k)]) > 1.0e-10) {
          F3(U->fce, 0, HX(g, i, j, k)) = 0.5 * P->dx[NX(g, i + 1, j, k)];
          F3(U->fcn, 0, HX(g, i, j, k)) = 0.0;
          F3(U->fct, 0, HX(g, i, j, k)) = 0.0;
        }
      F3(U->fcn, 0, HX(g, I, j, k)) = 0.0;
      F3(U->fct, 0, HX(g, I, j, k)) = 0.0;
    }
  for (int i = 1; i <= I; ++i)
    for (int k = 1; k <= K; ++k) {
      for (int j = 1; j <= J - 1; ++j)
        if (V->type[NX(g, i, j, k)] != 1 && V->type[NX(g, i, j + 1, k)] != 1 &&
            fabs(V->dy[NX(g, i, j, k)] - V->dy[NX(g, i, j + 1, k)]) > 1.0e-10) {
          F3(V->fcn, 1, HX(g, i, j, k)) = 0.5 * P->dy[NX(g, i, j + 1, k)];
          F3(V->fce, 1, HX(g, i, j, k)) = 0.0;
          F3(V->fct, 1, HX(g, i, j, k)) = 0.0;
        }
      F3(V->fce, 1, HX(g, i, J, k)) = 0.0;
      F3(V->fct, 1, HX(g, i, J, k)) = 0.0;
    }
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j) {
      for (int k = 1; k <= K - 1; ++k)
        if (W->type[NX(g, i, j, k)] != 1 && W->type[NX(g, i, j, k + 1)] != 1 &&
            fabs(W->dz[NX(g, i, j, k)] - W->dz[NX(g, i, j, k + 1)]) > 1.0e-10) {
          F3(W->fct, 2, HX(g, i, j, k)) = 0.5 * P->dz[NX(g, i, j, k + 1)];
          F3(W->fce, 2, HX(g, i, j, k)) = 0.0;
          F3(W->fcn, 2, HX(g, i, j, k)) = 0.0;
        }
      F3(W->fce, 2, HX(g, i, j, K)) = 0.0;
      F3(W->fcn, 2, HX(g, i, j, K)) = 0.0;
    }
  /* U/V/W faces through the P-cell centre (:163-223) */
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k);
        const int64_t hu = HX(g, i - 1, j, k), hv = HX(g, i, j - 1, k), hw = HX(g, i, j, k - 1);
        for (int m = 0; m < 3; ++m) P->cent[(int64_t)m * N + q] = 0.0;
        const double dx = P->dx[q], dy = P->dy[q], dz = P->dz[q], vf = P->vof[q];
        if (vf < 1.0 && vf > 0.0) {
          double F[3];
          orc_face_geo(dy, dz, P->ny[q], P->nz[q], P->nx[q], 0.0, P->phi[q], &U->ee[hu], F);
          F3(U->fce, 0, hu) = F[2] + dx / 2.0;
          F3(U->fce, 1, hu) = F[0]; F3(U->fce, 2, hu) = F[1];
          orc_face_geo(dz, dx, P->nz[q], P->nx[q], P->ny[q], 0.0, P->phi[q], &V->ne[hv], F);
          F3(V->fcn, 1, hv) = F[2] + dy / 2.0;
          F3(V->fcn, 0, hv) = F[1]; F3(V->fcn, 2, hv) = F[0];
          orc_face_geo(dx, dy, P->nx[q], P->ny[q], P->nz[q], 0.0, P->phi[q], &W->te[hw], F);
          F3(W->fct, 2, hw) = F[2] + dz / 2.0;
          F3(W->fct, 0, hw) = F[0]; F3(W->fct, 1, hw) = F[1];
        } else if (vf > 0.5) {
          F3(U->fce, 0, hu) = dx / 2.0; F3(U->fce, 1, hu) = 0.0; F3(U->fce, 2, hu) = 0.0;
          U->ee[hu] = 1.0;
          F3(V->fcn, 1, hv) = dy / 2.0; F3(V->fcn, 0, hv) = 0.0; F3(V->fcn, 2, hv) = 0.0;
          V->ne[hv] = 1.0;
          F3(W->fct, 2, hw) = dz / 2.0; F3(W->fct, 0, hw) = 0.0; F3(W->fct, 1, hw) = 0.0;
          W->te[hw] = 1.0;
        } else {
          U->ee[hu] = 0.0; V->ne[hv] = 0.0; W->te[hw] = 0.0;
          for (int m = 0; m < 3; ++m) {
            F3(U->fce, m, hu) = 0.0; F3(V->fcn, m, hv) = 0.0; F3(W->fct, m, hw) = 0.0;
          }
        }
      }
  /* P faces through the U/V/W cell centres and their cell centres (:225-323) */
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k), h = HX(g, i, j, k);
        double F[3], A;
        if (U->vof[q] < 1.0 && U->vof[q] > 0.0) {
          orc_face_geo(U->dy[q], U->dz[q], U->ny[q], U->nz[q], U->nx[q], 0.0, U->phi[q], &A, F);
          F3(P->fce, 0, h) = F[2] + P->dx[q] / 2.0;
          F3(P->fce, 1, h) = F[0]; F3(P->fce, 2, h) = F[1];
          U->cent[q] = F[2]; U->cent[N + q] = F[0]; U->cent[2 * N + q] = F[1];
        } else if (U->vof[q] > 0.5) {
          P->ee[h] = 1.0;
          F3(P->fce, 0, h) = P->dx[q] / 2.0; F3(P->fce, 1, h) = 0.0; F3(P->fce, 2, h) = 0.0;
          for (int m = 0; m < 3; ++m) U->cent[(int64_t)m * N + q] = 0.0;
        } else {
          P->ee[h] = 0.0;
          for (int m = 0; m < 3; ++m) { F3(P->fce, m, h) = 0.0; U->cent[(int64_t)m * N + q] = 0.0; }
        }
        if (V->vof[q] < 1.0 && V->vof[q] > 0.0) {
          orc_face_geo(V->dz[q], V->dx[q], V->nz[q], V->nx[q], V->ny[q], 0.0, V->phi[q], &A, F);
          F3(P->fcn, 0, h) = F[1]; F3(P->fcn, 2, h) = F[0];
          F3(P->fcn, 1, h) = F[2] + P->dy[q] / 2.0;
          V->cent[q] = F[1]; V->cent[N + q] = F[2]; V->cent[2 * N + q] = F[0];
        } else if (V->vof[q] > 0.5) {
          P->ne[h] = 1.0;
          F3(P->fcn, 0, h) = 0.0; F3(P->fcn, 2, h) = 0.0; F3(P->fcn, 1, h) = P->dy[q] / 2.0;
          for (int m = 0; m < 3; ++m) V->cent[(int64_t)m * N + q] = 0.0;
        } else {
          P->ne[h] = 0.0;
          for (int m = 0; m < 3; ++m) { F3(P->fcn, m, h) = 0.0; V->cent[(int64_t)m * N + q] = 0.0; }
        }
        if (W->vof[q] < 1.0 && W->vof[q] > 0.0) {
          orc_face_geo(W->dx[q], W->dy[q], W->nx[q], W->ny[q], W->nz[q], 0.0, W->phi[q], &A, F);
          F3(P->fct, 0, h) = F[0]; F3(P->fct, 1, h) = F[1];
          F3(P->fct, 2, h) = F[2] + P->dz[q] / 2.0;
          W->cent[q] = F[0]; W->cent[N + q] = F[1]; W->cent[2 * N + q] = F[2];
        } else if (W->vof[q] > 0.5) {
          P->te[h] = 1.0;
          F3(P->fct, 0, h) = 0.0; F3(P->fct, 1, h) = 0.0; F3(P->fct, 2, h) = P->dz[q] / 2.0;
          for (int m = 0; m < 3; ++m) W->cent[(int64_t)m * N + q] = 0.0;
        } else {
          P->te[h] = 0.0;
          for (int m = 0; m < 3; ++m) { F3(P->fct, m, h) = 0.0; W->cent[(int64_t)m * N + q] = 0.0; }
        }
      }
#undef F3
}

/* DefineMomentumExchangeCell (CutCell.f90:673-751): a cut U (V, W) cell next
 * to a solid P cell (either P cell of its face, only the own one in the last
 * cell of a line) gets MoExCell = 1; other entries keep their values. */
void orc_define_moex(int I, int J, int K, const int32_t* ctp, const int32_t* ctu,
                     const int32_t* ctv, const int32_t* ctw, int32_t* mu, int32_t* mv,
                     int32_t* mw) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k);
        if (ctu[q] == 1) {
          if (i < I) {
            if (ctp[q] == 2 || ctp[NX(g, i + 1, j, k)] == 2) mu[q] = 1;
          } else if (ctp[q] == 2) {
            mu[q] = 1;
          }
        }
        if (ctv[q] == 1) {
          if (j < J) {
            if (ctp[q] == 2 || ctp[NX(g, i, j + 1, k)] == 2) mv[q] = 1;
          } else if (ctp[q] == 2) {
            mv[q] = 1;
          }
        }
        if (ctw[q] == 1) {
          if (k < K) {
            if (ctp[q] == 2 || ctp[NX(g, i, j, k + 1)] == 2) mw[q] = 1;
          } else if (ctp[q] == 2) {
            mw[q] = 1;
          }
        }
      }
}

/* NumberExternalCell (CutCell.f90:753-786): PosNu (Imax-iu, Jmax-iv,
 * Kmax-iw) numbers the non-solid cells from 0 in the loop order i, j, k (k
 * innermost), -1 for solid cells; returns ExtCell = count - 1. */
int32_t orc_number_external(int I, int J, int K, int iu, int iv, int iw, const int32_t* ct,
                            int32_t* posnu) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  const int64_t I2 = I - iu, J2 = J - iv;
  int32_t ctr = 0;
  for (int i = 1; i <= I - iu; ++i)
    for (int j = 1; j <= J - iv; ++j)
      for (int k = 1; k <= K - iw; ++k) {
        const int64_t p = (i - 1) + I2 * ((j - 1) + J2 * (int64_t)(k - 1));
        if (ct[NX(g, i, j, k)] != 2) posnu[p] = ctr++;
        else posnu[p] = -1;
      }
  return ctr - 1;
}

/* ------------------------------------------------------------------ */
/* NewCellFace (CutCell.f90:788-881) with CellLinking (:1269-1355) and */
/* East/North/TopFaceInterpolationInf (:983-1267).                      */
/* ------------------------------------------------------------------ */

/* The interpolation data of one face direction `ax` (0 E, 1 N, 2 T) of grid
 * T, with B the grid whose spacing crosses the face and `ib` the offset of
 * its index (the iu/iv/iw argument). */
static void face_interp(int I, int J, int K, int ax, int ib, const orc_cell* T,
                        const double* bd, const int32_t* ct, double* S, double* Eta,
                        double* Al, double* DAl) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  const int64_t N = (int64_t)I * J * K, Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  const double* cent = T->cent;
  const double* FC = ax == 0 ? T->fce : (ax == 1 ? T->fcn : T->fct);
  const double* area = ax == 0 ? T->ee : (ax == 1 ? T->ne : T->te);
  const double* td[3] = {T->dx, T->dy, T->dz};
  const double* tn[3] = {T->nx, T->ny, T->nz};
  const int L = ax == 0 ? I : (ax == 1 ? J : K);
  const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;  /* the two transverse axes */
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int idx[3] = {i, j, k};
        if (idx[ax] > L - 1) continue;
        int n3[3] = {i, j, k};
        n3[ax] += 1;
        const int64_t q = NX(g, i, j, k), qn = NX(g, n3[0], n3[1], n3[2]), h = HX(g, i, j, k);
        int b3[3] = {i, j, k};
        b3[ax] += ib;
        const double dB = bd[NX(g, b3[0], b3[1], b3[2])];
        if (ct[q] != 2 && ct[qn] != 2) {
          const double ca = cent[ax * N + q], can = cent[ax * N + qn];
          S[h] = can + dB - ca;
          const double s1 = cent[o1 * N + qn] - cent[o1 * N + q];
          const double s2 = cent[o2 * N + qn] - cent[o2 * N + q];
          Eta[h] = fabs((FC[ax * Nh + h] - ca) / S[h]);
          if (Eta[h] > 1.0) Eta[h] = 0.5;
          double f[3];
          for (int m = 0; m < 3; ++m) {
            const double a = cent[m * N + q], b = cent[m * N + qn];
            f[m] = m == ax ? (1.0 - Eta[h]) * a + Eta[h] * (b + dB) : (1.0 - Eta[h]) * a + Eta[h] * b;
          }
          /* East tests > (:1013-1014), North and Top >= (:1107, :1202) */
          const int tr = ax == 0 ? (fabs(s1) > 1.0e-4 * td[o1][q] || fabs(s2) > 1.0e-4 * td[o2][q])
                                 : (fabs(s1) >= 1.0e-4 * td[o1][q] || fabs(s2) >= 1.0e-4 * td[o2][q]);
          if (tr) {
            const double nf1 = 0.5 * (tn[o1][q] + tn[o1][qn]);
            const double nf2 = 0.5 * (tn[o2][q] + tn[o2][qn]);
            const double nx0 = T->nx[q], ny0 = T->ny[q], nz0 = T->nz[q], ph0 = T->phi[q];
            const double nx1 = T->nx[qn], ny1 = T->ny[qn], nz1 = T->nz[qn], ph1 = T->phi[qn];
            const double delh1 = f[0] * nx0 + f[1] * ny0 + f[2] * nz0 + ph0;
            double g1[3] = {f[0], f[1], f[2]};
            g1[ax] = f[ax] - dB;
            const double delh2 = g1[0] * nx1 + g1[1] * ny1 + g1[2] * nz1 + ph1;
            double delh = 0.5 * (delh1 + delh2);
            const double F1 = FC[h], F2 = FC[Nh + h], F3 = FC[2 * Nh + h];
            double delhec1, delhec2;
            if (ax == 0) {          /* :1032-1037 */
              delhec1 = F1 * nx0 + F2 * ny0 + F3 * nz0 + ph0;
              delhec2 = (F1 - dB) * nx1 + F2 * ny1 + F3 * nz1 + ph1;
            } else if (ax == 1) {   /* :1127-1132, BGrid%dy(i,j+1,k) whatever iv is */
              const double dB1 = bd[NX(g, i, j + 1, k)];
              delhec1 = F1 * nx0 + F2 * ny0 + nz0 * F3 + ph0;
              delhec2 = F1 * nx1 + (F2 - dB1) * ny1 + nz1 * F3 + ph1;
            } else {                /* :1222-1227 */
              delhec1 = F1 * nx0 + F2 * ny0 + nz0 * F3 + ph0;
              delhec2 = F1 * nx1 + F2 * ny1 + nz1 * (F3 - dB) + ph1;
            }
            const double delhec = 0.5 * (delhec1 + delhec2);
            if (delh < CG_EPSI && area[h] > 0.0) delh = delhec;
            Al[h] = fabs(delhec) / fabs(delh);
            DAl[h] = (s1 * nf1 + s2 * nf2) / (S[h] * fabs(delh));
          } else {
            DAl[h] = 0.0;
            Al[h] = 1.0;
          }
        } else {
          Al[h] = 1.0;
          DAl[h] = 0.0;
          S[h] = dB;
          Eta[h] = 0.5;
        }
      }
  /* the entry after each line, index Imax (Jmax, Kmax), copies the last one
   * (the loop variable after the loop, :1067-1070), then plane 0 copies
   * plane 1 over the whole halo range (:1073-1076) */
  double* arrs[4] = {S, Eta, Al, DAl};
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int idx[3] = {i, j, k};
        if (idx[ax] != L) continue;
        int p3[3] = {i, j, k};
        p3[ax] -= 1;
        for (int m = 0; m < 4; ++m) arrs[m][HX(g, i, j, k)] = arrs[m][HX(g, p3[0], p3[1], p3[2])];
      }
  for (int m = 0; m < 4; ++m)
    for (int k = 0; k <= K + 1; ++k)
      for (int j = 0; j <= J + 1; ++j)
        for (int i = 0; i <= I + 1; ++i) {
          const int idx[3] = {i, j, k};
          if (idx[ax] != 0) continue;
          int p3[3] = {i, j, k};
          p3[ax] = 1;
          arrs[m][HX(g, i, j, k)] = arrs[m][HX(g, p3[0], p3[1], p3[2])];
        }
}

/* CellLinking (CutCell.f90:1269-1355) of cell (i,j,k) of grid T: the cell
 * centre moves to its master neighbour's, shifted by that cell's width (less
 * 1e-8 of it), along the dominant normal component.  A master outside the
 * grid (index 0, which the reference reads outside its arrays) leaves the
 * cell unchanged. */
static void cell_linking(int I, int J, int K, int i, int j, int k, const orc_cell* T,
                         double* cent, int32_t* msce) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  const int64_t N = (int64_t)I * J * K, q = NX(g, i, j, k);
  const double tol = 1.0e-8;
  const double n[3] = {T->nx[q], T->ny[q], T->nz[q]};
  const double* d[3] = {T->dx, T->dy, T->dz};
  double nmax = fabs(n[0]);
  if (fabs(n[1]) > nmax) nmax = fabs(n[1]);
  if (fabs(n[2]) > nmax) nmax = fabs(n[2]);
  int ax = -1;
  for (int a = 0; a < 3 && ax < 0; ++a)
    if (fabs(fabs(n[a]) - nmax) < 1.0e-10) ax = a;
  if (ax < 0) return;
  const int idx[3] = {i, j, k};
  int m3[3] = {i, j, k};
  m3[ax] += n[ax] < 0.0 ? -1 : 1;
  if (m3[ax] < 1) return;
  const int64_t qm = NX(g, m3[0], m3[1], m3[2]);
  const double dm = d[ax][qm];
  double c[3];
  for (int a = 0; a < 3; ++a) c[a] = cent[a * N + qm];
  if (n[ax] < 0.0) c[ax] = c[ax] - dm + tol * dm;
  else c[ax] = c[ax] + dm - tol * dm;
  for (int a = 0; a < 3; ++a) {
    cent[a * N + q] = c[a];
    msce[a * N + q] = a == ax ? m3[a] : idx[a];
  }
}

void orc_new_cell_face(int I, int J, int K, const orc_cell* P, orc_cell* U, orc_cell* V,
                       orc_cell* W, orc_interp* iu, orc_interp* iv, orc_interp* iw) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  const int64_t N = (int64_t)I * J * K, Nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  orc_cell* C[3] = {U, V, W};
  orc_interp* X[3] = {iu, iv, iw};
  /* cell centres of U/V/W from the P face centres (:798-822) */
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t q = NX(g, i, j, k), h = HX(g, i, j, k);
        if (P->ee[h] > 0.0) {
          U->cent[q] = 0.0; U->cent[N + q] = P->fce[Nh + h]; U->cent[2 * N + q] = P->fce[2 * Nh + h];
        }
        if (P->ne[h] > 0.0) {
          V->cent[q] = P->fcn[h]; V->cent[N + q] = 0.0; V->cent[2 * N + q] = P->fcn[2 * Nh + h];
        }
        if (P->te[h] > 0.0) {
          W->cent[q] = P->fct[h]; W->cent[N + q] = P->fct[Nh + h]; W->cent[2 * N + q] = 0.0;
        }
      }
  /* CellLinking of the momentum-exchange cells, in the reference's order
   * (:824-842) */
  for (int i = 1; i <= I - 1; ++i)
    for (int j = 1; j <= J - 1; ++j)
      for (int k = 1; k <= K - 1; ++k) {
        const int64_t q = NX(g, i, j, k);
        for (int m = 0; m < 3; ++m)
          if (X[m]->moex[q] == 1) cell_linking(I, J, K, i, j, k, C[m], C[m]->cent, X[m]->msce);
      }
  /* distance of the cell centre to the solid surface (:844-858) */
  for (int m = 0; m < 3; ++m)
    for (int64_t q = 0; q < N; ++q) {
      const orc_cell* T = C[m];
      T->delh[q] = fabs(T->cent[q] * T->nx[q] + T->cent[N + q] * T->ny[q] +
                        T->cent[2 * N + q] * T->nz[q] + T->phi[q]);
    }
  /* interpolation data (:862-872): face ax of grid m crosses grid B, which
   * is P along the grid's own axis (offset 1) and U/V/W otherwise */
  for (int m = 0; m < 3; ++m)
    for (int ax = 0; ax < 3; ++ax) {
      const orc_cell* B = ax == m ? P : C[ax];
      const double* bd = ax == 0 ? B->dx : (ax == 1 ? B->dy : B->dz);
      face_interp(I, J, K, ax, ax == m ? 1 : 0, C[m], bd, C[m]->type, X[m]->s[ax],
                  X[m]->eta[ax], X[m]->al[ax], X[m]->dal[ax]);
    }
}

/* ------------------------------------------------------------------ */
/* ComputeUVWLiquidField (Clsvof.f90:174-298) with                      */
/* ComputeVelocityCellLiquidField (:300-326), DirectionAverageArray     */
/* (:328-353), Volume_Fraction_Calc (:1475-1550) and                    */
/* ConvertTheCoordinateSystem (:2458-2465).                             */
/* ------------------------------------------------------------------ */

/* ConvertTheCoordinateSystem: the level set measured from the cell corner */
static double orc_corner_dist(double dx, double dy, double dz, double nx, double ny, double nz,
                              double phi) {
  return phi + 0.5 * (fabs(nx) * dx + fabs(ny) * dy + fabs(nz) * dz);
}

/* Volume_Fraction_Calc (Griebel and Klitz): the volume of the box on the
 * corner side of the plane at distance s.  x**3.d0 is pow(x, 3.0) (a real
 * exponent); x**2.d0 is x*x. */
double orc_volume_fraction(double dx, double dy, double dz, double nxx, double nyy, double nzz,
                           double s) {
  const double nx1 = fabs(nxx), ny1 = fabs(nyy), nz1 = fabs(nzz);
  const double a = nx1 * dx, b = ny1 * dy, c = nz1 * dz;
  double dx1 = a, dz1 = a;
  if (b > dx1) dx1 = b;
  if (c > dx1) dx1 = c;
  if (b < dz1) dz1 = b;
  if (c < dz1) dz1 = c;
  const double dy1 = (a + b + c) - dx1 - dz1;
  const double sm = dx1 + dy1 + dz1;
  const double sc = s < sm - s ? s : sm - s;
  if (s <= 0.0) return 0.0;
  if (s >= sm) return 1.0 * dx * dy * dz;
  double fc;
  if (sc <= dz1) {
    fc = pow(sc, 3.0) / (6.0 * dx1 * dy1 * dz1);
  } else if (sc <= dy1) {
    fc = (sc * sc - dz1 * sc + dz1 * dz1 / 3.0) / (2.0 * dx1 * dy1);
  } else if (dx1 >= dy1 + dz1) {
    if (sc >= dy1 + dz1) {
      fc = (2.0 * sc - dy1 - dz1) / (2.0 * dx1);
    } else {
      fc = (pow(sc, 3.0) - pow(sc - dz1, 3.0) - pow(sc - dy1, 3.0));
      fc = fc / (6.0 * dx1 * dy1 * dz1);
    }
  } else {
    if (sc > dx1) fc = (pow(sc, 3.0) - pow(sc - dz1, 3.0) - pow(sc - dy1, 3.0) - pow(sc - dx1, 3.0));
    else fc = (pow(sc, 3.0) - pow(sc - dz1, 3.0) - pow(sc - dy1, 3.0));
    fc = fc / (6.0 * dx1 * dy1 * dz1);
  }
  if (s <= 0.5 * sm) return fc * dx * dy * dz;
  return (1.0 - fc) * dx * dy * dz;
}

/* the liquid volume of the part (d1, d2, d3) of a P cell whose level set,
 * moved to that part as the reference moves it, is phin */
static double liquid_part(double d1, double d2, double d3, double nx, double ny, double nz,
                          double phin) {
  const double dist = orc_corner_dist(d1, d2, d3, nx, ny, nz, phin);
  const double v = orc_volume_fraction(d1, d2, d3, nx, ny, nz, dist);
  return d1 * d2 * d3 - v;
}

void orc_uvw_liquid_field(int I, int J, int K, const orc_liquid* P, orc_liquid* U, orc_liquid* V,
                          orc_liquid* W) {
  const orc_grid G = {.I = I, .J = J, .K = K};
  const orc_grid* g = &G;
  orc_liquid* C[3] = {U, V, W};
  const double* pd[3] = {P->dx, P->dy, P->dz};
  /* DirectionAverageArray of phiL, nxL, nyL, nzL (:311-315, :328-353) */
  for (int m = 0; m < 3; ++m) {
    const double* in[4] = {P->phi, P->nx, P->ny, P->nz};
    double* out[4] = {C[m]->phi, C[m]->nx, C[m]->ny, C[m]->nz};
    const int iu = m == 0, iv = m == 1, iw = m == 2;
    for (int f = 0; f < 4; ++f) {
      for (int i = 1; i <= I - iu; ++i)
        for (int j = 1; j <= J - iv; ++j)
          for (int k = 1; k <= K - iw; ++k) {
            const int64_t q = NX(g, i, j, k), qn = NX(g, i + iu, j + iv, k + iw);
            const double lamda = pd[m][qn] / (pd[m][q] + pd[m][qn]);
            out[f][q] = lamda * in[f][q] + (1.0 - lamda) * in[f][qn];
          }
      for (int i = 1; i <= I; ++i)
        for (int j = 1; j <= J; ++j)
          for (int k = 1; k <= K; ++k) {
            const int idx[3] = {i, j, k};
            const int L = m == 0 ? I : (m == 1 ? J : K);
            if (idx[m] != L) continue;
            int p3[3] = {i, j, k};
            p3[m] -= 1;
            out[f][NX(g, i, j, k)] = out[f][NX(g, p3[0], p3[1], p3[2])];
          }
    }
  }
  /* U (:189-216) */
  for (int j = 1; j <= J; ++j)
    for (int k = 1; k <= K; ++k) {
      for (int i = 1; i <= I - 1; ++i) {
        const int64_t q = NX(g, i, j, k), qn = NX(g, i + 1, j, k);
        double dc = P->dx[q] / 2.0;
        double phin = P->phi[q] + dc * P->nx[q];
        const double vl = liquid_part(dc, P->dy[q], P->dz[q], P->nx[q], P->ny[q], P->nz[q], phin);
        dc = P->dx[qn] / 2.0;
        phin = P->phi[qn] - dc * P->nx[qn];
        const double vr = liquid_part(dc, P->dy[qn], P->dz[qn], P->nx[qn], P->ny[qn], P->nz[qn], phin);
        U->vof[q] = (vl + vr) / U->dx[q] / U->dy[q] / U->dz[q];
      }
      U->vof[NX(g, I, j, k)] = P->vof[NX(g, I, j, k)];
    }
  /* V (:218-257): the south part uses phiL itself with the full dy for the
   * corner distance; the north part moves phiL(i,j,k) by nyL(i,j+1,k) */
  for (int i = 1; i <= I; ++i)
    for (int k = 1; k <= K; ++k) {
      for (int j = 1; j <= J - 1; ++j) {
        const int64_t q = NX(g, i, j, k), qn = NX(g, i, j + 1, k);
        double dc = P->dy[q] / 2.0;
        double dist = orc_corner_dist(P->dx[q], P->dy[q], P->dz[q], P->nx[q], P->ny[q], P->nz[q],
                                      P->phi[q]);
        double v = orc_volume_fraction(P->dx[q], dc, P->dz[q], P->nx[q], P->ny[q], P->nz[q], dist);
        const double vl = P->dx[q] * dc * P->dz[q] - v;
        dc = P->dy[qn] / 2.0;
        const double phin = P->phi[q] + dc * P->ny[qn];
        dist = orc_corner_dist(P->dx[qn], dc, P->dz[qn], P->nx[qn], P->ny[qn], P->nz[qn], phin);
        v = orc_volume_fraction(P->dx[qn], dc, P->dz[qn], P->nx[qn], P->ny[qn], P->nz[qn], dist);
        const double vr = P->dx[qn] * dc * P->dz[qn] - v;
        V->vof[q] = (vl + vr) / V->dx[q] / V->dy[q] / V->dz[q];
      }
      V->vof[NX(g, i, J, k)] = P->vof[NX(g, i, J, k)];
    }
  /* W (:259-296): the top part moves phiL(k+1) by +dz/2 nzL(k+1) */
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j) {
      for (int k = 1; k <= K - 1; ++k) {
        const int64_t q = NX(g, i, j, k), qn = NX(g, i, j, k + 1);
        double dc = P->dz[q] / 2.0;
        double phin = P->phi[q] + dc * P->nz[q];
        const double vl = liquid_part(P->dx[q], P->dy[q], dc, P->nx[q], P->ny[q], P->nz[q], phin);
        dc = P->dz[qn] / 2.0;
        phin = P->phi[qn] + dc * P->nz[qn];
        const double vr = liquid_part(P->dx[qn], P->dy[qn], dc, P->nx[qn], P->ny[qn], P->nz[qn], phin);
        W->vof[q] = (vl + vr) / W->dx[q] / W->dy[q] / W->dz[q];
      }
      W->vof[NX(g, i, j, K)] = P->vof[NX(g, i, j, K)];
    }
}
