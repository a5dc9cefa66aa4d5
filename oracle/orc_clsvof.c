/* CPU restatement of the level-set pieces of the interface transport and of
 * the moving-solid geometry update -- test infrastructure only (the tests,
 * smoke() and the bench's CPU leg use it as the checker):
 *
 *   orc_normal_vector_irre  Normal_Vector_Irre (Clsvof.f90:2361-2456)
 *   orc_interpolate_lvs     TSolidBody%interpolateLvS (SolidBody.f90:646-742)
 *   orc_find_distance       Find_Distance (Clsvof.f90:1311-1408)
 *   orc_redist_candidate    one candidate of Redistance_Levelset
 *                           (Clsvof.f90:1738-1860, 1900-2062)
 *   orc_clsvof_scheme       Clsvof_Scheme (Clsvof.f90:355-798) with its
 *                           sweeps, reconstruction, boundary conditions and
 *                           redistancing (:801-1310, 1552-2221)
 *
 * Loop orders and the reference's operation order are kept, with
 * floating-point contraction off (the Makefile's -ffp-contract=off), so the
 * values are those of the reference's loops; x**2.d0 and x**2 are x*x.
 * Halo arrays are (0:I+1, 0:J+1, 0:K+1), cell arrays (1:I, 1:J, 1:K), both
 * i fastest. */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define HX(i, j, k) ((int64_t)(i) + (int64_t)(I + 2) * ((int64_t)(j) + (int64_t)(J + 2) * (int64_t)(k)))
#define CX(i, j, k) \
  ((int64_t)((i)-1) + (int64_t)I * ((int64_t)((j)-1) + (int64_t)J * (int64_t)((k)-1)))

static const double kEta = 0.075; /* Clsvof.f90:32 */

static double nv_q(double vx, double vy, double vz) {
  return fabs(1.0 - sqrt(vx * vx + vy * vy + vz * vz));
}

/* Normal_Vector_Irre: the (unnormalised) level-set gradient at an interior
 * cell (1 < i < I, ...), one-sided in a direction where the two-cell
 * difference on the other side departs from |grad| = 1 by eta or more
 * (Clsvof.f90:2361-2456).  phi is a halo array, x, y, z cell arrays. */
void orc_normal_vector_irre(int I, int J, int K, const double* x, const double* y,
                            const double* z, const double* phi, int i, int j, int k, double* nxx,
                            double* nyy, double* nzz) {
#define P(a, b, c) phi[HX(a, b, c)]
#define X(a, b, c) x[CX(a, b, c)]
#define Y(a, b, c) y[CX(a, b, c)]
#define Z(a, b, c) z[CX(a, b, c)]
  double vx, vy, vz, qm, qp;
  int dx, dy, dz;
  /* qi */
  if (i >= 3) vx = (P(i, j, k) - P(i - 2, j, k)) / (X(i, j, k) - X(i - 2, j, k));
  else vx = (P(i, j, k) - P(i - 1, j, k)) / (X(i, j, k) - X(i - 1, j, k));
  vy = (P(i - 1, j + 1, k) - P(i - 1, j - 1, k)) / (Y(i - 1, j + 1, k) - Y(i - 1, j - 1, k));
  vz = (P(i - 1, j, k + 1) - P(i - 1, j, k - 1)) / (Z(i - 1, j, k + 1) - Z(i - 1, j, k - 1));
  qm = nv_q(vx, vy, vz);
  if (i <= I - 2) vx = (P(i + 2, j, k) - P(i, j, k)) / (X(i + 2, j, k) - X(i, j, k));
  else vx = (P(i + 1, j, k) - P(i, j, k)) / (X(i + 1, j, k) - X(i, j, k));
  vy = (P(i + 1, j + 1, k) - P(i + 1, j - 1, k)) / (Y(i + 1, j + 1, k) - Y(i + 1, j - 1, k));
  vz = (P(i + 1, j, k + 1) - P(i + 1, j, k - 1)) / (Z(i + 1, j, k + 1) - Z(i + 1, j, k - 1));
  qp = nv_q(vx, vy, vz);
  dx = (qm < kEta && qp >= kEta) ? -1 : ((qm >= kEta && qp < kEta) ? 1 : 0);
  /* qj */
  if (j >= 3) vy = (P(i, j, k) - P(i, j - 2, k)) / (Y(i, j, k) - Y(i, j - 2, k));
  else vy = (P(i, j, k) - P(i, j - 1, k)) / (Y(i, j, k) - Y(i, j - 1, k));
  vx = (P(i + 1, j - 1, k) - P(i - 1, j - 1, k)) / (X(i + 1, j - 1, k) - X(i - 1, j - 1, k));
  vz = (P(i, j - 1, k + 1) - P(i, j - 1, k - 1)) / (Z(i, j - 1, k + 1) - Z(i, j - 1, k - 1));
  qm = nv_q(vx, vy, vz);
  vx = (P(i + 1, j + 1, k) - P(i - 1, j + 1, k)) / (X(i + 1, j + 1, k) - X(i - 1, j + 1, k));
  if (j <= J - 2) vy = (P(i, j + 2, k) - P(i, j, k)) / (Y(i, j + 2, k) - Y(i, j, k));
  else vy = (P(i, j + 1, k) - P(i, j, k)) / (Y(i, j + 1, k) - Y(i, j, k));
  vz = (P(i, j + 1, k + 1) - P(i, j + 1, k - 1)) / (Z(i, j + 1, k + 1) - Z(i, j + 1, k - 1));
  qp = nv_q(vx, vy, vz);
  dy = (qm < kEta && qp >= kEta) ? -1 : ((qm >= kEta && qp < kEta) ? 1 : 0);
  /* qk */
  vx = (P(i + 1, j, k - 1) - P(i - 1, j, k - 1)) / (X(i + 1, j, k - 1) - X(i - 1, j, k - 1));
  vy = (P(i, j + 1, k - 1) - P(i, j - 1, k - 1)) / (Y(i, j + 1, k - 1) - Y(i, j - 1, k - 1));
  if (k >= 3) vz = (P(i, j, k) - P(i, j, k - 2)) / (Z(i, j, k) - Z(i, j, k - 2));
  else vz = (P(i, j, k) - P(i, j, k - 1)) / (Z(i, j, k) - Z(i, j, k - 1));
  qm = nv_q(vx, vy, vz);
  vx = (P(i + 1, j, k + 1) - P(i - 1, j, k + 1)) / (X(i + 1, j, k + 1) - X(i - 1, j, k + 1));
  vy = (P(i, j + 1, k + 1) - P(i, j - 1, k + 1)) / (Y(i, j + 1, k + 1) - Y(i, j - 1, k + 1));
  if (k <= K - 2) vz = (P(i, j, k + 2) - P(i, j, k)) / (Z(i, j, k + 2) - Z(i, j, k));
  else vz = (P(i, j, k + 1) - P(i, j, k)) / (Z(i, j, k + 1) - Z(i, j, k));
  qp = nv_q(vx, vy, vz);
  dz = (qm < kEta && qp >= kEta) ? -1 : ((qm >= kEta && qp < kEta) ? 1 : 0);
  if (dx == -1) *nxx = (P(i, j, k) - P(i - 1, j, k)) / (X(i, j, k) - X(i - 1, j, k));
  else if (dx == 1) *nxx = (P(i + 1, j, k) - P(i, j, k)) / (X(i + 1, j, k) - X(i, j, k));
  else *nxx = (P(i + 1, j, k) - P(i - 1, j, k)) / (X(i + 1, j, k) - X(i - 1, j, k));
  if (dy == -1) *nyy = (P(i, j, k) - P(i, j - 1, k)) / (Y(i, j, k) - Y(i, j - 1, k));
  else if (dy == 1) *nyy = (P(i, j + 1, k) - P(i, j, k)) / (Y(i, j + 1, k) - Y(i, j, k));
  else *nyy = (P(i, j + 1, k) - P(i, j - 1, k)) / (Y(i, j + 1, k) - Y(i, j - 1, k));
  if (dz == -1) *nzz = (P(i, j, k) - P(i, j, k - 1)) / (Z(i, j, k) - Z(i, j, k - 1));
  else if (dz == 1) *nzz = (P(i, j, k + 1) - P(i, j, k)) / (Z(i, j, k + 1) - Z(i, j, k));
  else *nzz = (P(i, j, k + 1) - P(i, j, k - 1)) / (Z(i, j, k + 1) - Z(i, j, k - 1));
#undef P
#undef X
#undef Y
#undef Z
}

/* interpolateLvS (SolidBody.f90:646-742) onto one grid.  phi and vof are
 * reset over the whole halo array (1e4 and 1); every cell centre is taken
 * into the body's frame (pointL = R^T (x - CM), transferMeshGtoL
 * :618-627) and, where its lower local node (CEILING of the offset over the
 * spacing) lies in [1, imax-1] x ..., phi is the trilinear interpolation of
 * the body's phi.  Then every cell with |phi| <= 4 max(dx, dy, dz) gets the
 * normal (Normal_Vector_Irre inside, one-sided clamped differences on the
 * outer layer), normalised ((0, 0, 1) below 1e-14), and vof from
 * Volume_Fraction_Calc of the corner distance.  Elsewhere nx, ny, nz keep
 * their values and vof stays 1, inside the body too (:660-661). */
void orc_interpolate_lvs(const orc_solid_body* B, int I, int J, int K, const double* gx,
                         const double* gy, const double* gz, const double* gdx, const double* gdy,
                         const double* gdz, double* phi, double* vof, double* nx, double* ny,
                         double* nz) {
  const int64_t nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  for (int64_t q = 0; q < nh; ++q) {
    phi[q] = 1.0e4;
    vof[q] = 1.0;
  }
  const double* R = B->rot; /* rotMatLtoG(a, b) = R[(a-1) + 3 (b-1)] */
  const int bI = B->imax, bJ = B->jmax;
#define BP(a, b, c) B->phi[(int64_t)((a)-1) + (int64_t)bI * ((int64_t)((b)-1) + (int64_t)bJ * ((c)-1))]
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t c = CX(i, j, k);
        const double d[3] = {gx[c] - B->cm[0], gy[c] - B->cm[1], gz[c] - B->cm[2]};
        double pl[3];
        for (int a = 0; a < 3; ++a) {
          double s = 0.0;
          for (int b = 0; b < 3; ++b) s = s + R[b + 3 * a] * d[b];
          pl[a] = s;
        }
        const int iL = (int)ceil((pl[0] - B->x[0]) / B->dx);
        const int jL = (int)ceil((pl[1] - B->y[0]) / B->dy);
        const int kL = (int)ceil((pl[2] - B->z[0]) / B->dz);
        if (iL < 1 || iL > B->imax - 1 || jL < 1 || jL > B->jmax - 1 || kL < 1 ||
            kL > B->kmax - 1)
          continue;
        double v = 0.0;
        for (int ii = 0; ii <= 1; ++ii)
          for (int jj = 0; jj <= 1; ++jj)
            for (int kk = 0; kk <= 1; ++kk) {
              double coeff = (B->dx - fabs(pl[0] - B->x[iL + ii - 1])) / B->dx;
              coeff = coeff * (B->dy - fabs(pl[1] - B->y[jL + jj - 1]));
              coeff = coeff / B->dy;
              coeff = coeff * (B->dz - fabs(pl[2] - B->z[kL + kk - 1]));
              coeff = coeff / B->dz;
              v = v + coeff * BP(iL + ii, jL + jj, kL + kk);
            }
        phi[HX(i, j, k)] = v;
      }
#undef BP
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t h = HX(i, j, k), c = CX(i, j, k);
        const double ddx = gdx[c], ddy = gdy[c], ddz = gdz[c];
        double dm = ddx;
        if (ddy > dm) dm = ddy;
        if (ddz > dm) dm = ddz;
        if (fabs(phi[h]) > 4.0 * dm) continue;
        double a, b, e;
        if (i > 1 && i < I && j > 1 && j < J && k > 1 && k < K) {
          orc_normal_vector_irre(I, J, K, gx, gy, gz, phi, i, j, k, &a, &b, &e);
        } else {
          const int ip = i + 1 < I ? i + 1 : I, im = i - 1 > 1 ? i - 1 : 1;
          const int jp = j + 1 < J ? j + 1 : J, jm = j - 1 > 1 ? j - 1 : 1;
          const int kp = k + 1 < K ? k + 1 : K, km = k - 1 > 1 ? k - 1 : 1;
          a = (phi[HX(ip, j, k)] - phi[HX(im, j, k)]) / (gx[CX(ip, j, k)] - gx[CX(im, j, k)]);
          b = (phi[HX(i, jp, k)] - phi[HX(i, jm, k)]) / (gy[CX(i, jp, k)] - gy[CX(i, jm, k)]);
          e = (phi[HX(i, j, kp)] - phi[HX(i, j, km)]) / (gz[CX(i, j, kp)] - gz[CX(i, j, km)]);
        }
        const double t = sqrt(a * a + b * b + e * e);
        if (t < 1.0e-14) {
          a = 0.0;
          b = 0.0;
          e = 1.0;
        } else {
          a = a / t;
          b = b / t;
          e = e / t;
        }
        nx[h] = a;
        ny[h] = b;
        nz[h] = e;
        const double s = phi[h] + 0.5 * (fabs(a) * ddx + fabs(b) * ddy + fabs(e) * ddz);
        const double vol = orc_volume_fraction(ddx, ddy, ddz, a, b, e, s);
        vof[h] = vol / (ddx * ddy * ddz);
      }
}

/* ------------------------------------------------------------------------
 * Clsvof_Scheme (Clsvof.f90:355-798): nv = 6 sub-steps of directionally
 * split CLSVOF transport of the liquid (vofL, phiL), each three sweeps in a
 * rotating order, the level-set and vof boundary conditions, the interface
 * reconstruction and the geometric redistancing of the level set.
 * Restated loop by loop; vofeps = 1e-14, tolpar = 1e-14, band_width = 4,
 * nv = 6 (Clsvof.f90:31-35).  Newton_Raphson has no iteration limit in the
 * reference; here it stops after 200 iterations (never reached by a
 * converging case).
 * ------------------------------------------------------------------------ */

#define VOFEPS 1.0e-14
#define TOLPAR 1.0e-14
#define BANDW 4
#define NV 6

/* Isinterface (Clsvof.f90:801-815) */
static int isinterface(int I, int J, const double* v, int i, int j, int k) {
  const double c = v[HX(i, j, k)], lim = 1.0 - 2.1 * VOFEPS;
  int f = 0;
  if (c > VOFEPS && c < 1.0 - VOFEPS) f = 1;
  if (fabs(v[HX(i - 1, j, k)] - c) >= lim) f = 1;
  if (fabs(v[HX(i + 1, j, k)] - c) >= lim) f = 1;
  if (fabs(v[HX(i, j - 1, k)] - c) >= lim) f = 1;
  if (fabs(v[HX(i, j + 1, k)] - c) >= lim) f = 1;
  if (fabs(v[HX(i, j, k - 1)] - c) >= lim) f = 1;
  if (fabs(v[HX(i, j, k + 1)] - c) >= lim) f = 1;
  return f;
}

/* Newton_Raphson (:1462-1473): x**3.d0 is pow, x**2.d0 is x*x */
static double newton(double a3, double a2, double a1, double a0, double sc) {
  double delvar = 1.0e6;
  for (int it = 0; it < 200 && fabs(delvar) > 1.0e-6 * fabs(sc); ++it) {
    delvar = (a3 * pow(sc, 3.0) + a2 * (sc * sc) + a1 * sc + a0) /
             (3.0 * a3 * (sc * sc) + 2.0 * a2 * sc + a1);
    sc = sc - delvar;
  }
  return sc;
}

/* Cubic_Equation_Solve (:1380-1398); case2 selects the first form */
static double cubic(double fc, double dx1, double dy1, double dz1, double sc, int case2) {
  if (case2) {
    const double a3 = 1.0, a2 = -3.0 * (dy1 + dz1), a1 = 3.0 * (dy1 * dy1 + dz1 * dz1);
    const double a0 = 6.0 * fc * dx1 * dy1 * dz1 - (pow(dy1, 3.0) + pow(dz1, 3.0));
    return newton(a3, a2, a1, a0, sc);
  }
  const double a3 = 2.0, a2 = -3.0 * (dx1 + dy1 + dz1);
  const double a1 = 3.0 * (dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
  const double a0 = 6.0 * fc * dx1 * dy1 * dz1 - (pow(dx1, 3.0) + pow(dy1, 3.0) + pow(dz1, 3.0));
  return newton(a3, a2, a1, a0, sc);
}

/* Find_Distance (:1311-1378) with Final_Distance (:1400-1408): the distance
 * s from the cell corner of the plane with normal n that leaves the volume
 * fraction f on the corner side */
double orc_find_distance(double dx, double dy, double dz, double nxx, double nyy, double nzz,
                         double f) {
  const double nx1 = fabs(nxx), ny1 = fabs(nyy), nz1 = fabs(nzz);
  const double a = nx1 * dx, b = ny1 * dy, c = nz1 * dz;
  double dx1 = a, dz1 = a;
  if (b > dx1) dx1 = b;
  if (c > dx1) dx1 = c;
  if (b < dz1) dz1 = b;
  if (c < dz1) dz1 = c;
  if (fabs(dz1) < TOLPAR) dz1 = 0.0;
  double dy1 = (a + b + c) - dx1 - dz1;
  if (fabs(dy1) < TOLPAR) dy1 = 0.0;
  const double fc = f < 1.0 - f ? f : 1.0 - f;
  const double sm = dx1 + dy1 + dz1;
  double sc = 6.0 * fc * dx1 * dy1 * dz1;
  sc = copysign(1.0, sc) * pow(fabs(sc), 1.0 / 3.0);
  if (!(sc < dz1)) {
    sc = 0.5 * dz1 + sqrt(2.0 * fc * dx1 * dy1 - dz1 * dz1 / 12.0);
    if (!(sc < dy1)) {
      if (dx1 >= dy1 + dz1) {
        sc = fc * dx1 + (dy1 + dz1) / 2.0;
        if (!(sc >= dy1 + dz1)) sc = cubic(fc, dx1, dy1, dz1, sc, 1);
      } else {
        sc = fc * dx1 + (dy1 + dz1) / 2.0;
        if (sc >= dx1) {
          sc = cubic(fc, dx1, dy1, dz1, sc, 0);
          if (sc < dx1) sc = cubic(fc, dx1, dy1, dz1, sc, 1);
        } else {
          sc = cubic(fc, dx1, dy1, dz1, sc, 1);
        }
      }
    }
  }
  return f <= 0.5 ? sc : sm - sc;
}

/* East_Flux .. Bottom_Flux (:1552-1687): the liquid volume that crosses a
 * face in one sub-step, over the face area; `ax` is the sweep axis and
 * `pos` selects the East/North/Top form (1) or the West/South/Bottom form
 * (0) */
static double face_flux(int ax, int pos, double nxx, double nyy, double nzz, double diss,
                        double volf, double dx, double dy, double dz, double udt) {
  if (udt == 0.0) return 0.0;
  const double nn = ax == 0 ? nxx : (ax == 1 ? nyy : nzz);
  const int near = pos ? (nn <= 0.0) : (nn >= 0.0);
  double vol, flux;
  if (ax == 0) {
    if (near) {
      vol = orc_volume_fraction(udt, dy, dz, nxx, nyy, nzz, diss);
      flux = vol / (udt * dy * dz);
    } else {
      vol = orc_volume_fraction(dx - udt, dy, dz, nxx, nyy, nzz, diss);
      flux = (volf * dx * dy * dz - vol) / (udt * dy * dz);
    }
  } else if (ax == 1) {
    if (near) {
      vol = orc_volume_fraction(dx, udt, dz, nxx, nyy, nzz, diss);
      flux = vol / (dx * udt * dz);
    } else {
      vol = orc_volume_fraction(dx, dy - udt, dz, nxx, nyy, nzz, diss);
      flux = (volf * dx * dy * dz - vol) / (dx * udt * dz);
    }
  } else {
    if (near) {
      vol = orc_volume_fraction(dx, dy, udt, nxx, nyy, nzz, diss);
      flux = vol / (dx * dy * udt);
    } else {
      vol = orc_volume_fraction(dx, dy, dz - udt, nxx, nyy, nzz, diss);
      flux = (volf * dx * dy * dz - vol) / (dx * dy * udt);
    }
  }
  return flux * udt;
}

typedef struct clsvof_work {
  int I, J, K;
  const orc_clsvof_in* in;
  double *nxx, *nyy, *nzz, *diss; /* halo */
} clsvof_work;

/* Interface_Reconstruct (:1216-1309) */
static void reconstruct(clsvof_work* W, const double* vfl, const double* phi) {
  const int I = W->I, J = W->J, K = W->K;
  const orc_clsvof_in* in = W->in;
  const int64_t nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  for (int64_t q = 0; q < nh; ++q) W->nxx[q] = W->nyy[q] = W->nzz[q] = 0.0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t h = HX(i, j, k), c = CX(i, j, k);
        double a = 0.0, b = 0.0, e = 0.0, d = 0.0;
        if (isinterface(I, J, vfl, i, j, k)) {
          if (i > 1 && i < I && j > 1 && j < J && k > 1 && k < K) {
            orc_normal_vector_irre(I, J, K, in->x, in->y, in->z, phi, i, j, k, &a, &b, &e);
          } else {
            const int ip = i + 1 < I ? i + 1 : I, im = i - 1 > 1 ? i - 1 : 1;
            const int jp = j + 1 < J ? j + 1 : J, jm = j - 1 > 1 ? j - 1 : 1;
            const int kp = k + 1 < K ? k + 1 : K, km = k - 1 > 1 ? k - 1 : 1;
            a = (phi[HX(ip, j, k)] - phi[HX(im, j, k)]) / (in->x[CX(ip, j, k)] - in->x[CX(im, j, k)]);
            b = (phi[HX(i, jp, k)] - phi[HX(i, jm, k)]) / (in->y[CX(i, jp, k)] - in->y[CX(i, jm, k)]);
            e = (phi[HX(i, j, kp)] - phi[HX(i, j, km)]) / (in->z[CX(i, j, kp)] - in->z[CX(i, j, km)]);
          }
          const double t = sqrt(a * a + b * b + e * e);
          if (t < 1.0e-14) {
            a = 0.0;
            b = 0.0;
            e = 1.0;
          } else {
            a = a / t;
            b = b / t;
            e = e / t;
          }
          d = orc_find_distance(in->dx[c], in->dy[c], in->dz[c], a, b, e, vfl[h]);
        }
        W->nxx[h] = a;
        W->nyy[h] = b;
        W->nzz[h] = e;
        W->diss[h] = d;
      }
}

/* one sweep (X_Sweep :817-948, Y_Sweep :950-1083, Z_Sweep :1085-1214) along
 * axis ax: temvf/temls (copies of vfl/phi) receive the fluxes */
static void sweep(clsvof_work* W, int ax, const double* vfl, const double* phi, double* temvf,
                  double* temls, const double* bcl, double dtv) {
  const int I = W->I, J = W->J, K = W->K;
  const orc_clsvof_in* in = W->in;
  reconstruct(W, vfl, phi);
  const int n = ax == 0 ? I : (ax == 1 ? J : K);
  const int64_t st = ax == 0 ? 1 : (ax == 1 ? (int64_t)(I + 2) : (int64_t)(I + 2) * (J + 2));
  const int64_t cs = ax == 0 ? 1 : (ax == 1 ? (int64_t)I : (int64_t)I * J);
  const double* vel = ax == 0 ? in->u : (ax == 1 ? in->v : in->w);
  const double* dd = ax == 0 ? in->dx : (ax == 1 ? in->dy : in->dz);
  const double* xx = ax == 0 ? in->x : (ax == 1 ? in->y : in->z);
  const double* bcv = ax == 0 ? in->bcu_w : (ax == 1 ? in->bcv_s : in->bcw_b);
  /* lines: the two other indices */
  const int na = ax == 0 ? J : I, nb = ax == 2 ? J : K;
#pragma omp parallel for collapse(2) schedule(static)
  for (int b2 = 1; b2 <= nb; ++b2)
    for (int a2 = 1; a2 <= na; ++a2) {
      int i0, j0, k0;
      if (ax == 0) { i0 = 0; j0 = a2; k0 = b2; }
      else if (ax == 1) { i0 = a2; j0 = 0; k0 = b2; }
      else { i0 = a2; j0 = b2; k0 = 0; }
      const int64_t h0 = HX(i0, j0, k0);  /* index of position 0 of the line */
      const int64_t c0 = CX(ax == 0 ? 1 : i0, ax == 1 ? 1 : j0, ax == 2 ? 1 : k0) - cs;
      /* face value index of the BC array (VarW (J,K), VarS (I,K), VarB (I,J)) */
      const int64_t fb = ax == 0 ? (int64_t)(j0 - 1) + (int64_t)J * (k0 - 1)
                                 : (ax == 1 ? (int64_t)(i0 - 1) + (int64_t)I * (k0 - 1)
                                            : (int64_t)(i0 - 1) + (int64_t)I * (j0 - 1));
#define V(p) vfl[h0 + (int64_t)(p) * st]
#define PH(p) phi[h0 + (int64_t)(p) * st]
#define U(p) vel[h0 + (int64_t)(p) * st]
#define D(p) dd[c0 + (int64_t)(p) * cs]
#define XC(p) xx[c0 + (int64_t)(p) * cs]
#define TV(p) temvf[h0 + (int64_t)(p) * st]
#define TL(p) temls[h0 + (int64_t)(p) * st]
      /* volume of fluid */
      for (int p = 1; p <= n - 1; ++p) {
        const double ue = U(p);
        const int pos = ax == 2 ? (ue >= 0.0) : (ue > 0.0);
        double flux;
        const int q = pos ? p : p + 1;
        const double vq = V(q);
        if (vq >= 1.0 - VOFEPS || vq <= VOFEPS) {
          flux = vq * ue * dtv;
        } else {
          const int64_t hq = h0 + (int64_t)q * st, cq = c0 + (int64_t)q * cs;
          const double udt = pos ? ue * dtv : -ue * dtv;
          flux = face_flux(ax, pos, W->nxx[hq], W->nyy[hq], W->nzz[hq], W->diss[hq], vq,
                           in->dx[cq], in->dy[cq], in->dz[cq], udt);
          if (!pos) flux = -flux;
        }
        TV(p) = TV(p) - flux / D(p);
        if (p < n) TV(p + 1) = TV(p + 1) + flux / D(p + 1);
      }
      {
        double flux = V(1) * U(0) * dtv;
        TV(1) = TV(1) + flux / D(1);
        flux = V(n) * U(n) * dtv;
        TV(n) = TV(n) - flux / D(n);
      }
      /* level set */
      for (int p = 2; p <= n - 1; ++p) {
        const double ue = U(p);
        double ls;
        if (ue >= 0.0) {
          ls = PH(p) + D(p) / 2.0 * (1.0 - ue * dtv / D(p)) * (PH(p + 1) - PH(p - 1)) / (2.0 * D(p));
        } else if (p <= n - 2) {
          ls = PH(p + 1) - D(p) / 2.0 * (1.0 + ue * dtv / D(p)) * (PH(p + 2) - PH(p)) / (2.0 * D(p));
        } else {
          ls = PH(p + 1) - D(p) / 2.0 * (1.0 + ue * dtv / D(p)) * (PH(p + 1) - PH(p)) / D(p);
        }
        const double flux = ue * ls * dtv;
        if (p > 1) TL(p) = TL(p) - flux / D(p);
        if (p < n) TL(p + 1) = TL(p + 1) + flux / D(p + 1);
      }
      {
        /* face 1, reduced to first order (the forms differ per axis) */
        const double u1 = U(1);
        double ls;
        if (u1 >= 0.0) {
          if (ax == 0) ls = PH(1) + D(1) / 2.0 * (1.0 - u1 * dtv / D(1)) * (PH(2) - PH(1)) / D(1);
          else ls = PH(1) + D(1) / 2.0 * (1.0 - u1 * dtv / D(1)) * (PH(2) - PH(1)) / (XC(2) - XC(1));
        } else {
          if (ax == 0)
            ls = PH(2) - D(2) / 2.0 * (1.0 + u1 * dtv / D(2)) * (PH(3) - PH(1)) / 2.0 * D(2);
          else if (ax == 1)
            ls = PH(2) - D(2) / 2.0 * (1.0 + u1 * dtv / D(2)) * (PH(3) - PH(1)) / (XC(3) - XC(1));
          else
            ls = PH(2) + D(2) / 2.0 * (1.0 + u1 * dtv / D(2)) * (PH(3) - PH(1)) / (XC(3) - XC(1));
        }
        double flux = ls * u1 * dtv;
        TL(1) = TL(1) - flux / D(1);
        TL(2) = TL(2) + flux / D(2);
        /* inflow of the boundary value */
        flux = bcv[fb] * bcl[fb] * dtv;
        TL(1) = TL(1) + flux / D(1);
        /* face n */
        const double un = U(n);
        if (un > 0.0) {
          ls = PH(n) + D(n) / 2.0 * (1.0 - un * dtv / D(n)) * (PH(n) - PH(n - 1)) / (XC(n) - XC(n - 1));
        } else {
          /* the z sweep divides by dy(kmax-1) here (:1204-1208) */
          const double dq = ax == 2 ? in->dy[c0 + (int64_t)(n - 1) * cs] : D(n - 1);
          ls = PH(n) - D(n - 1) / 2.0 * (1.0 + un * dtv / dq) * (PH(n) - PH(n - 1)) / (XC(n) - XC(n - 1));
        }
        flux = ls * un * dtv;
        TL(n) = TL(n) - flux / D(n);
      }
#undef V
#undef PH
#undef U
#undef D
#undef XC
#undef TV
#undef TL
    }
}

static double clip(double v) {
  if (v <= VOFEPS) v = 0.0;
  if (v >= 1.0 - VOFEPS) v = 1.0;
  return v;
}

/* TypicalBC / LvSBoundaryCondition (BoundaryFunction.f90:760-793) of one
 * face: out(a, b) for the face's (a, b) cells */
static void typical_bc(double* out, int64_t n, const double* var, const double* dxyz, int sgn,
                       int flag, double cst) {
  for (int64_t q = 0; q < n; ++q) out[q] = flag == 0 ? cst : var[q] - cst * (sgn * dxyz[q]) / 2.0;
}

/* BoundaryConditionLvsVof (:2087-2221): BCLvs and BCVof face values, phiL
 * ghosts from the solid normals of the boundary cells, vofL ghosts copied */
static void bc_lvs_vof(int I, int J, int K, const orc_clsvof_in* in, orc_clsvof_state* st) {
  double* phi = st->phi;
  double* vof = st->vof;
  for (int f = 0; f < 6; ++f) {
    const int ax = f / 2, hi = f % 2;
    const int na = ax == 0 ? J : I, nb = ax == 2 ? J : K;
    const int64_t n = (int64_t)na * nb;
    double *vl = malloc(sizeof(double) * n), *vv = malloc(sizeof(double) * n),
           *dd = malloc(sizeof(double) * n), *nl = malloc(sizeof(double) * n);
    for (int b = 1; b <= nb; ++b)
      for (int a = 1; a <= na; ++a) {
        int i, j, k;
        if (ax == 0) { i = hi ? I : 1; j = a; k = b; }
        else if (ax == 1) { i = a; j = hi ? J : 1; k = b; }
        else { i = a; j = b; k = hi ? K : 1; }
        const int64_t q = (int64_t)(a - 1) + (int64_t)na * (b - 1);
        vl[q] = phi[HX(i, j, k)];
        vv[q] = vof[HX(i, j, k)];
        dd[q] = ax == 0 ? in->dx[CX(i, j, k)] : (ax == 1 ? in->dy[CX(i, j, k)] : in->dz[CX(i, j, k)]);
        nl[q] = st->nx[HX(i, j, k)];
      }
    const int sgn = hi ? -1 : 1;
    if (st->lvs_var[f]) {
      typical_bc(st->lvs_var[f], n, vl, dd, sgn, in->lvs_flag[f], in->lvs_const[f]);
      if (ax == 0) /* West and East pass nxL: LvSBoundaryCondition */
        for (int64_t q = 0; q < n; ++q) st->lvs_var[f][q] = vl[q] - nl[q] * (sgn * dd[q]) / 2.0;
    }
    if (st->vof_var[f]) typical_bc(st->vof_var[f], n, vv, dd, sgn, in->vof_flag[f], in->vof_const[f]);
    free(vl);
    free(vv);
    free(dd);
    free(nl);
  }
  /* ghosts: phiL from the boundary cell and the solid normal; vofL copied */
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j) {
      phi[HX(0, j, k)] = phi[HX(1, j, k)] - in->nx[HX(1, j, k)] * in->dx[CX(1, j, k)];
      phi[HX(I + 1, j, k)] = phi[HX(I, j, k)] + in->nx[HX(I, j, k)] * in->dx[CX(I, j, k)];
    }
  for (int k = 1; k <= K; ++k)
    for (int i = 1; i <= I; ++i) {
      phi[HX(i, 0, k)] = phi[HX(i, 1, k)] - in->ny[HX(i, 1, k)] * in->dy[CX(i, 1, k)];
      phi[HX(i, J + 1, k)] = phi[HX(i, J, k)] - in->ny[HX(i, J, k)] * in->dy[CX(i, J, k)];
    }
  for (int j = 1; j <= J; ++j)
    for (int i = 1; i <= I; ++i) {
      phi[HX(i, j, 0)] = phi[HX(i, j, 1)] - in->nz[HX(i, j, 1)] * in->dz[CX(i, j, 1)];
      phi[HX(i, j, K + 1)] = phi[HX(i, j, K)] + in->nz[HX(i, j, K)] * in->dz[CX(i, j, K)];
    }
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j) {
      vof[HX(0, j, k)] = vof[HX(1, j, k)];
      vof[HX(I + 1, j, k)] = vof[HX(I, j, k)];
    }
  for (int k = 1; k <= K; ++k)
    for (int i = 1; i <= I; ++i) {
      vof[HX(i, 0, k)] = vof[HX(i, 1, k)];
      vof[HX(i, J + 1, k)] = vof[HX(i, J, k)];
    }
  for (int j = 1; j <= J; ++j)
    for (int i = 1; i <= I; ++i) {
      vof[HX(i, j, 0)] = vof[HX(i, j, 1)];
      vof[HX(i, j, K + 1)] = vof[HX(i, j, K)];
    }
}

/* Iscutface (:1929-1961) */
static int iscutface(double fp, double nxx, double nyy, double nzz, double diss, double dx,
                     double dy, double dz, int cas) {
  double p1, p2, p3, p4;
  if (cas == 1) {
    p1 = fp * nxx - 0.5 * dy * nyy - 0.5 * dz * nzz + diss;
    p2 = fp * nxx - 0.5 * dy * nyy + 0.5 * dz * nzz + diss;
    p3 = fp * nxx + 0.5 * dy * nyy + 0.5 * dz * nzz + diss;
    p4 = fp * nxx + 0.5 * dy * nyy - 0.5 * dz * nzz + diss;
  } else if (cas == 2) {
    p1 = -0.5 * dx * nxx + fp * nyy - 0.5 * dz * nzz + diss;
    p2 = -0.5 * dx * nxx + fp * nyy + 0.5 * dz * nzz + diss;
    p3 = 0.5 * dx * nxx + fp * nyy + 0.5 * dz * nzz + diss;
    p4 = 0.5 * dx * nxx + fp * nyy - 0.5 * dz * nzz + diss;
  } else {
    p1 = -0.5 * dx * nxx - 0.5 * dy * nyy + fp * nzz + diss;
    p2 = -0.5 * dx * nxx + 0.5 * dy * nyy + fp * nzz + diss;
    p3 = 0.5 * dx * nxx + 0.5 * dy * nyy + fp * nzz + diss;
    p4 = 0.5 * dx * nxx - 0.5 * dy * nyy + fp * nzz + diss;
  }
  return !(p1 * p2 >= 0.0 && p2 * p3 >= 0.0 && p3 * p4 >= 0.0 && p4 * p1 >= 0.0);
}

/* Shortest_Point_2d (:1963-2062): the point of the interface line in the
 * face plane `cas` nearest to the offset cell centre (ii, jj, kk) */
static void shortest_point_2d(double nx1, double ny1, double nz1, double dis1, int ii, int jj,
                              int kk, int cas, double dx, double dy, double dz, double* xs,
                              double* ys, double* zs) {
  const double epsi = 1.0e-30;
  if (cas == 1) {
    const double t = sqrt(ny1 * ny1 + nz1 * nz1 + epsi * epsi);
    const double nyy = ny1 / t, nzz = nz1 / t, diss = dis1 / t;
    const double dij1 = nyy * dy * (double)jj + nzz * dz * (double)kk + diss;
    const double yp = dy * (double)jj - dij1 * nyy, zp = dz * (double)kk - dij1 * nzz;
    if (-0.5 * dy <= yp && yp <= 0.5 * dy && -0.5 * dz <= zp && zp <= 0.5 * dz) {
      *ys = yp;
      *zs = zp;
      return;
    }
    const double yoff = fmax(fabs(yp) - 0.5 * dy, 0.0), zoff = fmax(fabs(zp) - 0.5 * dz, 0.0);
    const double yfc = copysign(1.0, yp) * 0.5 * dy, zfc = copysign(1.0, zp) * 0.5 * dz;
    if (yoff * fabs(nyy) > zoff * fabs(nzz)) {
      *ys = yfc;
      *zs = (diss + nyy * *ys) / (copysign(1.0, -nzz) * fmax(fabs(nzz), TOLPAR));
    } else {
      *zs = zfc;
      *ys = (diss + nzz * *zs) / (copysign(1.0, -nyy) * fmax(fabs(nyy), TOLPAR));
    }
  } else if (cas == 2) {
    const double t = sqrt(nx1 * nx1 + nz1 * nz1 + epsi * epsi);
    const double nxx = nx1 / t, nzz = nz1 / t, diss = dis1 / t;
    const double dij1 = nxx * dx * (double)ii + nzz * dz * (double)kk + diss;
    const double xp = dx * (double)ii - dij1 * nxx, zp = dz * (double)kk - dij1 * nzz;
    if (-0.5 * dx <= xp && xp <= 0.5 * dx && -0.5 * dz <= zp && zp <= 0.5 * dz) {
      *xs = xp;
      *zs = zp;
      return;
    }
    const double xoff = fmax(fabs(xp) - 0.5 * dx, 0.0), zoff = fmax(fabs(zp) - 0.5 * dz, 0.0);
    const double xfc = copysign(1.0, xp) * 0.5 * dx, zfc = copysign(1.0, zp) * 0.5 * dz;
    if (xoff * fabs(nxx) > zoff * fabs(nzz)) {
      *xs = xfc;
      *zs = (diss + nxx * *xs) / (copysign(1.0, -nzz) * fmax(fabs(nzz), TOLPAR));
    } else {
      /* the reference bounds the divisor with |nzz| here (:2035) */
      *zs = zfc;
      *xs = (diss + nzz * *zs) / (copysign(1.0, -nxx) * fmax(fabs(nzz), TOLPAR));
    }
  } else {
    const double t = sqrt(nx1 * nx1 + ny1 * ny1 + epsi * epsi);
    const double nxx = nx1 / t, nyy = ny1 / t, diss = dis1 / t;
    const double dij1 = nxx * dx * (double)ii + nyy * dy * (double)jj + diss;
    const double xp = dx * (double)ii - dij1 * nxx, yp = dy * (double)jj - dij1 * nyy;
    if (-0.5 * dx <= xp && xp <= 0.5 * dx && -0.5 * dy <= yp && yp <= 0.5 * dy) {
      *xs = xp;
      *ys = yp;
      return;
    }
    const double xoff = fmax(fabs(xp) - 0.5 * dx, 0.0), yoff = fmax(fabs(yp) - 0.5 * dy, 0.0);
    const double xfc = copysign(1.0, xp) * 0.5 * dx, yfc = copysign(1.0, yp) * 0.5 * dy;
    if (xoff * fabs(nxx) > yoff * fabs(nyy)) {
      *xs = xfc;
      *ys = (diss + nxx * *xs) / (copysign(1.0, -nyy) * fmax(fabs(nyy), TOLPAR));
    } else {
      *ys = yfc;
      *xs = (diss + nyy * *ys) / (copysign(1.0, -nxx) * fmax(fabs(nxx), TOLPAR));
    }
  }
}

/* The candidate distance Redistance_Levelset gives the cell at offset
 * (ii, jj, kk) from an interface cell with plane (n, diss) and size d
 * (:1738-1860); phit is the target's current level set (its sign picks the
 * branch). */
double orc_redist_candidate(double nxx, double nyy, double nzz, double diss, double dx, double dy,
                            double dz, int ii, int jj, int kk, double phit) {
  const int l = ii < -1 ? -1 : (ii > 1 ? 1 : ii);
  const int m = jj < -1 ? -1 : (jj > 1 ? 1 : jj);
  const int n = kk < -1 ? -1 : (kk > 1 ? 1 : kk);
  const double xv = dx * (double)l / 2.0, yv = dy * (double)m / 2.0, zv = dz * (double)n / 2.0;
  const double dv = nxx * xv + nyy * yv + nzz * zv + diss;
  const double ox = dx * (double)ii, oy = dy * (double)jj, oz = dz * (double)kk;
  if (dv * copysign(1.0, phit) <= 0.0)
    return sqrt((ox - xv) * (ox - xv) + (oy - yv) * (oy - yv) + (oz - zv) * (oz - zv));
  const double dij1 = nxx * dx * (double)ii + nyy * dy * (double)jj + nzz * dz * (double)kk + diss;
  const double xp = ox - dij1 * nxx, yp = oy - dij1 * nyy, zp = oz - dij1 * nzz;
  if (-0.5 * dx <= xp && xp <= 0.5 * dx && -0.5 * dy <= yp && yp <= 0.5 * dy && -0.5 * dz <= zp &&
      zp <= 0.5 * dz)
    return sqrt((ox - xp) * (ox - xp) + (oy - yp) * (oy - yp) + (oz - zp) * (oz - zp));
  const double xoff = fmax(fabs(xp) - 0.5 * dx, 0.0), yoff = fmax(fabs(yp) - 0.5 * dy, 0.0),
               zoff = fmax(fabs(zp) - 0.5 * dz, 0.0);
  const double xfc = copysign(1.0, xp) * 0.5 * dx, yfc = copysign(1.0, yp) * 0.5 * dy,
               zfc = copysign(1.0, zp) * 0.5 * dz;
  const double ax = xoff * fabs(nxx), ay = yoff * fabs(nyy), az = zoff * fabs(nzz);
  double xs = 0.0, ys = 0.0, zs = 0.0, d1;
  int cas;
  if (ax >= ay && ax >= az) {
    if (iscutface(xfc, nxx, nyy, nzz, diss, dx, dy, dz, 1)) {
      xs = xfc;
      d1 = nxx * xs + diss;
      cas = 1;
    } else if (ay > az) {
      ys = yfc;
      d1 = nyy * ys + diss;
      cas = 2;
    } else {
      zs = zfc;
      d1 = nzz * zs + diss;
      cas = 3;
    }
    shortest_point_2d(nxx, nyy, nzz, d1, ii, jj, kk, cas, dx, dy, dz, &xs, &ys, &zs);
  }
  if (ay >= ax && ay >= az) {
    if (iscutface(yfc, nxx, nyy, nzz, diss, dx, dy, dz, 2)) {
      ys = yfc;
      d1 = nyy * ys + diss;
      cas = 2;
    } else if (ax > az) {
      xs = xfc;
      d1 = nxx * xs + diss;
      cas = 1;
    } else {
      zs = zfc;
      d1 = nzz * zs + diss;
      cas = 3;
    }
    shortest_point_2d(nxx, nyy, nzz, d1, ii, jj, kk, cas, dx, dy, dz, &xs, &ys, &zs);
  }
  if (az >= ax && az >= ay) {
    if (iscutface(zfc, nxx, nyy, nzz, diss, dx, dy, dz, 3)) {
      zs = zfc;
      d1 = nzz * zs + diss;
      cas = 3;
    } else if (ax > ay) {
      xs = xfc;
      d1 = nxx * xs + diss;
      cas = 1;
    } else {
      ys = yfc;
      d1 = nyy * ys + diss;
      cas = 2;
    }
    shortest_point_2d(nxx, nyy, nzz, d1, ii, jj, kk, cas, dx, dy, dz, &xs, &ys, &zs);
  }
  return sqrt((ox - xs) * (ox - xs) + (oy - ys) * (oy - ys) + (oz - zs) * (oz - zs));
}

/* Redistance_Levelset (:1689-1889) */
static void redistance(int I, int J, int K, const orc_clsvof_in* in, const double* vfl,
                       double* phi, const double* nx, const double* ny, const double* nz,
                       const double* dis) {
  const double epsi = 1.0e-15;
  const int64_t N = (int64_t)I * J * K;
  int* fix = calloc((size_t)N, sizeof(int));
  double* aux = malloc(sizeof(double) * (size_t)N);
  for (int64_t q = 0; q < N; ++q) aux[q] = 1.0e4;
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        const int64_t h = HX(i, j, k), c = CX(i, j, k);
        const double vc = vfl[h];
        if (!((0.0 + epsi) < vc && vc < (1.0 - epsi))) continue;
        aux[c] = copysign(1.0, 0.5 - vc) * fabs(dis[h]);
        fix[c] = 1;
        for (int ii = -BANDW; ii <= BANDW; ++ii)
          for (int jj = -BANDW; jj <= BANDW; ++jj)
            for (int kk = -BANDW; kk <= BANDW; ++kk) {
              const int a = i + ii, b = j + jj, e = k + kk;
              if (!(1 <= a && a <= I && 1 <= b && b <= J && 1 <= e && e <= K)) continue;
              const double vt = vfl[HX(a, b, e)];
              if (!(vt < 0.0 + epsi || vt > 1.0 - epsi)) continue;
              const int64_t ct = CX(a, b, e);
              if (fix[ct] == 0) fix[ct] = 1;
              const double deuc = orc_redist_candidate(nx[h], ny[h], nz[h], dis[h], in->dx[c],
                                                       in->dy[c], in->dz[c], ii, jj, kk,
                                                       phi[HX(a, b, e)]);
              const double cur = fabs(aux[ct]);
              aux[ct] = copysign(1.0, 0.5 - vt) * (deuc < cur ? deuc : cur);
            }
      }
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        const int64_t c = CX(i, j, k), h = HX(i, j, k);
        if (fix[c] == 1 || fix[c] == 2) phi[h] = aux[c];
        else phi[h] = (double)(BANDW + 1) * in->dx[c] * sqrt(3.0) * copysign(1.0, 0.5 - vfl[h]);
      }
  free(fix);
  free(aux);
}

int orc_clsvof_scheme(int I, int J, int K, const orc_clsvof_in* in, orc_clsvof_state* st) {
  if (I < 3 || J < 3 || K < 3) return -1;
  const int64_t nh = (int64_t)(I + 2) * (J + 2) * (K + 2);
  double* buf = calloc((size_t)(10 * nh), sizeof(double));
  if (!buf) return -2;
  clsvof_work W = {I, J, K, in, buf, buf + nh, buf + 2 * nh, buf + 3 * nh};
  double* tv[3] = {buf + 4 * nh, buf + 5 * nh, buf + 6 * nh};
  double* tl[3] = {buf + 7 * nh, buf + 8 * nh, buf + 9 * nh};
  double* vfl = st->vof;
  double* phi = st->phi;
  const double dtv = in->dt / (double)NV;
  const double* ve[3] = {in->u, in->v, in->w};
  const double* dd[3] = {in->dx, in->dy, in->dz};
  /* BCLvs%VarW, VarS, VarB feed the x, y, z sweeps */
  for (int kk = 1; kk <= NV; ++kk) {
    const int r = kk % 3;
    /* sweep order: 0 -> x y z, 1 -> y z x, 2 -> z x y */
    const int order[3] = {r, (r + 1) % 3, (r + 2) % 3};
    const int first = order[0];
    for (int s = 0; s < 3; ++s) {
      const int ax = order[s];
      memcpy(tv[ax], vfl, sizeof(double) * (size_t)nh);
      memcpy(tl[ax], phi, sizeof(double) * (size_t)nh);
      sweep(&W, ax, vfl, phi, tv[ax], tl[ax], st->lvs_var[2 * ax], dtv);
      const int64_t sx = ax == 0 ? 1 : (ax == 1 ? (int64_t)(I + 2) : (int64_t)(I + 2) * (J + 2));
      for (int i = 1; i <= I; ++i)
        for (int j = 1; j <= J; ++j)
          for (int k = 1; k <= K; ++k) {
            const int64_t h = HX(i, j, k), c = CX(i, j, k);
            const double du = ve[ax][h] - ve[ax][h - sx];
            const double q = dtv / dd[ax][c];
            if (s == 0) {
              tv[ax][h] = clip(tv[ax][h] / (1.0 - q * du));
              tl[ax][h] = tl[ax][h] / (1.0 - q * du);
              vfl[h] = tv[ax][h];
              phi[h] = tl[ax][h];
            } else if (s == 1) {
              tv[ax][h] = clip(tv[ax][h] + q * tv[first][h] * du);
              tl[ax][h] = tl[ax][h] + q * tl[first][h] * du;
              vfl[h] = tv[ax][h];
              phi[h] = tl[ax][h];
            } else {
              vfl[h] = clip(tv[ax][h] + q * tv[first][h] * du);
              phi[h] = tl[ax][h] + q * tl[first][h] * du;
            }
          }
    }
    bc_lvs_vof(I, J, K, in, st);
    reconstruct(&W, vfl, phi);
    for (int i = 1; i <= I; ++i)
      for (int j = 1; j <= J; ++j)
        for (int k = 1; k <= K; ++k) {
          const int64_t h = HX(i, j, k), c = CX(i, j, k);
          st->nx[h] = W.nxx[h];
          st->ny[h] = W.nyy[h];
          st->nz[h] = W.nzz[h];
          W.diss[h] = (0.5 * in->dx[c] * fabs(W.nxx[h]) + 0.5 * in->dy[c] * fabs(W.nyy[h]) +
                       0.5 * in->dz[c] * fabs(W.nzz[h])) - W.diss[h];
        }
    redistance(I, J, K, in, vfl, phi, W.nxx, W.nyy, W.nzz, W.diss);
  }
  free(buf);
  return 0;
}
