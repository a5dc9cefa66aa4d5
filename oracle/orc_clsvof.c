/* CPU restatement of the level-set pieces of the interface transport and of
 * the moving-solid geometry update -- test infrastructure only (the tests,
 * smoke() and the bench's CPU leg use it as the checker):
 *
 *   orc_normal_vector_irre  Normal_Vector_Irre (Clsvof.f90:2361-2456)
 *   orc_interpolate_lvs     TSolidBody%interpolateLvS (SolidBody.f90:646-742)
 *
 * Loop orders and the reference's operation order are kept, with
 * floating-point contraction off (the Makefile's -ffp-contract=off), so the
 * values are those of the reference's loops; x**2.d0 and x**2 are x*x.
 * Halo arrays are (0:I+1, 0:J+1, 0:K+1), cell arrays (1:I, 1:J, 1:K), both
 * i fastest. */

#include <math.h>
#include <stdint.h>

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
