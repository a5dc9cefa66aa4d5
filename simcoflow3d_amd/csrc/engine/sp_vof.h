// Device restatements of the geometric helpers of the interface transport
// (Clsvof.f90) shared by the liquid-field, solid level-set and transport
// kernels.  Arithmetic in the reference's operation order with
// floating-point contraction off; x**2.d0 is x*x, x**3.d0 (a real exponent)
// a pow call, as in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sp {

// ConvertTheCoordinateSystem (Clsvof.f90:2458-2465): the level set measured
// from the cell corner
__device__ inline double corner_dist(double dx, double dy, double dz, double nx, double ny,
                                     double nz, double phi) {
#pragma clang fp contract(off)
  return phi + 0.5 * (fabs(nx) * dx + fabs(ny) * dy + fabs(nz) * dz);
}

// Volume_Fraction_Calc (Clsvof.f90:1475-1550): the volume of the box on the
// corner side of the plane at distance s (Griebel and Klitz)
__device__ inline double volume_fraction(double dx, double dy, double dz, double nxx,
                                         double nyy, double nzz, double s) {
#pragma clang fp contract(off)
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
    if (sc > dx1)
      fc = (pow(sc, 3.0) - pow(sc - dz1, 3.0) - pow(sc - dy1, 3.0) - pow(sc - dx1, 3.0));
    else
      fc = (pow(sc, 3.0) - pow(sc - dz1, 3.0) - pow(sc - dy1, 3.0));
    fc = fc / (6.0 * dx1 * dy1 * dz1);
  }
  if (s <= 0.5 * sm) return fc * dx * dy * dz;
  return (1.0 - fc) * dx * dy * dz;
}

// index of (i, j, k) in a halo array (0:I+1, 0:J+1, 0:K+1) and in a cell
// array (1:I, 1:J, 1:K)
__device__ inline int64_t hidx(int I, int J, int i, int j, int k) {
  return (int64_t)i + (int64_t)(I + 2) * ((int64_t)j + (int64_t)(J + 2) * k);
}
__device__ inline int64_t cidx(int I, int J, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)I * ((int64_t)(j - 1) + (int64_t)J * (k - 1));
}

__device__ inline double nv_q(double vx, double vy, double vz) {
#pragma clang fp contract(off)
  return fabs(1.0 - sqrt(vx * vx + vy * vy + vz * vz));
}

// Normal_Vector_Irre (Clsvof.f90:2361-2456): the level-set gradient at an
// interior cell, one-sided towards a side whose two-cell difference keeps
// |grad| within eta = 0.075 of 1 while the other side's does not.  phi is a
// halo array, x, y, z cell arrays of the cell centres.
__device__ inline void normal_vector_irre(int I, int J, int K, const double* __restrict__ x,
                                          const double* __restrict__ y,
                                          const double* __restrict__ z,
                                          const double* __restrict__ phi, int i, int j, int k,
                                          double& nxx, double& nyy, double& nzz) {
#pragma clang fp contract(off)
  constexpr double eta = 0.075;
  auto P = [&](int a, int b, int c) { return phi[hidx(I, J, a, b, c)]; };
  auto X = [&](int a, int b, int c) { return x[cidx(I, J, a, b, c)]; };
  auto Y = [&](int a, int b, int c) { return y[cidx(I, J, a, b, c)]; };
  auto Z = [&](int a, int b, int c) { return z[cidx(I, J, a, b, c)]; };
  double vx, vy, vz, qm, qp;
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
  const int sx = (qm < eta && qp >= eta) ? -1 : ((qm >= eta && qp < eta) ? 1 : 0);
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
  const int sy = (qm < eta && qp >= eta) ? -1 : ((qm >= eta && qp < eta) ? 1 : 0);
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
  const int sz = (qm < eta && qp >= eta) ? -1 : ((qm >= eta && qp < eta) ? 1 : 0);
  if (sx == -1) nxx = (P(i, j, k) - P(i - 1, j, k)) / (X(i, j, k) - X(i - 1, j, k));
  else if (sx == 1) nxx = (P(i + 1, j, k) - P(i, j, k)) / (X(i + 1, j, k) - X(i, j, k));
  else nxx = (P(i + 1, j, k) - P(i - 1, j, k)) / (X(i + 1, j, k) - X(i - 1, j, k));
  if (sy == -1) nyy = (P(i, j, k) - P(i, j - 1, k)) / (Y(i, j, k) - Y(i, j - 1, k));
  else if (sy == 1) nyy = (P(i, j + 1, k) - P(i, j, k)) / (Y(i, j + 1, k) - Y(i, j, k));
  else nyy = (P(i, j + 1, k) - P(i, j - 1, k)) / (Y(i, j + 1, k) - Y(i, j - 1, k));
  if (sz == -1) nzz = (P(i, j, k) - P(i, j, k - 1)) / (Z(i, j, k) - Z(i, j, k - 1));
  else if (sz == 1) nzz = (P(i, j, k + 1) - P(i, j, k)) / (Z(i, j, k + 1) - Z(i, j, k));
  else nzz = (P(i, j, k + 1) - P(i, j, k - 1)) / (Z(i, j, k + 1) - Z(i, j, k - 1));
}

}  // namespace sp
