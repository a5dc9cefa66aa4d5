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

// ---- interface transport (Clsvof.f90) ----
constexpr double kVofEps = 1.0e-14;  // vofeps (Clsvof.f90:32)
constexpr double kTolPar = 1.0e-14;  // tolpar (Clsvof.f90:35)

// Newton_Raphson (Clsvof.f90:1462-1473); the reference has no iteration
// limit, here 200 (never reached by a converging case)
__device__ inline double newton_raphson(double a3, double a2, double a1, double a0, double sc) {
#pragma clang fp contract(off)
  double delvar = 1.0e6;
  for (int it = 0; it < 200 && fabs(delvar) > 1.0e-6 * fabs(sc); ++it) {
    delvar = (a3 * pow(sc, 3.0) + a2 * (sc * sc) + a1 * sc + a0) /
             (3.0 * a3 * (sc * sc) + 2.0 * a2 * sc + a1);
    sc = sc - delvar;
  }
  return sc;
}

// Cubic_Equation_Solve (Clsvof.f90:1380-1398); case2 selects the first form
__device__ inline double cubic_solve(double fc, double dx1, double dy1, double dz1, double sc,
                                     bool case2) {
#pragma clang fp contract(off)
  if (case2) {
    const double a3 = 1.0, a2 = -3.0 * (dy1 + dz1), a1 = 3.0 * (dy1 * dy1 + dz1 * dz1);
    const double a0 = 6.0 * fc * dx1 * dy1 * dz1 - (pow(dy1, 3.0) + pow(dz1, 3.0));
    return newton_raphson(a3, a2, a1, a0, sc);
  }
  const double a3 = 2.0, a2 = -3.0 * (dx1 + dy1 + dz1);
  const double a1 = 3.0 * (dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
  const double a0 = 6.0 * fc * dx1 * dy1 * dz1 - (pow(dx1, 3.0) + pow(dy1, 3.0) + pow(dz1, 3.0));
  return newton_raphson(a3, a2, a1, a0, sc);
}

// Find_Distance with Final_Distance (Clsvof.f90:1311-1408): the distance
// from the cell corner of the plane of normal n that leaves the fraction f
// on the corner side
__device__ inline double find_distance(double dx, double dy, double dz, double nxx, double nyy,
                                       double nzz, double f) {
#pragma clang fp contract(off)
  const double nx1 = fabs(nxx), ny1 = fabs(nyy), nz1 = fabs(nzz);
  const double a = nx1 * dx, b = ny1 * dy, c = nz1 * dz;
  double dx1 = a, dz1 = a;
  if (b > dx1) dx1 = b;
  if (c > dx1) dx1 = c;
  if (b < dz1) dz1 = b;
  if (c < dz1) dz1 = c;
  if (fabs(dz1) < kTolPar) dz1 = 0.0;
  double dy1 = (a + b + c) - dx1 - dz1;
  if (fabs(dy1) < kTolPar) dy1 = 0.0;
  const double fc = f < 1.0 - f ? f : 1.0 - f;
  const double sm = dx1 + dy1 + dz1;
  double sc = 6.0 * fc * dx1 * dy1 * dz1;
  sc = copysign(1.0, sc) * pow(fabs(sc), 1.0 / 3.0);
  if (!(sc < dz1)) {
    sc = 0.5 * dz1 + sqrt(2.0 * fc * dx1 * dy1 - dz1 * dz1 / 12.0);
    if (!(sc < dy1)) {
      if (dx1 >= dy1 + dz1) {
        sc = fc * dx1 + (dy1 + dz1) / 2.0;
        if (!(sc >= dy1 + dz1)) sc = cubic_solve(fc, dx1, dy1, dz1, sc, true);
      } else {
        sc = fc * dx1 + (dy1 + dz1) / 2.0;
        if (sc >= dx1) {
          sc = cubic_solve(fc, dx1, dy1, dz1, sc, false);
          if (sc < dx1) sc = cubic_solve(fc, dx1, dy1, dz1, sc, true);
        } else {
          sc = cubic_solve(fc, dx1, dy1, dz1, sc, true);
        }
      }
    }
  }
  return f <= 0.5 ? sc : sm - sc;
}

// East_Flux .. Bottom_Flux (Clsvof.f90:1552-1687): the liquid volume over
// the face area crossing a face of the sweep axis ax in one sub-step; pos
// selects the East/North/Top form, else West/South/Bottom
__device__ inline double face_flux(int ax, bool pos, double nxx, double nyy, double nzz,
                                   double diss, double volf, double dx, double dy, double dz,
                                   double udt) {
#pragma clang fp contract(off)
  if (udt == 0.0) return 0.0;
  const double nn = ax == 0 ? nxx : (ax == 1 ? nyy : nzz);
  const bool near = pos ? (nn <= 0.0) : (nn >= 0.0);
  double vol, flux;
  if (ax == 0) {
    if (near) {
      vol = volume_fraction(udt, dy, dz, nxx, nyy, nzz, diss);
      flux = vol / (udt * dy * dz);
    } else {
      vol = volume_fraction(dx - udt, dy, dz, nxx, nyy, nzz, diss);
      flux = (volf * dx * dy * dz - vol) / (udt * dy * dz);
    }
  } else if (ax == 1) {
    if (near) {
      vol = volume_fraction(dx, udt, dz, nxx, nyy, nzz, diss);
      flux = vol / (dx * udt * dz);
    } else {
      vol = volume_fraction(dx, dy - udt, dz, nxx, nyy, nzz, diss);
      flux = (volf * dx * dy * dz - vol) / (dx * udt * dz);
    }
  } else {
    if (near) {
      vol = volume_fraction(dx, dy, udt, nxx, nyy, nzz, diss);
      flux = vol / (dx * dy * udt);
    } else {
      vol = volume_fraction(dx, dy, dz - udt, nxx, nyy, nzz, diss);
      flux = (volf * dx * dy * dz - vol) / (dx * dy * udt);
    }
  }
  return flux * udt;
}

// Iscutface (Clsvof.f90:1929-1961)
__device__ inline bool iscutface(double fp, double nxx, double nyy, double nzz, double diss,
                                 double dx, double dy, double dz, int cas) {
#pragma clang fp contract(off)
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

// Shortest_Point_2d (Clsvof.f90:1963-2062): the point of the interface line
// in face plane `cas` nearest to the offset cell centre (ii, jj, kk); the
// y-face form bounds its last divisor with |nzz| as the reference does
__device__ inline void shortest_point_2d(double nx1, double ny1, double nz1, double dis1, int ii,
                                         int jj, int kk, int cas, double dx, double dy, double dz,
                                         double& xs, double& ys, double& zs) {
#pragma clang fp contract(off)
  const double epsi = 1.0e-30;
  if (cas == 1) {
    const double t = sqrt(ny1 * ny1 + nz1 * nz1 + epsi * epsi);
    const double nyy = ny1 / t, nzz = nz1 / t, diss = dis1 / t;
    const double dij1 = nyy * dy * (double)jj + nzz * dz * (double)kk + diss;
    const double yp = dy * (double)jj - dij1 * nyy, zp = dz * (double)kk - dij1 * nzz;
    if (-0.5 * dy <= yp && yp <= 0.5 * dy && -0.5 * dz <= zp && zp <= 0.5 * dz) {
      ys = yp;
      zs = zp;
      return;
    }
    const double yoff = fmax(fabs(yp) - 0.5 * dy, 0.0), zoff = fmax(fabs(zp) - 0.5 * dz, 0.0);
    const double yfc = copysign(1.0, yp) * 0.5 * dy, zfc = copysign(1.0, zp) * 0.5 * dz;
    if (yoff * fabs(nyy) > zoff * fabs(nzz)) {
      ys = yfc;
      zs = (diss + nyy * ys) / (copysign(1.0, -nzz) * fmax(fabs(nzz), kTolPar));
    } else {
      zs = zfc;
      ys = (diss + nzz * zs) / (copysign(1.0, -nyy) * fmax(fabs(nyy), kTolPar));
    }
  } else if (cas == 2) {
    const double t = sqrt(nx1 * nx1 + nz1 * nz1 + epsi * epsi);
    const double nxx = nx1 / t, nzz = nz1 / t, diss = dis1 / t;
    const double dij1 = nxx * dx * (double)ii + nzz * dz * (double)kk + diss;
    const double xp = dx * (double)ii - dij1 * nxx, zp = dz * (double)kk - dij1 * nzz;
    if (-0.5 * dx <= xp && xp <= 0.5 * dx && -0.5 * dz <= zp && zp <= 0.5 * dz) {
      xs = xp;
      zs = zp;
      return;
    }
    const double xoff = fmax(fabs(xp) - 0.5 * dx, 0.0), zoff = fmax(fabs(zp) - 0.5 * dz, 0.0);
    const double xfc = copysign(1.0, xp) * 0.5 * dx, zfc = copysign(1.0, zp) * 0.5 * dz;
    if (xoff * fabs(nxx) > zoff * fabs(nzz)) {
      xs = xfc;
      zs = (diss + nxx * xs) / (copysign(1.0, -nzz) * fmax(fabs(nzz), kTolPar));
    } else {
      zs = zfc;
      xs = (diss + nzz * zs) / (copysign(1.0, -nxx) * fmax(fabs(nzz), kTolPar));
    }
  } else {
    const double t = sqrt(nx1 * nx1 + ny1 * ny1 + epsi * epsi);
    const double nxx = nx1 / t, nyy = ny1 / t, diss = dis1 / t;
    const double dij1 = nxx * dx * (double)ii + nyy * dy * (double)jj + diss;
    const double xp = dx * (double)ii - dij1 * nxx, yp = dy * (double)jj - dij1 * nyy;
    if (-0.5 * dx <= xp && xp <= 0.5 * dx && -0.5 * dy <= yp && yp <= 0.5 * dy) {
      xs = xp;
      ys = yp;
      return;
    }
    const double xoff = fmax(fabs(xp) - 0.5 * dx, 0.0), yoff = fmax(fabs(yp) - 0.5 * dy, 0.0);
    const double xfc = copysign(1.0, xp) * 0.5 * dx, yfc = copysign(1.0, yp) * 0.5 * dy;
    if (xoff * fabs(nxx) > yoff * fabs(nyy)) {
      xs = xfc;
      ys = (diss + nxx * xs) / (copysign(1.0, -nyy) * fmax(fabs(nyy), kTolPar));
    } else {
      ys = yfc;
      xs = (diss + nyy * ys) / (copysign(1.0, -nxx) * fmax(fabs(nxx), kTolPar));
    }
  }
}

// The candidate distance Redistance_Levelset gives the cell at offset
// (ii, jj, kk) from an interface cell of plane (n, diss) and size d
// (Clsvof.f90:1738-1860); phit is the target's level set
__device__ inline double redist_candidate(double nxx, double nyy, double nzz, double diss,
                                          double dx, double dy, double dz, int ii, int jj, int kk,
                                          double phit) {
#pragma clang fp contract(off)
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
    if (iscutface(xfc, nxx, nyy, nzz, diss, dx, dy, dz, 1)) { xs = xfc; d1 = nxx * xs + diss; cas = 1; }
    else if (ay > az) { ys = yfc; d1 = nyy * ys + diss; cas = 2; }
    else { zs = zfc; d1 = nzz * zs + diss; cas = 3; }
    shortest_point_2d(nxx, nyy, nzz, d1, ii, jj, kk, cas, dx, dy, dz, xs, ys, zs);
  }
  if (ay >= ax && ay >= az) {
    if (iscutface(yfc, nxx, nyy, nzz, diss, dx, dy, dz, 2)) { ys = yfc; d1 = nyy * ys + diss; cas = 2; }
    else if (ax > az) { xs = xfc; d1 = nxx * xs + diss; cas = 1; }
    else { zs = zfc; d1 = nzz * zs + diss; cas = 3; }
    shortest_point_2d(nxx, nyy, nzz, d1, ii, jj, kk, cas, dx, dy, dz, xs, ys, zs);
  }
  if (az >= ax && az >= ay) {
    if (iscutface(zfc, nxx, nyy, nzz, diss, dx, dy, dz, 3)) { zs = zfc; d1 = nzz * zs + diss; cas = 3; }
    else if (ax > ay) { xs = xfc; d1 = nxx * xs + diss; cas = 1; }
    else { ys = yfc; d1 = nyy * ys + diss; cas = 2; }
    shortest_point_2d(nxx, nyy, nzz, d1, ii, jj, kk, cas, dx, dy, dz, xs, ys, zs);
  }
  return sqrt((ox - xs) * (ox - xs) + (oy - ys) * (oy - ys) + (oz - zs) * (oz - zs));
}

}  // namespace sp
