// ComputeUVWLiquidField (Clsvof.f90:174-298): the liquid level set, normal
// and volume fraction of the U, V, W cells from the P cells, which the
// interface transport hands to the GFM coefficients of the projection
// (Solver.f90:209).
//
//   ComputeVelocityCellLiquidField / DirectionAverageArray (:300-353):
//     phiL, nxL, nyL, nzL averaged to the face between two P centres with
//     the spacing weight lambda = d(i+1) / (d(i) + d(i+1)); the last cell
//     of a line copies the one before it.  (Its vof + vofL > 1 clamp acts
//     on a vofL that the loops below overwrite everywhere, so it has no
//     effect on the result and is not repeated.)
//   vofL: the liquid volume of two half-width boxes, one from each P cell
//     of the face (Volume_Fraction_Calc, :1475-1550, measured from the cell
//     corner by ConvertTheCoordinateSystem, :2458-2465), over the velocity
//     cell's volume; the last cell of a line copies the P cell's vofL.  The
//     V and W loops keep the reference's own forms: the south part of V
//     reads phiL with the full dy for the corner distance, the north part
//     moves phiL(i,j,k) by nyL(i,j+1,k), the top part of W moves
//     phiL(i,j,k+1) upward.
//
// One thread per cell writes the U, V and W values of that cell (15
// arrays); every value depends on the P cell and its +1 neighbours only.
// The reference's x**3.d0 (a real exponent) is a pow call, here as there.

#include <hip/hip_runtime.h>

#include "sp_device.h"
#include "sp_internal.h"
#include "sp_vof.h"

namespace sp {
namespace {

struct LiquidArgs {
  int I, J, K;
  const double *pdx, *pdy, *pdz, *phi, *nx, *ny, *nz, *pvof;
  const double *vd[3][3];  // U/V/W spacing dx, dy, dz
  double *out[3][5];       // U/V/W phiL, nxL, nyL, nzL, vofL
};

// liquid volume of a box (d1, d2, d3) whose centre level set is phin
__device__ inline double liquid_part(double d1, double d2, double d3, double nx, double ny,
                                     double nz, double phin) {
#pragma clang fp contract(off)
  const double dist = corner_dist(d1, d2, d3, nx, ny, nz, phin);
  const double v = volume_fraction(d1, d2, d3, nx, ny, nz, dist);
  return d1 * d2 * d3 - v;
}

__global__ void __launch_bounds__(256) k_liquid(LiquidArgs A) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > A.I || j > A.J) return;
  const int64_t IJ = (int64_t)A.I * A.J;
  const int64_t q = (int64_t)(i - 1) + (int64_t)A.I * (j - 1) + IJ * (k - 1);
  const int64_t step[3] = {1, (int64_t)A.I, IJ};
  const int idx[3] = {i, j, k};
  const int lim[3] = {A.I, A.J, A.K};
  const double* pd[3] = {A.pdx, A.pdy, A.pdz};
  const double* in[4] = {A.phi, A.nx, A.ny, A.nz};
  for (int m = 0; m < 3; ++m) {
    // DirectionAverageArray: the last cell of a line takes the value of the
    // one before it, which is the average of the same two P cells
    const int64_t qa = idx[m] < lim[m] ? q : q - step[m];
    const int64_t qb = qa + step[m];
    const double lamda = pd[m][qb] / (pd[m][qa] + pd[m][qb]);
#pragma unroll
    for (int f = 0; f < 4; ++f) A.out[m][f][q] = lamda * in[f][qa] + (1.0 - lamda) * in[f][qb];
  }
  const double dx = A.pdx[q], dy = A.pdy[q], dz = A.pdz[q], ph = A.phi[q];
  const double n1 = A.nx[q], n2 = A.ny[q], n3 = A.nz[q];
  // U (:189-216)
  if (i < A.I) {
    const int64_t r = q + 1;
    double dc = dx / 2.0;
    double phin = ph + dc * n1;
    const double vl = liquid_part(dc, dy, dz, n1, n2, n3, phin);
    dc = A.pdx[r] / 2.0;
    phin = A.phi[r] - dc * A.nx[r];
    const double vr = liquid_part(dc, A.pdy[r], A.pdz[r], A.nx[r], A.ny[r], A.nz[r], phin);
    A.out[0][4][q] = (vl + vr) / A.vd[0][0][q] / A.vd[0][1][q] / A.vd[0][2][q];
  } else {
    A.out[0][4][q] = A.pvof[q];
  }
  // V (:218-257)
  if (j < A.J) {
    const int64_t r = q + A.I;
    double dc = dy / 2.0;
    double dist = corner_dist(dx, dy, dz, n1, n2, n3, ph);
    double v = volume_fraction(dx, dc, dz, n1, n2, n3, dist);
    const double vl = dx * dc * dz - v;
    dc = A.pdy[r] / 2.0;
    const double phin = ph + dc * A.ny[r];
    dist = corner_dist(A.pdx[r], dc, A.pdz[r], A.nx[r], A.ny[r], A.nz[r], phin);
    v = volume_fraction(A.pdx[r], dc, A.pdz[r], A.nx[r], A.ny[r], A.nz[r], dist);
    const double vr = A.pdx[r] * dc * A.pdz[r] - v;
    A.out[1][4][q] = (vl + vr) / A.vd[1][0][q] / A.vd[1][1][q] / A.vd[1][2][q];
  } else {
    A.out[1][4][q] = A.pvof[q];
  }
  // W (:259-296)
  if (k < A.K) {
    const int64_t r = q + IJ;
    double dc = dz / 2.0;
    double phin = ph + dc * n3;
    const double vl = liquid_part(dx, dy, dc, n1, n2, n3, phin);
    dc = A.pdz[r] / 2.0;
    phin = A.phi[r] + dc * A.nz[r];
    const double vr = liquid_part(A.pdx[r], A.pdy[r], dc, A.nx[r], A.ny[r], A.nz[r], phin);
    A.out[2][4][q] = (vl + vr) / A.vd[2][0][q] / A.vd[2][1][q] / A.vd[2][2][q];
  } else {
    A.out[2][4][q] = A.pvof[q];
  }
}

}  // namespace

sp_status uvw_liquid_field(sp_ctx* c, const sp_liquid_field* const g[4]) {
  const Dims& D = c->dm;
  LiquidArgs A;
  A.I = D.I; A.J = D.J; A.K = D.K;
  const sp_liquid_field& P = *g[0];
  A.pdx = P.dx; A.pdy = P.dy; A.pdz = P.dz;
  A.phi = P.phi_l; A.nx = P.nx_l; A.ny = P.ny_l; A.nz = P.nz_l; A.pvof = P.vof_l;
  for (int m = 0; m < 3; ++m) {
    const sp_liquid_field& T = *g[1 + m];
    A.vd[m][0] = T.dx; A.vd[m][1] = T.dy; A.vd[m][2] = T.dz;
    A.out[m][0] = T.phi_l; A.out[m][1] = T.nx_l; A.out[m][2] = T.ny_l; A.out[m][3] = T.nz_l;
    A.out[m][4] = T.vof_l;
  }
  const dim3 blk(64, 4, 1);
  const dim3 grid((unsigned)((D.I + 63) / 64), (unsigned)((D.J + 3) / 4), (unsigned)D.K);
  int slot;
  // per cell: P spacing x3, phiL, nxL, nyL, nzL, vofL and U/V/W spacing x9
  // read; 15 doubles written
  timer_begin(c, "liquid_uvw", (double)D.N * (17 * 8 + 15 * 8), &slot);
  hipLaunchKernelGGL(k_liquid, grid, blk, 0, c->stream, A);
  timer_end(c, slot);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "ComputeUVWLiquidField launch");
}

}  // namespace sp
