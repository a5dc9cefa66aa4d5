// Level-set geometry on the GPU: the solid body's level set interpolated
// onto a global grid (interpolateLvS, SolidBody.f90:646-742), the first step
// of the moving-solid geometry update (SolidBody.f90:810-813, followed by
// GridPreProcess, DefineMomentumExchangeCell, NumberExternalCell and
// NewCellFace, which the engine also runs).
//
//   k_lvs_interp  one thread per halo cell: phi = 1e4 and vof = 1 on the
//                 whole array (:659-661); an interior cell centre taken into
//                 the body frame (pointL = R^T (x - CM), :618-627) whose
//                 lower local node lies in [1, imax-1] x [1, jmax-1] x
//                 [1, kmax-1] gets the trilinear interpolation of the body's
//                 phi, its eight terms summed in the reference's order.
//   k_lvs_band    one thread per interior cell with |phi| <= 4 max(dx, dy,
//                 dz) (all interpolation done): the normal (Normal_Vector_Irre
//                 inside, clamped central differences on the outer layer),
//                 normalised ((0, 0, 1) below 1e-14), and vof from the
//                 plane-cut volume of the corner distance.  Outside the band
//                 the normals keep their values and vof stays 1 -- inside
//                 the body too, as in the reference.
//
// Arithmetic in the reference's order with contraction off; the values are
// those of the reference's loops except where Volume_Fraction_Calc's
// x**3.d0 (a pow call) rounds differently on the GPU.

#include <hip/hip_runtime.h>

#include <vector>

#include "sp_internal.h"
#include "sp_vof.h"

namespace sp {
namespace {

struct LvsArgs {
  int I, J, K;
  // body
  int bI, bJ, bK;
  double bdx, bdy, bdz;
  const double *bx, *by, *bz, *bphi;
  double cm[3], rot[9];
  // grid
  const double *x, *y, *z, *dx, *dy, *dz;
  double *phi, *vof, *nx, *ny, *nz;
};

__global__ void __launch_bounds__(256) k_lvs_interp(LvsArgs A) {
#pragma clang fp contract(off)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nh = (int64_t)(A.I + 2) * (A.J + 2) * (A.K + 2);
  if (q >= nh) return;
  const int i = (int)(q % (A.I + 2));
  const int64_t r = q / (A.I + 2);
  const int j = (int)(r % (A.J + 2)), k = (int)(r / (A.J + 2));
  double v = 1.0e4;
  if (i >= 1 && i <= A.I && j >= 1 && j <= A.J && k >= 1 && k <= A.K) {
    const int64_t c = cidx(A.I, A.J, i, j, k);
    const double d[3] = {A.x[c] - A.cm[0], A.y[c] - A.cm[1], A.z[c] - A.cm[2]};
    double pl[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) s = s + A.rot[b + 3 * a] * d[b];
      pl[a] = s;
    }
    const int iL = (int)ceil((pl[0] - A.bx[0]) / A.bdx);
    const int jL = (int)ceil((pl[1] - A.by[0]) / A.bdy);
    const int kL = (int)ceil((pl[2] - A.bz[0]) / A.bdz);
    if (iL >= 1 && iL <= A.bI - 1 && jL >= 1 && jL <= A.bJ - 1 && kL >= 1 && kL <= A.bK - 1) {
      v = 0.0;
      for (int ii = 0; ii <= 1; ++ii)
        for (int jj = 0; jj <= 1; ++jj)
          for (int kk = 0; kk <= 1; ++kk) {
            double coeff = (A.bdx - fabs(pl[0] - A.bx[iL + ii - 1])) / A.bdx;
            coeff = coeff * (A.bdy - fabs(pl[1] - A.by[jL + jj - 1]));
            coeff = coeff / A.bdy;
            coeff = coeff * (A.bdz - fabs(pl[2] - A.bz[kL + kk - 1]));
            coeff = coeff / A.bdz;
            const int64_t b = (int64_t)(iL + ii - 1) +
                              (int64_t)A.bI * ((int64_t)(jL + jj - 1) + (int64_t)A.bJ * (kL + kk - 1));
            v = v + coeff * A.bphi[b];
          }
    }
  }
  A.phi[q] = v;
  A.vof[q] = 1.0;
}

__global__ void __launch_bounds__(256) k_lvs_band(LvsArgs A) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)A.I * A.J * A.K) return;
  const int i = (int)(t % A.I) + 1;
  const int64_t r = t / A.I;
  const int j = (int)(r % A.J) + 1, k = (int)(r / A.J) + 1;
  const int I = A.I, J = A.J, K = A.K;
  const int64_t h = hidx(I, J, i, j, k), c = t;
  const double ddx = A.dx[c], ddy = A.dy[c], ddz = A.dz[c];
  double dm = ddx;
  if (ddy > dm) dm = ddy;
  if (ddz > dm) dm = ddz;
  const double ph = A.phi[h];
  if (fabs(ph) > 4.0 * dm) return;
  double a, b, e;
  if (i > 1 && i < I && j > 1 && j < J && k > 1 && k < K) {
    normal_vector_irre(I, J, K, A.x, A.y, A.z, A.phi, i, j, k, a, b, e);
  } else {
    const int ip = i + 1 < I ? i + 1 : I, im = i - 1 > 1 ? i - 1 : 1;
    const int jp = j + 1 < J ? j + 1 : J, jm = j - 1 > 1 ? j - 1 : 1;
    const int kp = k + 1 < K ? k + 1 : K, km = k - 1 > 1 ? k - 1 : 1;
    a = (A.phi[hidx(I, J, ip, j, k)] - A.phi[hidx(I, J, im, j, k)]) /
        (A.x[cidx(I, J, ip, j, k)] - A.x[cidx(I, J, im, j, k)]);
    b = (A.phi[hidx(I, J, i, jp, k)] - A.phi[hidx(I, J, i, jm, k)]) /
        (A.y[cidx(I, J, i, jp, k)] - A.y[cidx(I, J, i, jm, k)]);
    e = (A.phi[hidx(I, J, i, j, kp)] - A.phi[hidx(I, J, i, j, km)]) /
        (A.z[cidx(I, J, i, j, kp)] - A.z[cidx(I, J, i, j, km)]);
  }
  const double tn = sqrt(a * a + b * b + e * e);
  if (tn < 1.0e-14) {
    a = 0.0;
    b = 0.0;
    e = 1.0;
  } else {
    a = a / tn;
    b = b / tn;
    e = e / tn;
  }
  A.nx[h] = a;
  A.ny[h] = b;
  A.nz[h] = e;
  const double s = corner_dist(ddx, ddy, ddz, a, b, e, ph);
  const double vol = volume_fraction(ddx, ddy, ddz, a, b, e, s);
  A.vof[h] = vol / (ddx * ddy * ddz);
}

}  // namespace

sp_status interpolate_lvs(sp_ctx* c, const sp_solid_body& B, const sp_lvs_grid& g, bool device) {
  const Dims& D = c->dm;
  const int I = D.I, J = D.J, K = D.K;
  const int64_t N = D.N, Nh = D.Nh;
  if (B.imax < 2 || B.jmax < 2 || B.kmax < 2)
    return fail(c, SP_ERR_INVALID_ARG, "solid body mesh must be at least 2 cells in each direction");
  if (!(B.dx > 0.0) || !(B.dy > 0.0) || !(B.dz > 0.0))
    return fail(c, SP_ERR_INVALID_ARG, "solid body spacings must be positive");
  const void* req[] = {B.x, B.y, B.z, B.phi, g.x, g.y, g.z, g.dx, g.dy, g.dz,
                       g.phi, g.vof, g.nx, g.ny, g.nz};
  for (const void* p : req)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "null solid body or grid member");
  const int64_t nb = (int64_t)B.imax * B.jmax * B.kmax;

  LvsArgs A{};
  A.I = I; A.J = J; A.K = K;
  A.bI = B.imax; A.bJ = B.jmax; A.bK = B.kmax;
  A.bdx = B.dx; A.bdy = B.dy; A.bdz = B.dz;
  for (int q = 0; q < 3; ++q) A.cm[q] = B.cm[q];
  for (int q = 0; q < 9; ++q) A.rot[q] = B.rot[q];

  // device copies of host arrays: body x, y, z, phi; grid x..dz; phi, vof
  // (written), nx, ny, nz (read and written)
  std::vector<void*> tmp;
  auto release = [&]() { for (void* p : tmp) hipFree(p); tmp.clear(); };
  auto stage = [&](const void* src, size_t bytes, bool upload, void** dst) -> sp_status {
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return hip_check(c, e, "hipMalloc(interpolate_lvs staging)");
    tmp.push_back(*dst);
    if (upload) e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    return e == hipSuccess ? SP_OK : hip_check(c, e, "hipMemcpyAsync(interpolate_lvs)");
  };
  if (device) {
    A.bx = B.x; A.by = B.y; A.bz = B.z; A.bphi = B.phi;
    A.x = g.x; A.y = g.y; A.z = g.z; A.dx = g.dx; A.dy = g.dy; A.dz = g.dz;
    A.phi = g.phi; A.vof = g.vof; A.nx = g.nx; A.ny = g.ny; A.nz = g.nz;
  } else {
    const double* bsrc[4] = {B.x, B.y, B.z, B.phi};
    const int64_t bn[4] = {B.imax, B.jmax, B.kmax, nb};
    const double** bdst[4] = {&A.bx, &A.by, &A.bz, &A.bphi};
    const double* gsrc[6] = {g.x, g.y, g.z, g.dx, g.dy, g.dz};
    const double** gdst[6] = {&A.x, &A.y, &A.z, &A.dx, &A.dy, &A.dz};
    double* osrc[5] = {g.phi, g.vof, g.nx, g.ny, g.nz};
    double** odst[5] = {&A.phi, &A.vof, &A.nx, &A.ny, &A.nz};
    sp_status s = SP_OK;
    for (int q = 0; q < 4 && !s; ++q) {
      void* d = nullptr;
      s = stage(bsrc[q], sizeof(double) * (size_t)bn[q], true, &d);
      *bdst[q] = static_cast<const double*>(d);
    }
    for (int q = 0; q < 6 && !s; ++q) {
      void* d = nullptr;
      s = stage(gsrc[q], sizeof(double) * (size_t)N, true, &d);
      *gdst[q] = static_cast<const double*>(d);
    }
    for (int q = 0; q < 5 && !s; ++q) {
      void* d = nullptr;
      s = stage(osrc[q], sizeof(double) * (size_t)Nh, q >= 2, &d);
      *odst[q] = static_cast<double*>(d);
    }
    if (s) { release(); return s; }
  }

  int slot;
  // algorithmic bytes: interpolation reads 8 body values and 3 centres and
  // writes phi and vof per interior cell
  timer_begin(c, "lvs_interp", (double)N * (8 * 8 + 3 * 8) + (double)Nh * 16, &slot);
  hipLaunchKernelGGL(k_lvs_interp, dim3((unsigned)((Nh + 255) / 256)), dim3(256), 0, c->stream,
                     A);
  timer_end(c, slot);
  timer_begin(c, "lvs_band", (double)N * 8 * 4, &slot);
  hipLaunchKernelGGL(k_lvs_band, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, A);
  timer_end(c, slot);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !device) {
    double* odst[5] = {g.phi, g.vof, g.nx, g.ny, g.nz};
    const double* osrc[5] = {A.phi, A.vof, A.nx, A.ny, A.nz};
    for (int q = 0; q < 5 && e == hipSuccess; ++q)
      e = hipMemcpyAsync(odst[q], osrc[q], sizeof(double) * (size_t)Nh, hipMemcpyDeviceToHost,
                         c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  release();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "interpolate_lvs");
}

}  // namespace sp
