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

#include "sp_geom.h"
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


// ---------------------------------------------------------------------------
// Clsvof_Scheme (Clsvof.f90:355-798): nv = 6 sub-steps, each three
// directional sweeps in the order x y z, y z x or z x y (kk mod 3 = 0, 1,
// 2), then BoundaryConditionLvsVof, Interface_Reconstruct and
// Redistance_Levelset.
//
//   k_cls_classify     Isinterface per cell (workgroups walk k planes);
//                      interface cells go to a list (one atomic per
//                      wavefront; the order does not matter), the others get
//                      nxL = 0 at the end of a sub-step.
//   k_cls_interface    grid-stride over the list: the normal
//                      (Normal_Vector_Irre or clamped differences) and
//                      Find_Distance; the last reconstruction of a sub-step
//                      also writes nxL and the centre distance.  Before a
//                      sweep it also forms the cell's geometric vof fluxes
//                      (East_Flux..Bottom_Flux) through the faces of the
//                      sweep axis where it is the upwind cell.  Keeping the
//                      heavy path (pow, Newton) out of the per-cell kernels
//                      keeps their registers, and occupancy, at the level of
//                      streaming kernels.
//   k_cls_sweep        one thread per cell: the cell's two face fluxes of
//                      vof (the upwind cell's, full/empty or from the list
//                      kernel) and of the level set (the reference's Lax-Wendroff
//                      form with its first-order boundary faces), summed in
//                      the order the reference's face loop adds them, then
//                      the sweep's update (division by 1 - dt du/dx in the
//                      first sweep, + dt/dx T1 du after it) and the clip to
//                      [vofeps, 1 - vofeps].  The three sweeps rotate over
//                      three buffers, so every sweep reads the fields as the
//                      reference's loop does.
//   k_cls_bc           one thread per boundary face cell: BCLvs and BCVof
//                      face values, phiL ghosts from the solid normal, vofL
//                      ghosts copied (into all three buffers).
//   k_cls_redist_scatter  grid-stride over (listed cell, x and y offsets):
//                      the distance candidates of the 9^3 neighbourhood, kept
//                      by an atomic minimum on their bit patterns (a
//                      non-negative double orders as its bits), which is the
//                      reference's sequential minimum whatever the order.
//   k_cls_redist_final one thread per cell: interface cells take their own
//                      plane distance, reached cells the minimum (at most
//                      1e4), the others +-5 sqrt(3) dx; the minimum is
//                      emptied again for the next sub-step.
// ---------------------------------------------------------------------------
namespace {

constexpr unsigned long long kAuxEmpty = 0x7FF0000000000000ull;  // +inf: not reached

struct ClsArgs {
  int I, J, K;
  const double *u, *v, *w;
  const double *x, *y, *z, *dx, *dy, *dz;
  const double *snx, *sny, *snz;
  const double* bcv[3];        // BCu%VarW, BCv%VarS, BCw%VarB
  int lvs_flag[6], vof_flag[6];
  double lvs_const[6], vof_const[6];
  double dtv;
  double *nx, *ny, *nz, *dis;  // reconstruction (cell arrays)
  double *flo, *fhi;           // a mixed cell's vof flux through its lower / upper face
  unsigned long long* aux;     // redistancing minimum (cell array)
  double *nxl, *nyl, *nzl;     // PCell%nxL, nyL, nzL (halo)
  double* lvs_var[6];
  double* vof_var[6];
  double* vof_buf[3];          // A, B, C vof buffers (ghosts kept equal)
};

// Isinterface for every cell: interface cells are appended to a list for
// k_cls_interface (one global atomic per wavefront); the last
// reconstruction of a sub-step writes nxL = 0 for the others.  A workgroup
// walks `kper` k planes of a 64 x 4 column, keeping the k-1, k, k+1 values
// of its cell in registers.
__global__ void __launch_bounds__(256) k_cls_classify(ClsArgs A, const double* __restrict__ vfl,
                                                       int final_, int kper,
                                                       int* __restrict__ list,
                                                       int* __restrict__ count) {
#pragma clang fp contract(off)
  const int I = A.I, J = A.J, K = A.K;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k0 = blockIdx.z * kper + 1;
  const int k1 = k0 + kper - 1 < K ? k0 + kper - 1 : K;
  const bool in = i <= I && j <= J;
  const int ic = in ? i : 1, jc = in ? j : 1;  // lanes past the edge read a valid cell
  const int64_t sk = (int64_t)(I + 2) * (J + 2);
  int64_t h = hidx(I, J, ic, jc, k0);
  double vm = vfl[h - sk], vc = vfl[h];
  const double lim = 1.0 - 2.1 * kVofEps;
  for (int k = k0; k <= k1; ++k, h += sk) {
    const double vp = vfl[h + sk];
    bool f = vc > kVofEps && vc < 1.0 - kVofEps;
    f = f || fabs(vfl[h - 1] - vc) >= lim;
    f = f || fabs(vfl[h + 1] - vc) >= lim;
    f = f || fabs(vfl[h - (I + 2)] - vc) >= lim;
    f = f || fabs(vfl[h + (I + 2)] - vc) >= lim;
    f = f || fabs(vm - vc) >= lim;
    f = f || fabs(vp - vc) >= lim;
    f = f && in;
    // one atomic per wavefront: the lanes of interface cells take
    // consecutive list slots after the wave's base
    const unsigned long long m = __ballot(f);
    if (m) {
      const int lane = __lane_id();
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(count, __popcll(m));
      base = __shfl(base, leader);
      if (f) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)cidx(I, J, i, j, k);
    }
    if (final_ && in && !f) {
      A.nxl[h] = 0.0;
      A.nyl[h] = 0.0;
      A.nzl[h] = 0.0;
    }
    vm = vc;
    vc = vp;
  }
}

// Interface_Reconstruct of the listed cells (grid-stride): the normal
// (Normal_Vector_Irre or clamped differences), normalised, and Find_Distance;
// the last reconstruction of a sub-step also writes nxL and the centre
// distance the redistancing uses.
__global__ void __launch_bounds__(256) k_cls_interface(ClsArgs A, const double* __restrict__ vfl,
                                                        const double* __restrict__ phi,
                                                        int final_, int ax,
                                                        const int* __restrict__ list,
                                                        const int* __restrict__ count) {
#pragma clang fp contract(off)
  const int I = A.I, J = A.J, K = A.K;
  const int n = *count;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int64_t t = list[q];
    const int i = (int)(t % I) + 1;
    const int64_t r = t / I;
    const int j = (int)(r % J) + 1, k = (int)(r / J) + 1;
    const int64_t h = hidx(I, J, i, j, k);
    double a, b, e;
    if (i > 1 && i < I && j > 1 && j < J && k > 1 && k < K) {
      normal_vector_irre(I, J, K, A.x, A.y, A.z, phi, i, j, k, a, b, e);
    } else {
      const int ip = i + 1 < I ? i + 1 : I, im = i - 1 > 1 ? i - 1 : 1;
      const int jp = j + 1 < J ? j + 1 : J, jm = j - 1 > 1 ? j - 1 : 1;
      const int kp = k + 1 < K ? k + 1 : K, km = k - 1 > 1 ? k - 1 : 1;
      a = (phi[hidx(I, J, ip, j, k)] - phi[hidx(I, J, im, j, k)]) /
          (A.x[cidx(I, J, ip, j, k)] - A.x[cidx(I, J, im, j, k)]);
      b = (phi[hidx(I, J, i, jp, k)] - phi[hidx(I, J, i, jm, k)]) /
          (A.y[cidx(I, J, i, jp, k)] - A.y[cidx(I, J, i, jm, k)]);
      e = (phi[hidx(I, J, i, j, kp)] - phi[hidx(I, J, i, j, km)]) /
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
    double d = find_distance(A.dx[t], A.dy[t], A.dz[t], a, b, e, vfl[h]);
    A.nx[t] = a;
    A.ny[t] = b;
    A.nz[t] = e;
    if (final_) {
      A.nxl[h] = a;
      A.nyl[h] = b;
      A.nzl[h] = e;
      d = (0.5 * A.dx[t] * fabs(a) + 0.5 * A.dy[t] * fabs(b) + 0.5 * A.dz[t] * fabs(e)) - d;
    }
    A.dis[t] = d;
    // before a sweep along ax: the fluxes through this cell's two faces of
    // that axis where it is the upwind cell and neither full nor empty (what
    // the sweep would compute, so the sweep kernel itself stays light)
    const double v = vfl[h];
    if (ax >= 0 && !(v >= 1.0 - kVofEps || v <= kVofEps)) {
      const int64_t st = ax == 0 ? 1 : (ax == 1 ? (int64_t)(I + 2) : (int64_t)(I + 2) * (J + 2));
      const double* vel = ax == 0 ? A.u : (ax == 1 ? A.v : A.w);
      const double uh = vel[h], ul = vel[h - st];
      const bool pos_h = ax == 2 ? (uh >= 0.0) : (uh > 0.0);
      const bool pos_l = ax == 2 ? (ul >= 0.0) : (ul > 0.0);
      if (pos_h)
        A.fhi[t] = face_flux(ax, true, a, b, e, d, v, A.dx[t], A.dy[t], A.dz[t], uh * A.dtv);
      if (!pos_l)
        A.flo[t] = face_flux(ax, false, a, b, e, d, v, A.dx[t], A.dy[t], A.dz[t], -ul * A.dtv);
    }
  }
}

__global__ void __launch_bounds__(256) k_cls_sweep(ClsArgs A, int ax, int s,
                                                    const double* __restrict__ vin,
                                                    const double* __restrict__ pin,
                                                    double* __restrict__ vout,
                                                    double* __restrict__ pout,
                                                    const double* __restrict__ t1v,
                                                    const double* __restrict__ t1p,
                                                    const double* __restrict__ bcl) {
#pragma clang fp contract(off)
  const int I = A.I, J = A.J, K = A.K;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > I || j > J) return;
  const int64_t t = cidx(I, J, i, j, k);
  const int n = ax == 0 ? I : (ax == 1 ? J : K);
  const int p = ax == 0 ? i : (ax == 1 ? j : k);
  const int64_t st = ax == 0 ? 1 : (ax == 1 ? (int64_t)(I + 2) : (int64_t)(I + 2) * (J + 2));
  const int64_t cs = ax == 0 ? 1 : (ax == 1 ? (int64_t)I : (int64_t)I * J);
  const int64_t h = hidx(I, J, i, j, k);
  const int64_t h0 = h - (int64_t)p * st, c0 = t - (int64_t)p * cs;
  const double* vel = ax == 0 ? A.u : (ax == 1 ? A.v : A.w);
  const double* dd = ax == 0 ? A.dx : (ax == 1 ? A.dy : A.dz);
  const double* xx = ax == 0 ? A.x : (ax == 1 ? A.y : A.z);
  const double dtv = A.dtv;
  auto V = [&](int q) { return vin[h0 + (int64_t)q * st]; };
  auto PH = [&](int q) { return pin[h0 + (int64_t)q * st]; };
  auto U = [&](int q) { return vel[h0 + (int64_t)q * st]; };
  auto D = [&](int q) { return dd[c0 + (int64_t)q * cs]; };
  auto XC = [&](int q) { return xx[c0 + (int64_t)q * cs]; };
  // vof flux through face f (between positions f and f+1), 1 <= f <= n-1
  auto vof_face = [&](int f) {
    const double ue = U(f);
    const bool pos = ax == 2 ? (ue >= 0.0) : (ue > 0.0);
    const int q = pos ? f : f + 1;
    const double vq = V(q);
    if (vq >= 1.0 - kVofEps || vq <= kVofEps) return vq * ue * dtv;
    const int64_t cq = c0 + (int64_t)q * cs;
    return pos ? A.fhi[cq] : -A.flo[cq];
  };
  // level-set flux through face f, 2 <= f <= n-1
  auto ls_face = [&](int f) {
    const double ue = U(f);
    double ls;
    if (ue >= 0.0)
      ls = PH(f) + D(f) / 2.0 * (1.0 - ue * dtv / D(f)) * (PH(f + 1) - PH(f - 1)) / (2.0 * D(f));
    else if (f <= n - 2)
      ls = PH(f + 1) - D(f) / 2.0 * (1.0 + ue * dtv / D(f)) * (PH(f + 2) - PH(f)) / (2.0 * D(f));
    else
      ls = PH(f + 1) - D(f) / 2.0 * (1.0 + ue * dtv / D(f)) * (PH(f + 1) - PH(f)) / D(f);
    return ue * ls * dtv;
  };
  // face 1, reduced to first order (the forms differ per axis)
  auto ls_face1 = [&]() {
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
    return ls * u1 * dtv;
  };
  // face n (the z sweep divides by dy(kmax-1) on its else branch)
  auto ls_facen = [&]() {
    const double un = U(n);
    double ls;
    if (un > 0.0) {
      ls = PH(n) + D(n) / 2.0 * (1.0 - un * dtv / D(n)) * (PH(n) - PH(n - 1)) / (XC(n) - XC(n - 1));
    } else {
      const double dq = ax == 2 ? A.dy[c0 + (int64_t)(n - 1) * cs] : D(n - 1);
      ls = PH(n) - D(n - 1) / 2.0 * (1.0 + un * dtv / dq) * (PH(n) - PH(n - 1)) / (XC(n) - XC(n - 1));
    }
    return ls * un * dtv;
  };

  double T = V(p), L = PH(p);
  if (p == 1) {
    T = T - vof_face(1) / D(1);
    T = T + V(1) * U(0) * dtv / D(1);
  } else if (p < n) {
    T = T + vof_face(p - 1) / D(p);
    T = T - vof_face(p) / D(p);
  } else {
    T = T + vof_face(n - 1) / D(n);
    T = T - V(n) * U(n) * dtv / D(n);
  }
  if (p == 1) {
    L = L - ls_face1() / D(1);
    const int64_t fb = ax == 0 ? (int64_t)(j - 1) + (int64_t)J * (k - 1)
                               : (ax == 1 ? (int64_t)(i - 1) + (int64_t)I * (k - 1)
                                          : (int64_t)(i - 1) + (int64_t)I * (j - 1));
    L = L + A.bcv[ax][fb] * bcl[fb] * dtv / D(1);
  } else if (p == 2) {
    L = L - ls_face(2) / D(2);
    L = L + ls_face1() / D(2);
  } else if (p < n) {
    L = L + ls_face(p - 1) / D(p);
    L = L - ls_face(p) / D(p);
  } else {
    L = L + ls_face(n - 1) / D(n);
    L = L - ls_facen() / D(n);
  }
  const double du = U(p) - U(p - 1);
  const double q = dtv / D(p);
  double Tn, Ln;
  if (s == 0) {
    Tn = T / (1.0 - q * du);
    Ln = L / (1.0 - q * du);
  } else {
    Tn = T + q * t1v[h] * du;
    Ln = L + q * t1p[h] * du;
  }
  if (Tn <= kVofEps) Tn = 0.0;
  if (Tn >= 1.0 - kVofEps) Tn = 1.0;
  vout[h] = Tn;
  pout[h] = Ln;
}

__global__ void __launch_bounds__(256) k_cls_bc(ClsArgs A, double* __restrict__ phi) {
#pragma clang fp contract(off)
  const int I = A.I, J = A.J, K = A.K;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf[3] = {(int64_t)J * K, (int64_t)I * K, (int64_t)I * J};
  int f = -1;
  for (int q = 0; q < 6; ++q) {
    if (t < nf[q / 2]) { f = q; break; }
    t -= nf[q / 2];
  }
  if (f < 0) return;
  const int ax = f / 2, hi = f % 2;
  const int na = ax == 0 ? J : I;
  const int a = (int)(t % na) + 1, b = (int)(t / na) + 1;
  int i, j, k;
  if (ax == 0) { i = hi ? I : 1; j = a; k = b; }
  else if (ax == 1) { i = a; j = hi ? J : 1; k = b; }
  else { i = a; j = b; k = hi ? K : 1; }
  const int64_t h = hidx(I, J, i, j, k), c = cidx(I, J, i, j, k);
  const double vl = phi[h], vv = A.vof_buf[0][h];
  const double d = ax == 0 ? A.dx[c] : (ax == 1 ? A.dy[c] : A.dz[c]);
  const double sd = hi ? -d : d;
  double lv = A.lvs_flag[f] == 0 ? A.lvs_const[f] : vl - A.lvs_const[f] * sd / 2.0;
  if (ax == 0) lv = vl - A.nxl[h] * sd / 2.0;  // West and East pass nxL
  A.lvs_var[f][t] = lv;
  if (A.vof_var[f])
    A.vof_var[f][t] = A.vof_flag[f] == 0 ? A.vof_const[f] : vv - A.vof_const[f] * sd / 2.0;
  // ghosts
  const int64_t step = ax == 0 ? 1 : (ax == 1 ? (int64_t)(I + 2) : (int64_t)(I + 2) * (J + 2));
  const int64_t g = hi ? h + step : h - step;
  double gp;
  if (ax == 0) gp = hi ? vl + A.snx[h] * A.dx[c] : vl - A.snx[h] * A.dx[c];
  else if (ax == 1) gp = vl - A.sny[h] * A.dy[c];  // North subtracts too (:2136-2137)
  else gp = hi ? vl + A.snz[h] * A.dz[c] : vl - A.snz[h] * A.dz[c];
  phi[g] = gp;
  for (int q = 0; q < 3; ++q) A.vof_buf[q][g] = vv;
}

// Redistancing candidates from the listed interface cells (grid-stride over
// (cell, x offset, y offset)): the 9 targets of one (x, y) offset each
__global__ void __launch_bounds__(256) k_cls_redist_scatter(ClsArgs A,
                                                             const double* __restrict__ vfl,
                                                             const double* __restrict__ phi,
                                                             const int* __restrict__ list,
                                                             const int* __restrict__ count) {
#pragma clang fp contract(off)
  const int I = A.I, J = A.J, K = A.K;
  const int64_t n81 = 81 * (int64_t)*count;
  const double epsi = 1.0e-15;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n81;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = list[q / 81];
    const int ii = (int)(q % 81) / 9 - 4, jj = (int)(q % 9) - 4;
    const int i = (int)(c % I) + 1;
    const int64_t r = c / I;
    const int j = (int)(r % J) + 1, k = (int)(r / J) + 1;
    const double vc = vfl[hidx(I, J, i, j, k)];
    if (!((0.0 + epsi) < vc && vc < (1.0 - epsi))) continue;
    const int a = i + ii;
    if (a < 1 || a > I) continue;
    const int b = j + jj;
    if (b < 1 || b > J) continue;
    const double nxx = A.nx[c], nyy = A.ny[c], nzz = A.nz[c], diss = A.dis[c];
    const double ddx = A.dx[c], ddy = A.dy[c], ddz = A.dz[c];
    for (int kk = -4; kk <= 4; ++kk) {
      const int e = k + kk;
      if (e < 1 || e > K) continue;
      const int64_t ht = hidx(I, J, a, b, e);
      const double vt = vfl[ht];
      if (!(vt < 0.0 + epsi || vt > 1.0 - epsi)) continue;
      const double deuc =
          redist_candidate(nxx, nyy, nzz, diss, ddx, ddy, ddz, ii, jj, kk, phi[ht]);
      atomicMin(&A.aux[cidx(I, J, a, b, e)], (unsigned long long)__double_as_longlong(deuc));
    }
  }
}

__global__ void __launch_bounds__(256) k_cls_redist_final(ClsArgs A, const double* __restrict__ vfl,
                                                           double* __restrict__ phi) {
#pragma clang fp contract(off)
  const int I = A.I, J = A.J;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > I || j > J) return;
  const int64_t t = cidx(I, J, i, j, k);
  const int64_t h = hidx(I, J, i, j, k);
  const double epsi = 1.0e-15;
  const double vc = vfl[h];
  const double sg = copysign(1.0, 0.5 - vc);
  double out;
  if ((0.0 + epsi) < vc && vc < (1.0 - epsi)) {
    out = sg * fabs(A.dis[t]);
  } else if (A.aux[t] != kAuxEmpty) {
    const double m = __longlong_as_double((long long)A.aux[t]);
    out = sg * (m < 1.0e4 ? m : 1.0e4);
    A.aux[t] = kAuxEmpty;  // empty again for the next sub-step
  } else {
    out = 5.0 * A.dx[t] * sqrt(3.0) * sg;
  }
  phi[h] = out;
}

__global__ void __launch_bounds__(256) k_cls_fill_aux(unsigned long long* __restrict__ aux,
                                                       int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) aux[t] = kAuxEmpty;
}

// the classification launch: about 2048 workgroups of 64 x 4 columns, each
// walking kper k planes
Geom classify_geom(const Dims& D) { return make_geom(D.I, D.J, D.K, 2048); }

constexpr int kClsCounters = 32;  // interface-list counters, one per reconstruction of a call

}  // namespace

// B and C buffers (vof, level set), the reconstruction (4 N), the
// redistancing minimum (N), the face fluxes (2 N), the interface list (N
// int32) and its counters
size_t clsvof_work_doubles(const Dims& D) {
  return 4 * (size_t)D.Nh + 7 * (size_t)D.N + ((size_t)D.N + 1) / 2 + kClsCounters / 2;
}

sp_status clsvof_scheme(sp_ctx* c, const sp_clsvof_inputs& in, const sp_clsvof_state& st,
                        double* work) {
  const Dims& D = c->dm;
  const int I = D.I, J = D.J, K = D.K;
  const int64_t N = D.N, Nh = D.Nh;
  ClsArgs A{};
  A.I = I; A.J = J; A.K = K;
  A.u = in.u; A.v = in.v; A.w = in.w;
  A.x = in.x; A.y = in.y; A.z = in.z; A.dx = in.dx; A.dy = in.dy; A.dz = in.dz;
  A.snx = in.nx; A.sny = in.ny; A.snz = in.nz;
  A.bcv[0] = in.bcu_w; A.bcv[1] = in.bcv_s; A.bcv[2] = in.bcw_b;
  for (int q = 0; q < 6; ++q) {
    A.lvs_flag[q] = in.lvs_flag[q];
    A.lvs_const[q] = in.lvs_const[q];
    A.vof_flag[q] = in.vof_flag[q];
    A.vof_const[q] = in.vof_const[q];
    A.lvs_var[q] = st.lvs_var[q];
    A.vof_var[q] = st.vof_var[q];
  }
  A.dtv = in.dt / 6.0;
  A.nxl = st.nx_l; A.nyl = st.ny_l; A.nzl = st.nz_l;
  double* vb[3] = {st.vof_l, work, work + Nh};
  double* pb[3] = {st.phi_l, work + 2 * Nh, work + 3 * Nh};
  double* w = work + 4 * Nh;
  A.nx = w; A.ny = w + N; A.nz = w + 2 * N; A.dis = w + 3 * N;
  A.aux = reinterpret_cast<unsigned long long*>(w + 4 * N);
  A.flo = w + 5 * N; A.fhi = w + 6 * N;
  for (int q = 0; q < 3; ++q) A.vof_buf[q] = vb[q];
  // the B and C buffers start with A's ghosts (the sweeps write interiors)
  hipError_t e = hipMemcpyAsync(vb[1], vb[0], sizeof(double) * (size_t)Nh,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(vb[2], vb[0], sizeof(double) * (size_t)Nh, hipMemcpyDeviceToDevice,
                       c->stream);
  for (int q = 1; q < 3 && e == hipSuccess; ++q)
    e = hipMemcpyAsync(pb[q], pb[0], sizeof(double) * (size_t)Nh, hipMemcpyDeviceToDevice,
                       c->stream);
  if (e != hipSuccess) return hip_check(c, e, "clsvof buffers");
  const Geom gK = classify_geom(D);
  int* list = reinterpret_cast<int*>(w + 7 * N);
  int* counters = list + 2 * (((size_t)N + 1) / 2);
  if (e == hipSuccess)
    e = hipMemsetAsync(counters, 0, sizeof(int) * kClsCounters, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "clsvof counters");
  // per-cell kernels: 64 x 4 threads per workgroup over (i, j), one k plane
  // per grid row
  const dim3 bC(64, 4, 1), gC((unsigned)((I + 63) / 64), (unsigned)((J + 3) / 4), (unsigned)K);
  const unsigned gN = (unsigned)((N + 255) / 256);
  // grid-stride kernels over the interface list: enough workgroups for a few
  // per CU
  const unsigned gL = (unsigned)(c->num_cu > 0 ? 8 * c->num_cu : 2048);
  const int64_t nbc = 2 * ((int64_t)J * K + (int64_t)I * K + (int64_t)I * J);
  int slot;
  hipLaunchKernelGGL(k_cls_fill_aux, dim3(gN), dim3(256), 0, c->stream, A.aux, N);
  int ctr = 0;
  auto reconstruct = [&](const double* v, const double* p, int final_, int ax) {
    int* cnt = counters + ctr++;
    // algorithmic bytes per cell: vof read (and nxL, nyL, nzL written at the
    // end of a sub-step)
    timer_begin(c, final_ ? "cls_classify_final" : "cls_classify", (double)N * (final_ ? 32 : 8),
                &slot);
    hipLaunchKernelGGL(k_cls_classify, gK.grid, gK.block, 0, c->stream, A, v, final_, gK.kper,
                       list, cnt);
    timer_end(c, slot);
    timer_begin(c, "cls_interface", 0.0, &slot);
    hipLaunchKernelGGL(k_cls_interface, dim3(gL), dim3(256), 0, c->stream, A, v, p, final_, ax,
                       list, cnt);
    timer_end(c, slot);
    return cnt;
  };
  for (int kk = 1; kk <= 6; ++kk) {
    const int rr = kk % 3;
    for (int s = 0; s < 3; ++s) {
      const int ax = (rr + s) % 3;
      const int src = s == 0 ? 0 : s, dst = s == 2 ? 0 : s + 1;  // A -> B -> C -> A
      reconstruct(vb[src], pb[src], 0, ax);
      // vof, level set, face velocity, width, reconstruction (4) read, the
      // first sweep's vof and level set read after it, vof and level set
      // written
      timer_begin(c, "cls_sweep", (double)N * (s == 0 ? 88 : 104), &slot);
      hipLaunchKernelGGL(k_cls_sweep, gC, bC, 0, c->stream, A, ax, s, vb[src],
                         pb[src], vb[dst], pb[dst], vb[1], pb[1], st.lvs_var[2 * ax]);
      timer_end(c, slot);
    }
    timer_begin(c, "cls_bc", (double)nbc * 8 * 8, &slot);
    hipLaunchKernelGGL(k_cls_bc, dim3((unsigned)((nbc + 255) / 256)), dim3(256), 0, c->stream, A,
                       pb[0]);
    timer_end(c, slot);
    const int* cnt = reconstruct(vb[0], pb[0], 1, -1);
    timer_begin(c, "cls_redist_scatter", 0.0, &slot);
    hipLaunchKernelGGL(k_cls_redist_scatter, dim3(gL), dim3(256), 0, c->stream, A, vb[0], pb[0],
                       list, cnt);
    timer_end(c, slot);
    // vof, widths, centre distance, minimum read; level set written
    timer_begin(c, "cls_redist_final", (double)N * 48, &slot);
    hipLaunchKernelGGL(k_cls_redist_final, gC, bC, 0, c->stream, A, vb[0], pb[0]);
    timer_end(c, slot);
  }
  e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "clsvof kernels");
}

}  // namespace sp
