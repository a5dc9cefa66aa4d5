// The wall step that follows the velocity correction in UpdatePUVW:
//
//   BoundaryConditionVarNew (ComputePUVW.f90:418-694): for u, v, w, p and the
//     faces W,E,S,N,B,T in that order, the TypicalBC face value
//     (BoundaryFunction.f90:760-781), the wall-normal face (or, for p, the
//     top cell layer) set to it, and the ghost by ComputeGhostVarBoundary
//     (ComputePUVW.f90:696-708);
//   the mass fix of the first Dirichlet pressure face (ComputePUVW.f90:256-384);
//   VariablesInternalCellCondition (Solver.f90:467-501);
//   mres of the result (Solver.f90:236-250).
//
// Face values follow the reference's argument passing: the BC routines
// receive whole halo slices such as Vari%u(1,:,:) as assumed-shape dummies,
// so element (a,b) of a face (1-based) reads the slice's halo entry
// (a-1, b-1).  The faces of one variable depend on each other through those
// entries (a later face reads ghosts an earlier one wrote), so the six faces
// run as six launch pairs; the four variables of one face are independent
// and share a launch (blockIdx.z).  Each pair first writes every face value
// of the face, then applies them, because the top pressure layer is both
// read and overwritten by its own face.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "sp_device.h"
#include "sp_geom.h"
#include "sp_internal.h"

namespace sp {
namespace {

struct WallArgs {
  int I, J, K, SX, SY;
  double* X[4];         // u, v, w, p (halo arrays)
  double* var;          // face values: variable q, face f at var + q*vstride + voff[f]
  int64_t vstride;
  int64_t voff[6];
  int32_t fl[24];       // flags, BCu W..T, BCv, BCw, BCp
  double cst[24];
  const double *dxp, *dyp, *dzp;
  const double *ee, *ne, *te;
  const int32_t *ctp, *ctu, *ctv, *ctw;
  RedSite site;
  Scalars* S;
};

__device__ inline int plane_n1(const WallArgs& A, int f) { return f < 2 ? A.J : A.I; }
__device__ inline int plane_n2(const WallArgs& A, int f) { return f < 4 ? A.K : A.J; }

// halo index of face element (a,b) (1-based in the plane) on plane `pl`
__device__ inline int64_t plane_at(const WallArgs& A, int f, int pl, int a, int b) {
  const int64_t SX = A.SX, SXY = (int64_t)A.SX * A.SY;
  if (f < 2) return pl + SX * a + SXY * b;
  if (f < 4) return a + SX * pl + SXY * b;
  return a + SX * b + SXY * pl;
}

// boundary cell plane of face f
__device__ inline int last_plane(const WallArgs& A, int f) {
  switch (f) {
    case 1: return A.I;
    case 3: return A.J;
    case 5: return A.K;
    default: return 1;
  }
}

// TypicalBC face values of face f for the four variables.  u E, v N and w T
// read the slice one cell further in with twice the spacing
// (BoundaryFunction.f90:102, 251, 399); the other faces read the boundary
// cell's own plane with +spacing on W/S/B and -spacing on E/N/T.
__global__ void __launch_bounds__(256) k_wall_values(WallArgs A, int f) {
#pragma clang fp contract(off)
  const int q = blockIdx.z;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int n1 = plane_n1(A, f), n2 = plane_n2(A, f);
  if (a >= n1 || b >= n2) return;
  const int flag = A.fl[6 * q + f];
  const double c = A.cst[6 * q + f];
  double val = c;
  if (flag != 0) {
    int src = last_plane(A, f);
    double mult = (f & 1) ? -1.0 : 1.0;
    if ((q == 0 && f == 1) || (q == 1 && f == 3) || (q == 2 && f == 5)) {
      src -= 1;
      mult = -2.0;
    }
    double d;
    if (f < 2) d = f == 0 ? A.dxp[0] : A.dxp[A.I - 1];
    else if (f < 4) d = f == 2 ? A.dyp[0] : A.dyp[A.J - 1];
    else d = f == 4 ? A.dzp[0] : A.dzp[A.K - 1];
    const double dxyz = mult * d;
    // element (a+1, b+1) of the dummy is the halo entry (a, b) of the slice
    val = A.X[q][plane_at(A, f, src, a, b)] - c * dxyz / 2.0;
  }
  A.var[q * A.vstride + A.voff[f] + a + (int64_t)n1 * b] = val;
}

// Assignment of the wall-normal face (u W/E, v S/N, w B/T, the top layer of
// p) and the ghost (every face except u W, v S, w B).
__global__ void __launch_bounds__(256) k_wall_apply(WallArgs A, int f) {
#pragma clang fp contract(off)
  const int q = blockIdx.z;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int n1 = plane_n1(A, f), n2 = plane_n2(A, f);
  if (a >= n1 || b >= n2) return;
  const double V = A.var[q * A.vstride + A.voff[f] + a + (int64_t)n1 * b];
  double* X = A.X[q];
  const int last = last_plane(A, f);
  const bool assign = (q == 0 && f < 2) || (q == 1 && (f == 2 || f == 3)) ||
                      (q == 2 && f >= 4) || (q == 3 && f == 5);
  if (assign) X[plane_at(A, f, (f & 1) ? last : 0, a + 1, b + 1)] = V;
  const bool no_ghost = (q == 0 && f == 0) || (q == 1 && f == 2) || (q == 2 && f == 4);
  if (no_ghost) return;
  const double in = X[plane_at(A, f, last, a + 1, b + 1)];
  const int ghost = (f & 1) ? last + 1 : 0;
  X[plane_at(A, f, ghost, a + 1, b + 1)] = A.fl[6 * q + f] == 1 ? V - in : 2.0 * V - in;
}

// Mass fix on face f (ComputePUVW.f90:260-384).  The reference's north
// branch loops j over 1..Jmax with k past Kmax (reading outside its arrays);
// this is the branch's evident intent, the mirror of the south branch.
__global__ void __launch_bounds__(256) k_wall_massfix(WallArgs A, int f) {
#pragma clang fp contract(off)
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int n1 = plane_n1(A, f), n2 = plane_n2(A, f);
  if (a >= n1 || b >= n2) return;
  int i, j, k;
  if (f < 2) { i = f == 0 ? 1 : A.I; j = a + 1; k = b + 1; }
  else if (f < 4) { i = a + 1; j = f == 2 ? 1 : A.J; k = b + 1; }
  else { i = a + 1; j = b + 1; k = f == 4 ? 1 : A.K; }
  const int64_t SX = A.SX, SXY = (int64_t)A.SX * A.SY;
  const int64_t h = i + SX * j + SXY * k;
  const double dx = A.dxp[i - 1], dy = A.dyp[j - 1], dz = A.dzp[k - 1];
  const double e = A.ee[h], n = A.ne[h], t = A.te[h];
  double *u = A.X[0], *v = A.X[1], *w = A.X[2];
  switch (f) {
    case 0:
      u[h - 1] = (dx * dz * (n * v[h] - n * v[h - SX]) + dx * dy * (t * w[h] - t * w[h - SXY]) +
                  dy * dz * e * u[h]) / e / dy / dz;
      break;
    case 1:
      u[h] = (dx * dz * (-n * v[h] + n * v[h - SX]) + dx * dy * (-t * w[h] + t * w[h - SXY]) +
              dy * dz * e * u[h - 1]) / e / dy / dz;
      break;
    case 2:
      v[h - SX] = (dy * dz * (e * u[h] - e * u[h - 1]) + dx * dy * (t * w[h] - t * w[h - SXY]) +
                   dx * dz * n * v[h]) / n / dx / dz;
      break;
    case 3:
      v[h] = (-dy * dz * (e * u[h] - e * u[h - 1]) - dx * dy * (t * w[h] - t * w[h - SXY]) +
              dx * dz * n * v[h - SX]) / n / dx / dz;
      break;
    case 4:
      w[h - SXY] = (dy * dz * (e * u[h] - e * u[h - 1]) + dx * dz * (n * v[h] - n * v[h - SX]) +
                    dx * dy * t * w[h]) / t / dx / dy;
      break;
    default:
      w[h] = (-dy * dz * (e * u[h] - e * u[h - 1]) - dx * dz * (n * v[h] - n * v[h - SX]) +
              dx * dy * t * w[h - SXY]) / t / dx / dy;
      break;
  }
}

// VariablesInternalCellCondition: p, u, v, w = 0 in solid P/U/V/W cells.
__global__ void __launch_bounds__(256) k_wall_solid(WallArgs A) {
  const Col3 q = col_of(A.I, A.J, A.K);
  if (!q.in) return;
  const int64_t IJ = (int64_t)A.I * A.J, SXY = (int64_t)A.SX * A.SY;
  int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
  int64_t h = (q.a + 1) + (int64_t)A.SX * ((q.b + 1) + (int64_t)A.SY * (q.c0 + 1));
  for (int c = q.c0; c < q.c1; ++c, n += IJ, h += SXY) {
    if (A.ctp[n] == 2) A.X[3][h] = 0.0;
    if (A.ctu[n] == 2) A.X[0][h] = 0.0;
    if (A.ctv[n] == 2) A.X[1][h] = 0.0;
    if (A.ctw[n] == 2) A.X[2][h] = 0.0;
  }
}

// max |mres|/V of the final field (Solver.f90:239-247); u(0), v(:,0,:),
// w(:,:,0) are the values the wall step left there.
__global__ void __launch_bounds__(256) k_wall_div(WallArgs A) {
#pragma clang fp contract(off)
  const Col3 q = col_of(A.I, A.J, A.K);
  const int64_t SX = A.SX, SXY = (int64_t)A.SX * A.SY;
  double m[1] = {0.0};
  if (q.in) {
    const double dx = A.dxp[q.a], dy = A.dyp[q.b];
    const double *u = A.X[0], *v = A.X[1], *w = A.X[2];
    int64_t h = (q.a + 1) + SX * (q.b + 1) + SXY * (q.c0 + 1);
    for (int c = q.c0; c < q.c1; ++c, h += SXY) {
      const double dz = A.dzp[c];
      const double r = dy * dz * (A.ee[h] * u[h] - A.ee[h - 1] * u[h - 1]) +
                       dx * dz * (A.ne[h] * v[h] - A.ne[h - SX] * v[h - SX]) +
                       dx * dy * (A.te[h] * w[h] - A.te[h - SXY] * w[h - SXY]);
      double a = fabs(r) / (dx * dy * dz);
      if (!(a <= 1.0e300)) a = 1.0e300;  // NaN / Inf
      m[0] = a > m[0] ? a : m[0];
    }
  }
  block_partials<kTileThreads, 1, 1u>(A.site, m);
}

__global__ void __launch_bounds__(kFinThreads) k_wall_div_fin(WallArgs A, uint32_t nb) {
  double t[1];
  finalize_partials<kFinThreads, 1, 1u>(A.site, nb, t);
  if (flat_tid() == 0) A.S->walldiv = t[0];
}

}  // namespace

sp_status wall_step(sp_ctx* c, const sp_inputs& in, const sp_outputs& out, bool host_var) {
  const Dims& D = c->dm;
  const sp_wall_bc& W = *in.wall;
  WallArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K; A.SX = D.SX; A.SY = D.SY;
  A.X[0] = out.u; A.X[1] = out.v; A.X[2] = out.w; A.X[3] = out.p;
  const int64_t pl[6] = {(int64_t)D.J * D.K, (int64_t)D.J * D.K, (int64_t)D.I * D.K,
                         (int64_t)D.I * D.K, (int64_t)D.I * D.J, (int64_t)D.I * D.J};
  int64_t off = 0;
  for (int f = 0; f < 6; ++f) { A.voff[f] = off; off += pl[f]; }
  A.vstride = off;
  A.var = c->wall_var;
  for (int f = 0; f < 6; ++f) {
    A.fl[f] = W.flag_u[f]; A.fl[6 + f] = W.flag_v[f]; A.fl[12 + f] = W.flag_w[f];
    A.fl[18 + f] = in.bc_flag[f];
    A.cst[f] = W.const_u[f]; A.cst[6 + f] = W.const_v[f]; A.cst[12 + f] = W.const_w[f];
    A.cst[18 + f] = W.const_p[f];
  }
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp;
  A.ee = in.ee_area; A.ne = in.ne_area; A.te = in.te_area;
  A.ctp = in.cell_type_p; A.ctu = in.cell_type_u; A.ctv = in.cell_type_v; A.ctw = in.cell_type_w;
  A.S = c->S;
  A.site = red_site(c, kSiteDiv);
  const dim3 blk(64, 4, 1);
  auto plane_grid = [&](int f, int nz) {
    const int n1 = f < 2 ? D.J : D.I, n2 = f < 4 ? D.K : D.J;
    return dim3((unsigned)((n1 + 63) / 64), (unsigned)((n2 + 3) / 4), (unsigned)nz);
  };
  int slot;
  timer_begin(c, "wall_bc", 0.0, &slot);
  for (int f = 0; f < 6; ++f) {
    hipLaunchKernelGGL(k_wall_values, plane_grid(f, 4), blk, 0, c->stream, A, f);
    hipLaunchKernelGGL(k_wall_apply, plane_grid(f, 4), blk, 0, c->stream, A, f);
  }
  int fix = 0;
  while (fix < 6 && in.bc_flag[fix] != 0) ++fix;
  if (fix < 6) hipLaunchKernelGGL(k_wall_massfix, plane_grid(fix, 1), blk, 0, c->stream, A, fix);
  timer_end(c, slot);
  const double NB = (double)D.N;
  const dim3 cg = col_grid(D.I, D.J, D.K);
  const uint32_t ncol = col_blocks(D.I, D.J, D.K);
  // cell types x4 read; zeros written only in solid cells
  timer_begin(c, "wall_solid", NB * 16, &slot);
  hipLaunchKernelGGL(k_wall_solid, cg, blk, 0, c->stream, A);
  timer_end(c, slot);
  timer_begin(c, "wall_div", NB * 6 * 8, &slot);
  hipLaunchKernelGGL(k_wall_div, cg, blk, 0, c->stream, A);
  hipLaunchKernelGGL(k_wall_div_fin, dim3(1), dim3(kFinThreads), 0, c->stream, A, ncol);
  timer_end(c, slot);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_check(c, e, "wall step launch");
  for (int q = 0; q < 24; ++q) {
    double* dst = out.bc_var[q];
    if (!dst) continue;
    e = hipMemcpyAsync(dst, c->wall_var + (q / 6) * A.vstride + A.voff[q % 6],
                       sizeof(double) * pl[q % 6],
                       host_var ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(bc_var)");
  }
  return SP_OK;
}

}  // namespace sp
