// NewCellFace (CutCell.f90:788-881), the last step of the geometry update
// after a solid moves (SolidBody.f90:824, Main.f90:218):
//
//   the U/V/W cell centres from the P face centres (:798-822);
//   CellLinking (:1269-1355) of the momentum-exchange cells, which moves a
//     tiny cut cell's centre next to its master neighbour's;
//   delh, the distance of each centre to the solid surface (:844-858);
//   East/North/TopFaceInterpolationInf (:983-1267) for the three grids.
//
// CellLinking runs in the reference in loop order (i, j, k, k innermost),
// so a link to the -1 neighbour along an axis reads that neighbour's
// already linked centre and a link to the +1 neighbour its centre before
// linking.  Here every linked cell follows its chain of -1 links back to
// the first centre that is not relinked before it, then applies the links
// forward, reading only the centres as they were before linking (a copy):
// the same values in the same order as the sequential loop.

#include <hip/hip_runtime.h>

#include "sp_device.h"
#include "sp_internal.h"

namespace sp {
namespace {

constexpr double kEpsiI = 1.0e-3;  // CutCell.f90:11

struct NcfDims {
  int I, J, K, SX, SY;
  int64_t N, Nh;
};

__device__ inline int64_t hxi(const NcfDims& D, int i, int j, int k) {
  return (int64_t)i + (int64_t)D.SX * ((int64_t)j + (int64_t)D.SY * (int64_t)k);
}
__device__ inline int64_t cxi(const NcfDims& D, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)D.I * ((int64_t)(j - 1) + (int64_t)D.J * (int64_t)(k - 1));
}

struct LinkGrid {
  const double *d[3], *n[3];
  const int32_t* moex;
  const double* orig;  // centres before linking (3 components)
  double* cent;
  int32_t* msce;
};

// :798-822 for U (m = 0), V (1), W (2) at once
__global__ void __launch_bounds__(256) k_ncf_cent(NcfDims D, const double* pee, const double* pne,
                                                  const double* pte, const double* pfce,
                                                  const double* pfcn, const double* pfct,
                                                  double* cu, double* cv, double* cw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  const int64_t q = cxi(D, i, j, k), h = hxi(D, i, j, k), N = D.N, Nh = D.Nh;
  if (pee[h] > 0.0) { cu[q] = 0.0; cu[N + q] = pfce[Nh + h]; cu[2 * N + q] = pfce[2 * Nh + h]; }
  if (pne[h] > 0.0) { cv[q] = pfcn[h]; cv[N + q] = 0.0; cv[2 * N + q] = pfcn[2 * Nh + h]; }
  if (pte[h] > 0.0) { cw[q] = pfct[h]; cw[N + q] = pfct[Nh + h]; cw[2 * N + q] = 0.0; }
}

// The link of cell (i,j,k) as CellLinking would apply it: the axis of the
// dominant normal component (x, y, z tested in that order) and the master
// index along it; false when the cell is not linked (outside the loop
// range 1..Imax-1 x 1..Jmax-1 x 1..Kmax-1, MoExCell /= 1, no axis within
// 1e-10 of the largest |n|, or a master at index 0, outside the grid).
__device__ inline bool link_of(const NcfDims& D, const LinkGrid& G, const int c[3], int* ax,
                               int* dir) {
  if (c[0] > D.I - 1 || c[1] > D.J - 1 || c[2] > D.K - 1) return false;
  const int64_t q = cxi(D, c[0], c[1], c[2]);
  if (G.moex[q] != 1) return false;
  const double n0 = G.n[0][q], n1 = G.n[1][q], n2 = G.n[2][q];
  double nmax = fabs(n0);
  if (fabs(n1) > nmax) nmax = fabs(n1);
  if (fabs(n2) > nmax) nmax = fabs(n2);
  int a;
  double na;
  if (fabs(fabs(n0) - nmax) < 1.0e-10) { a = 0; na = n0; }
  else if (fabs(fabs(n1) - nmax) < 1.0e-10) { a = 1; na = n1; }
  else if (fabs(fabs(n2) - nmax) < 1.0e-10) { a = 2; na = n2; }
  else return false;
  const int s = na < 0.0 ? -1 : 1;
  if (c[a] + s < 1) return false;
  *ax = a;
  *dir = s;
  return true;
}

// CellLinking with the reference's sequential semantics, one thread per
// linked cell of one grid.
__global__ void __launch_bounds__(256) k_ncf_link(NcfDims D, LinkGrid G) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  const int c0[3] = {i, j, k};
  int ax, dir;
  if (!link_of(D, G, c0, &ax, &dir)) return;
  const double tol = 1.0e-8;
  const int64_t N = D.N;
  // chain length: follow -1 links while the master is itself linked (it was
  // relinked earlier in the loop); each step lowers one index, so the walk
  // ends within Imax + Jmax + Kmax steps
  int n = 0;
  {
    int c[3] = {i, j, k}, a = ax, s = dir;
    const int cap = D.I + D.J + D.K;
    while (s < 0 && n < cap) {
      int m[3] = {c[0], c[1], c[2]};
      m[a] += s;
      int a2, s2;
      if (!link_of(D, G, m, &a2, &s2)) break;
      c[0] = m[0]; c[1] = m[1]; c[2] = m[2];
      a = a2; s = s2;
      ++n;
    }
  }
  // apply the links from the end of the chain back to this cell
  double v[3];
  for (int t = n; t >= 0; --t) {
    int c[3] = {i, j, k}, a = ax, s = dir;
    for (int u = 0; u < t; ++u) {  // cell t of the chain
      c[a] += s;
      link_of(D, G, c, &a, &s);
    }
    int m[3] = {c[0], c[1], c[2]};
    m[a] += s;
    const int64_t qm = cxi(D, m[0], m[1], m[2]);
    if (t == n) {
      v[0] = G.orig[qm]; v[1] = G.orig[N + qm]; v[2] = G.orig[2 * N + qm];
    }
    const double dm = G.d[a][qm];
    if (s < 0) v[a] = v[a] - dm + tol * dm;
    else v[a] = v[a] + dm - tol * dm;
  }
  const int64_t q = cxi(D, i, j, k);
  G.cent[q] = v[0];
  G.cent[N + q] = v[1];
  G.cent[2 * N + q] = v[2];
  int mm[3] = {i, j, k};
  mm[ax] += dir;
  G.msce[q] = mm[0];
  G.msce[N + q] = mm[1];
  G.msce[2 * N + q] = mm[2];
}

// :844-858
__global__ void __launch_bounds__(256) k_ncf_delh(int64_t N, const double* cent, const double* nx,
                                                  const double* ny, const double* nz,
                                                  const double* phi, double* delh) {
#pragma clang fp contract(off)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  delh[q] = fabs(cent[q] * nx[q] + cent[N + q] * ny[q] + cent[2 * N + q] * nz[q] + phi[q]);
}

struct InterpArgs {
  const double *d[3], *n[3], *phi, *cent, *fc, *area, *bd;
  const int32_t* ct;
  double *S, *Eta, *Al, *DAl;
  int ax, ib;
};

// Face*InterpolationInf for face direction A.ax of one grid at face
// (i,j,k), index < Imax (Jmax, Kmax) along ax.
__device__ inline void interp_face(const NcfDims& D, const InterpArgs& A, int i, int j, int k) {
#pragma clang fp contract(off)
  const int ax = A.ax;
  const int idx[3] = {i, j, k};
  const int L = ax == 0 ? D.I : (ax == 1 ? D.J : D.K);
  if (idx[ax] > L - 1) return;
  const int64_t N = D.N, Nh = D.Nh;
  const int64_t step = ax == 0 ? 1 : (ax == 1 ? (int64_t)D.I : (int64_t)D.I * D.J);
  const int64_t q = cxi(D, i, j, k), qn = q + step, h = hxi(D, i, j, k);
  const double dB = A.bd[q + (A.ib ? step : 0)];
  const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
  if (A.ct[q] == 2 || A.ct[qn] == 2) {
    A.Al[h] = 1.0;
    A.DAl[h] = 0.0;
    A.S[h] = dB;
    A.Eta[h] = 0.5;
    return;
  }
  const double* C = A.cent;
  const double ca = C[ax * N + q], can = C[ax * N + qn];
  const double S = can + dB - ca;
  A.S[h] = S;
  const double s1 = C[o1 * N + qn] - C[o1 * N + q];
  const double s2 = C[o2 * N + qn] - C[o2 * N + q];
  double eta = fabs((A.fc[ax * Nh + h] - ca) / S);
  if (eta > 1.0) eta = 0.5;
  A.Eta[h] = eta;
  double f[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double a = C[m * N + q], b = C[m * N + qn];
    f[m] = m == ax ? (1.0 - eta) * a + eta * (b + dB) : (1.0 - eta) * a + eta * b;
  }
  const double t1 = 1.0e-4 * A.d[o1][q], t2 = 1.0e-4 * A.d[o2][q];
  const bool tr = ax == 0 ? (fabs(s1) > t1 || fabs(s2) > t2)     // :1013 tests >
                          : (fabs(s1) >= t1 || fabs(s2) >= t2);  // :1107, :1202 test >=
  if (!tr) {
    A.DAl[h] = 0.0;
    A.Al[h] = 1.0;
    return;
  }
  const double nf1 = 0.5 * (A.n[o1][q] + A.n[o1][qn]);
  const double nf2 = 0.5 * (A.n[o2][q] + A.n[o2][qn]);
  const double nx0 = A.n[0][q], ny0 = A.n[1][q], nz0 = A.n[2][q], ph0 = A.phi[q];
  const double nx1 = A.n[0][qn], ny1 = A.n[1][qn], nz1 = A.n[2][qn], ph1 = A.phi[qn];
  const double delh1 = f[0] * nx0 + f[1] * ny0 + f[2] * nz0 + ph0;
  double g[3] = {f[0], f[1], f[2]};
  g[ax] = f[ax] - dB;
  const double delh2 = g[0] * nx1 + g[1] * ny1 + g[2] * nz1 + ph1;
  double delh = 0.5 * (delh1 + delh2);
  const double F1 = A.fc[h], F2 = A.fc[Nh + h], F3 = A.fc[2 * Nh + h];
  double e1, e2;
  if (ax == 0) {         // :1032-1037
    e1 = F1 * nx0 + F2 * ny0 + F3 * nz0 + ph0;
    e2 = (F1 - dB) * nx1 + F2 * ny1 + F3 * nz1 + ph1;
  } else if (ax == 1) {  // :1127-1132 subtract BGrid%dy(i,j+1,k) whatever iv is
    const double dB1 = A.bd[qn];
    e1 = F1 * nx0 + F2 * ny0 + nz0 * F3 + ph0;
    e2 = F1 * nx1 + (F2 - dB1) * ny1 + nz1 * F3 + ph1;
  } else {               // :1222-1227
    e1 = F1 * nx0 + F2 * ny0 + nz0 * F3 + ph0;
    e2 = F1 * nx1 + F2 * ny1 + nz1 * (F3 - dB) + ph1;
  }
  const double delhec = 0.5 * (e1 + e2);
  if (delh < kEpsiI && A.area[h] > 0.0) delh = delhec;
  A.Al[h] = fabs(delhec) / fabs(delh);
  A.DAl[h] = (s1 * nf1 + s2 * nf2) / (S * fabs(delh));
}

// The three face directions of one grid in one launch: a thread handles
// the E, N and T faces of its cell, so the cell's own centre, normal,
// level set and type are read once for the three.
__global__ void __launch_bounds__(256) k_ncf_interp(NcfDims D, InterpArgs AE, InterpArgs AN,
                                                    InterpArgs AT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  interp_face(D, AE, i, j, k);
  interp_face(D, AN, i, j, k);
  interp_face(D, AT, i, j, k);
}

// The entry at index Imax (Jmax, Kmax) along ax copies the one before it
// (:1067-1070), over the interior of the other two axes.
__global__ void __launch_bounds__(256) k_ncf_last(NcfDims D, InterpArgs A) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int b = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int ax = A.ax;
  const int na = ax == 0 ? D.J : D.I, nb = ax == 2 ? D.J : D.K;
  if (a > na || b > nb) return;
  int64_t h, hp;
  if (ax == 0) { h = hxi(D, D.I, a, b); hp = h - 1; }
  else if (ax == 1) { h = hxi(D, a, D.J, b); hp = h - D.SX; }
  else { h = hxi(D, a, b, D.K); hp = h - (int64_t)D.SX * D.SY; }
  A.S[h] = A.S[hp];
  A.Eta[h] = A.Eta[hp];
  A.Al[h] = A.Al[hp];
  A.DAl[h] = A.DAl[hp];
}

// Plane 0 along ax copies plane 1 over the whole halo range (:1073-1076).
__global__ void __launch_bounds__(256) k_ncf_plane0(NcfDims D, InterpArgs A) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int ax = A.ax;
  const int na = ax == 0 ? D.SY : D.SX, nb = ax == 2 ? D.SY : D.K + 2;
  if (a >= na || b >= nb) return;
  int64_t h, hp;
  if (ax == 0) { h = hxi(D, 0, a, b); hp = h + 1; }
  else if (ax == 1) { h = hxi(D, a, 0, b); hp = h + D.SX; }
  else { h = hxi(D, a, b, 0); hp = h + (int64_t)D.SX * D.SY; }
  A.S[h] = A.S[hp];
  A.Eta[h] = A.Eta[hp];
  A.Al[h] = A.Al[hp];
  A.DAl[h] = A.DAl[hp];
}

}  // namespace

sp_status new_cell_face(sp_ctx* c, const sp_cell_geom* const g[4],
                        const sp_cell_interp* const x[3], double* scratch) {
  const Dims& Dm = c->dm;
  NcfDims D{Dm.I, Dm.J, Dm.K, Dm.SX, Dm.SY, Dm.N, Dm.Nh};
  const dim3 blk(64, 4, 1);
  const dim3 cells((unsigned)((D.I + 63) / 64), (unsigned)((D.J + 3) / 4), (unsigned)D.K);
  const sp_cell_geom& P = *g[0];
  int slot;
  timer_begin(c, "ncf_centres", 0.0, &slot);
  hipLaunchKernelGGL(k_ncf_cent, cells, blk, 0, c->stream, D, P.ee_area, P.ne_area, P.te_area,
                     P.fce, P.fcn, P.fct, g[1]->cell_cent, g[2]->cell_cent, g[3]->cell_cent);
  for (int m = 0; m < 3; ++m) {
    const sp_cell_geom& T = *g[1 + m];
    double* orig = scratch + (size_t)m * 3 * D.N;
    hipError_t e = hipMemcpyAsync(orig, T.cell_cent, sizeof(double) * 3 * D.N,
                                  hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "NewCellFace centre copy");
    LinkGrid L;
    L.d[0] = T.dx; L.d[1] = T.dy; L.d[2] = T.dz;
    L.n[0] = T.nx; L.n[1] = T.ny; L.n[2] = T.nz;
    L.moex = x[m]->moex; L.orig = orig; L.cent = T.cell_cent; L.msce = x[m]->msce;
    hipLaunchKernelGGL(k_ncf_link, cells, blk, 0, c->stream, D, L);
    hipLaunchKernelGGL(k_ncf_delh, dim3((unsigned)((D.N + 255) / 256)), dim3(256), 0, c->stream,
                       D.N, T.cell_cent, T.nx, T.ny, T.nz, T.phi, T.delh);
  }
  timer_end(c, slot);
  timer_begin(c, "ncf_interp", 0.0, &slot);
  for (int m = 0; m < 3; ++m) {
    const sp_cell_geom& T = *g[1 + m];
    const sp_cell_interp& X = *x[m];
    double* S[3] = {X.sxe, X.syn, X.szt};
    double* Eta[3] = {X.etae, X.etan, X.etat};
    double* Al[3] = {X.ale, X.aln, X.alt};
    double* DAl[3] = {X.dale, X.daln, X.dalt};
    InterpArgs AA[3];
    for (int ax = 0; ax < 3; ++ax) {
      // the grid crossed by the face: P along the grid's own axis, else U/V/W
      const sp_cell_geom& B = ax == m ? P : *g[1 + ax];
      InterpArgs& A = AA[ax];
      A.d[0] = T.dx; A.d[1] = T.dy; A.d[2] = T.dz;
      A.n[0] = T.nx; A.n[1] = T.ny; A.n[2] = T.nz;
      A.phi = T.phi; A.cent = T.cell_cent;
      A.fc = ax == 0 ? T.fce : (ax == 1 ? T.fcn : T.fct);
      A.area = ax == 0 ? T.ee_area : (ax == 1 ? T.ne_area : T.te_area);
      A.bd = ax == 0 ? B.dx : (ax == 1 ? B.dy : B.dz);
      A.ct = T.cell_type;
      A.S = S[ax]; A.Eta = Eta[ax]; A.Al = Al[ax]; A.DAl = DAl[ax];
      A.ax = ax;
      A.ib = ax == m ? 1 : 0;
    }
    hipLaunchKernelGGL(k_ncf_interp, cells, blk, 0, c->stream, D, AA[0], AA[1], AA[2]);
    for (int ax = 0; ax < 3; ++ax) {
      const InterpArgs& A = AA[ax];
      const int na = ax == 0 ? D.J : D.I, nb = ax == 2 ? D.J : D.K;
      hipLaunchKernelGGL(k_ncf_last, dim3((unsigned)((na + 63) / 64), (unsigned)((nb + 3) / 4)),
                         blk, 0, c->stream, D, A);
      const int ha = ax == 0 ? D.SY : D.SX, hb = ax == 2 ? D.SY : D.K + 2;
      hipLaunchKernelGGL(k_ncf_plane0, dim3((unsigned)((ha + 63) / 64), (unsigned)((hb + 3) / 4)),
                         blk, 0, c->stream, D, A);
    }
  }
  timer_end(c, slot);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "NewCellFace launch");
}

}  // namespace sp
