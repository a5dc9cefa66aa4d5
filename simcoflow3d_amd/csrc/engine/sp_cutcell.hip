// Cut-cell geometry of the four staggered grids: GridPreprocess
// (CutCell.f90:62-325), run once at start-up (Main.f90:203) and after every
// move of a solid body (SolidBody.f90:815).
//
//   DefineCell (CutCell.f90:329-365): cell class from the solid vof;
//   CellGeoCal (CutCell.f90:369-671): the six face apertures of every cell
//     from the solid level set and normal (FaceGeoCal, :883-981), the
//     averaging of the two sides of each shared face, the MaxFace < 1e-2
//     demotion of P cells, zero apertures next to solid cells, the boundary
//     copies of the face centres and the wall length WlLh / distance delh;
//   the face-centre rule of U/V/W on stretched spacing (:89-159), the U/V/W
//     faces through the P-cell centre (:163-223) and the P faces through the
//     U/V/W cell centres (:225-323).
//
// All arithmetic follows the reference's operation order with contraction
// off, so every output is bit-identical to the sequential loops.  The
// reference keeps the raw face areas and centres of all six faces in
// scratch arrays; here a cell recomputes the three neighbour faces it shares
// (9 face clippings per cell instead of 6), which costs a few hundred FP64
// operations and saves 18 doubles of scratch per cell in HBM.  Steps that
// read a neighbour's result of an earlier step (demotion, zeroing, the
// final wall-length pass, the ordered boundary copies) are separate
// launches in the reference's order.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "sp_device.h"
#include "sp_internal.h"

namespace sp {
namespace {

constexpr double kEpsi = 1.0e-3;  // CutCell.f90:11

struct GeoDims {
  int I, J, K, SX, SY;
  int64_t N, Nh;
};

struct GeoGrid {
  const double *dx, *dy, *dz, *phi, *nx, *ny, *nz;
  double *vof, *vofl;
  int32_t* type;
  double *ee, *ne, *te, *fce, *fcn, *fct, *cent, *wllh, *delh;
};

__device__ inline int64_t hx(const GeoDims& D, int i, int j, int k) {
  return (int64_t)i + (int64_t)D.SX * ((int64_t)j + (int64_t)D.SY * (int64_t)k);
}
__device__ inline int64_t nx_(const GeoDims& D, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)D.I * ((int64_t)(j - 1) + (int64_t)D.J * (int64_t)(k - 1));
}

// The polygon vertices of FaceGeoCal (CutCell.f90:900-956) in the
// reference's order, handed to emit(x, y): the face corners on the positive
// side of n1*x + n2*y + n3*d3 + phi and the crossings of the zero line with
// the edges whose end values have opposite signs.
template <class Emit>
__device__ inline void face_nodes(double d11, double d22, double n1, double n2, double n3,
                                  double d3, double phi, Emit&& emit) {
#pragma clang fp contract(off)
  const double dp0 = d11 * n1 - d22 * n2 + d3 * n3 + phi;
  const double dp1 = -d11 * n1 - d22 * n2 + d3 * n3 + phi;
  const double dp2 = -d11 * n1 + d22 * n2 + d3 * n3 + phi;
  const double dp3 = d11 * n1 + d22 * n2 + d3 * n3 + phi;
  if (dp0 > 0.0) emit(d11, -d22);
  if (dp0 * dp1 < 0.0) emit((d22 * n2 - d3 * n3 - phi) / n1, -d22);
  if (dp1 > 0.0) emit(-d11, -d22);
  if (dp1 * dp2 < 0.0) emit(-d11, (d11 * n1 - d3 * n3 - phi) / n2);
  if (dp2 > 0.0) emit(-d11, d22);
  if (dp2 * dp3 < 0.0) emit((-d22 * n2 - d3 * n3 - phi) / n1, d22);
  if (dp3 > 0.0) emit(d11, d22);
  if (dp3 * dp0 < 0.0) emit(d11, (-d11 * n1 - d3 * n3 - phi) / n2);
}

// FaceGeoCal (CutCell.f90:883-981): area fraction and centroid of the part
// of [-d1/2,d1/2] x [-d2/2,d2/2] where n1*x + n2*y + n3*d3 + phi > 0.  The
// reference stores the vertices and sums over the closed polygon twice
// (signed area, then centroid with 1/6/vol); here the vertices are
// generated twice instead, so nothing is indexed dynamically (no private
// memory), and the sums run over the same edges in the same order.
__device__ inline double face_geo(double d1, double d2, double n1, double n2, double n3,
                                  double d3, double phi, double cent[3]) {
#pragma clang fp contract(off)
  const double d11 = d1 / 2.0, d22 = d2 / 2.0;
  double fx = 0.0, fy = 0.0, lx = 0.0, ly = 0.0, vol = 0.0;
  int t = 0;
  face_nodes(d11, d22, n1, n2, n3, d3, phi, [&](double x, double y) {
    if (t) vol = vol + 0.5 * (lx * y - x * ly);
    else { fx = x; fy = y; }
    lx = x; ly = y; ++t;
  });
  if (t) vol = vol + 0.5 * (lx * fy - fx * ly);  // closing edge
  const double area = fabs(vol / (d1 * d2));
  double c0 = 0.0, c1 = 0.0;
  if (area > 0.0) {
    const double s = 1.0 / 6.0 / vol;  // the same value in every term of :969-975
    int u = 0;
    face_nodes(d11, d22, n1, n2, n3, d3, phi, [&](double x, double y) {
      if (u) {
        const double cr = lx * y - x * ly;
        c0 = c0 + s * (lx + x) * cr;
        c1 = c1 + s * (ly + y) * cr;
      }
      lx = x; ly = y; ++u;
    });
    const double cr = lx * fy - fx * ly;
    c0 = c0 + s * (lx + fx) * cr;
    c1 = c1 + s * (ly + fy) * cr;
  }
  cent[0] = c0;
  cent[1] = c1;
  cent[2] = d3;
  return area;
}

// The west and south faces of cell q, as its -1 neighbour needs them for
// the shared-face average.
__device__ inline double west_face(const GeoGrid& G, int64_t q, double F[3]) {
#pragma clang fp contract(off)
  return face_geo(G.dy[q], G.dz[q], G.ny[q], G.nz[q], G.nx[q], -G.dx[q] / 2.0, G.phi[q], F);
}
__device__ inline double south_face(const GeoGrid& G, int64_t q, double F[3]) {
#pragma clang fp contract(off)
  return face_geo(G.dz[q], G.dx[q], G.nz[q], G.nx[q], G.ny[q], -G.dy[q] / 2.0, G.phi[q], F);
}

// Cells per thread along k in k_geo_cells.
constexpr int kGeoChunk = 32;

struct CellIn {
  double dx, dy, dz, phi, n1, n2, n3;
};

__device__ inline CellIn load_cell(const GeoGrid& G, int64_t q) {
  return CellIn{G.dx[q], G.dy[q], G.dz[q], G.phi[q], G.nx[q], G.ny[q], G.nz[q]};
}

// Initial aperture 1 on the halo entries of EEArea/NEArea/TEArea (:394-396):
// blockIdx.z picks one of the six boundary planes of the halo box (edges
// and corners are written by two or three planes, with the same value).
__global__ void __launch_bounds__(256) k_geo_halo(GeoDims D, GeoGrid G) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int f = blockIdx.z;
  const int ax = f >> 1;
  const int na = ax == 0 ? D.SY : D.SX;
  const int nb = ax == 2 ? D.SY : D.K + 2;
  if (a >= na || b >= nb) return;
  const int pl = (f & 1) ? (ax == 0 ? D.I + 1 : (ax == 1 ? D.J + 1 : D.K + 1)) : 0;
  const int64_t h = ax == 0 ? hx(D, pl, a, b) : (ax == 1 ? hx(D, a, pl, b) : hx(D, a, b, pl));
  G.ee[h] = 1.0;
  G.ne[h] = 1.0;
  G.te[h] = 1.0;
}

// DefineCell and the face pass of CellGeoCal (:387-555) for the interior.
// A thread walks kGeoChunk cells up a k column: each cell classifies
// itself, computes its faces and WlLh, and writes the averaged aperture and
// centre of its E, N, T faces from its own side and its +1 neighbour's W,
// S, B side.  The +1 neighbour along k is the next cell of the walk, so its
// inputs are loaded once and its bottom face, computed for the average, is
// the next cell's own bottom face (8 face clippings per cell).  The last
// cell of a line keeps its raw E (N, T) aperture: the averaging loops start
// at 2 and write index i-1.
__global__ void __launch_bounds__(256) k_geo_cells(GeoDims D, GeoGrid G) {
#pragma clang fp contract(off)
  // XCD-aware block order: the blocks of one XCD cover a contiguous range
  // of rows, so the j+1 neighbour row and the cache lines shared by two
  // blocks of a row stay in that XCD's L2
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  const int i = bx * blockDim.x + threadIdx.x + 1;
  const int j = by * blockDim.y + threadIdx.y + 1;
  const int k0 = bz * kGeoChunk + 1;
  if (i > D.I || j > D.J) return;
  const int k1 = min(D.K, k0 + kGeoChunk - 1);
  const double tol = 1.0e-20;
  const int64_t N = D.N, Nh = D.Nh, IJ = (int64_t)D.I * D.J, SXY = (int64_t)D.SX * D.SY;
  int64_t q = nx_(D, i, j, k0), h = hx(D, i, j, k0);
  CellIn c = load_cell(G, q);
  double F[3];
  double BE = face_geo(c.dx, c.dy, c.n1, c.n2, c.n3, -c.dz / 2.0, c.phi, F);
  for (int k = k0; k <= k1; ++k, q += IJ, h += SXY) {
    // DefineCell
    double v = G.vof[q], vl = G.vofl[q];
    int32_t ty = G.type[q];
    if (v >= 1.0 - kEpsi) { ty = 0; vl = vl / v; v = 1.0; }
    if (v <= kEpsi) { ty = 2; v = 0.0; vl = 0.0; }
    if (v >= kEpsi && v <= 1.0 - kEpsi) ty = 1;
    G.vof[q] = v;
    G.vofl[q] = vl;
    G.type[q] = ty;
    G.cent[q] = 0.0;
    G.cent[N + q] = 0.0;
    G.cent[2 * N + q] = 0.0;
    double Ft[3], Fe[3], Fn[3];
    const double TE = face_geo(c.dx, c.dy, c.n1, c.n2, c.n3, c.dz / 2.0, c.phi, Ft);
    const double EE = face_geo(c.dy, c.dz, c.n2, c.n3, c.n1, c.dx / 2.0, c.phi, Fe);
    const double WE = face_geo(c.dy, c.dz, c.n2, c.n3, c.n1, -c.dx / 2.0, c.phi, F);
    const double NE = face_geo(c.dz, c.dx, c.n3, c.n1, c.n2, c.dy / 2.0, c.phi, Fn);
    const double SE = face_geo(c.dz, c.dx, c.n3, c.n1, c.n2, -c.dy / 2.0, c.phi, F);
    const double a = (EE - WE) * c.dy * c.dz;
    const double b = (NE - SE) * c.dx * c.dz;
    const double e = (TE - BE) * c.dx * c.dy;
    G.wllh[q] = sqrt(a * a + b * b + e * e);
    if (i < D.I) {  // CeE = (F3, F1, F2), CeW of the neighbour likewise
      const double w = west_face(G, q + 1, F);
      G.ee[h] = (w * w + EE * EE) / (w + EE + tol);
      G.fce[Nh + h] = (F[0] * w + Fe[0] * EE) / (w + EE + tol);
      G.fce[2 * Nh + h] = (F[1] * w + Fe[1] * EE) / (w + EE + tol);
      G.fce[h] = Fe[2];
    } else {
      G.ee[h] = EE;
    }
    if (j < D.J) {  // CeN = (F2, F3, F1)
      const double s = south_face(G, q + D.I, F);
      G.ne[h] = (s * s + NE * NE) / (s + NE + tol);
      G.fcn[h] = (F[1] * s + Fn[1] * NE) / (s + NE + tol);
      G.fcn[2 * Nh + h] = (F[0] * s + Fn[0] * NE) / (s + NE + tol);
      G.fcn[Nh + h] = Fn[2];
    } else {
      G.ne[h] = NE;
    }
    if (k < D.K) {  // CeT = (F1, F2, F3); :541-543 weighs CeB(:,1) in both sums
      c = load_cell(G, q + IJ);
      const double bb = face_geo(c.dx, c.dy, c.n1, c.n2, c.n3, -c.dz / 2.0, c.phi, F);
      G.te[h] = (bb * bb + TE * TE) / (bb + TE + tol);
      G.fct[h] = (F[0] * bb + Ft[0] * TE) / (bb + TE + tol);
      G.fct[Nh + h] = (F[0] * bb + Ft[1] * TE) / (bb + TE + tol);
      G.fct[2 * Nh + h] = Ft[2];
      BE = bb;
    } else {
      G.te[h] = TE;
    }
  }
}

// :557-569: a P cell (i, j, k >= 2) whose six averaged apertures are all
// below 1e-2 becomes solid.
__global__ void __launch_bounds__(256) k_geo_demote(GeoDims D, GeoGrid G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
  const int k = blockIdx.z + 2;
  if (i > D.I || j > D.J) return;
  double mf = G.ee[hx(D, i - 1, j, k)];
  const double o[5] = {G.ee[hx(D, i, j, k)], G.ne[hx(D, i, j - 1, k)], G.ne[hx(D, i, j, k)],
                       G.te[hx(D, i, j, k - 1)], G.te[hx(D, i, j, k)]};
#pragma unroll
  for (int m = 0; m < 5; ++m) mf = o[m] > mf ? o[m] : mf;
  if (mf < 1.0e-2) G.type[nx_(D, i, j, k)] = 2;
}

// :571-599: a face next to a solid cell has zero aperture (the solid cell
// zeroes its six faces, a cut cell its E/N/T face towards a solid cell:
// together, a face is zero when either side is solid).  Each cell owns its
// E, N, T faces, and the first cell of a line also its W, S, B face.
__global__ void __launch_bounds__(256) k_geo_zero(GeoDims D, GeoGrid G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  const int64_t q = nx_(D, i, j, k), h = hx(D, i, j, k);
  const bool s = G.type[q] == 2;
  if (s || (i < D.I && G.type[q + 1] == 2)) G.ee[h] = 0.0;
  if (s || (j < D.J && G.type[q + D.I] == 2)) G.ne[h] = 0.0;
  if (s || (k < D.K && G.type[q + (int64_t)D.I * D.J] == 2)) G.te[h] = 0.0;
  if (s && i == 1) G.ee[h - 1] = 0.0;
  if (s && j == 1) G.ne[h - D.SX] = 0.0;
  if (s && k == 1) G.te[h - (int64_t)D.SX * D.SY] = 0.0;
}

// One of the boundary copies of :601-640: plane `dst` of axis `ax` takes
// plane `src`, over the whole halo range of the other two axes, for the
// face-centre arrays selected by `mask` (1 FCE, 2 FCN, 4 FCT), all three
// components.
__global__ void __launch_bounds__(256) k_geo_fc_copy(GeoDims D, GeoGrid G, int ax, int dst,
                                                     int src, int mask) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int na = ax == 0 ? D.SY : D.SX;
  const int nb = ax == 2 ? D.SY : D.K + 2;
  if (a >= na || b >= nb) return;
  int64_t hd, hs;
  if (ax == 0) { hd = hx(D, dst, a, b); hs = hx(D, src, a, b); }
  else if (ax == 1) { hd = hx(D, a, dst, b); hs = hx(D, a, src, b); }
  else { hd = hx(D, a, b, dst); hs = hx(D, a, b, src); }
  double* arr[3] = {G.fce, G.fcn, G.fct};
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (!(mask & (1 << f))) continue;
#pragma unroll
    for (int m = 0; m < 3; ++m) arr[f][m * D.Nh + hd] = arr[f][m * D.Nh + hs];
  }
}

// :642-666: a full cell (i, j, k >= 2) with a zero aperture gets the wall
// length and distance of that face; later directions win.
__global__ void __launch_bounds__(256) k_geo_final(GeoDims D, GeoGrid G) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
  const int k = blockIdx.z + 2;
  if (i > D.I || j > D.J) return;
  const int64_t q = nx_(D, i, j, k), h = hx(D, i, j, k);
  if (!(fabs(G.vof[q] - 1.0) < kEpsi)) return;
  const double dx = G.dx[q], dy = G.dy[q], dz = G.dz[q];
  double wl = 0.0, dl = 0.0;
  bool set = false;
  if (G.ee[h - 1] == 0.0 || G.ee[h] == 0.0) { wl = dy * dz; dl = dx / 2.0; set = true; }
  if (G.ne[h - D.SX] == 0.0 || G.ne[h] == 0.0) { wl = dx * dz; dl = dy / 2.0; set = true; }
  if (G.te[h - (int64_t)D.SX * D.SY] == 0.0 || G.te[h] == 0.0) {
    wl = dx * dy; dl = dz / 2.0; set = true;
  }
  if (set) { G.wllh[q] = wl; G.delh[q] = dl; }
}

// :89-159: on stretched spacing, the E (N, T) face centre of a U (V, W)
// cell whose neighbour along the axis is not cut sits half a P cell ahead;
// the last cell of a line has zero transverse centre components.
__global__ void __launch_bounds__(256) k_geo_stretch(GeoDims D, GeoGrid G, GeoGrid P, int ax) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  const int64_t q = nx_(D, i, j, k), h = hx(D, i, j, k), Nh = D.Nh;
  const int pos = ax == 0 ? i : (ax == 1 ? j : k);
  const int last = ax == 0 ? D.I : (ax == 1 ? D.J : D.K);
  const int64_t step = ax == 0 ? 1 : (ax == 1 ? (int64_t)D.I : (int64_t)D.I * D.J);
  double* own = ax == 0 ? G.fce : (ax == 1 ? G.fcn : G.fct);           // the axis' own face
  double* o1 = ax == 0 ? G.fcn : G.fce;                                 // the other two
  double* o2 = ax == 2 ? G.fcn : G.fct;
  const int64_t c = (int64_t)ax * Nh;
  if (pos < last) {
    const double* d = ax == 0 ? G.dx : (ax == 1 ? G.dy : G.dz);
    const double* pd = ax == 0 ? P.dx : (ax == 1 ? P.dy : P.dz);
    if (G.type[q] != 1 && G.type[q + step] != 1 && fabs(d[q] - d[q + step]) > 1.0e-10) {
      own[c + h] = 0.5 * pd[q + step];
      o1[c + h] = 0.0;
      o2[c + h] = 0.0;
    }
  } else {
    o1[c + h] = 0.0;
    o2[c + h] = 0.0;
  }
}

// :163-223: the U (V, W) face through the centre of P cell (i,j,k), i.e.
// entry i-1 (j-1, k-1) of the U (V, W) face arrays.
__global__ void __launch_bounds__(256) k_geo_pcent(GeoDims D, GeoGrid P, GeoGrid U, GeoGrid V,
                                                   GeoGrid W) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  const int64_t q = nx_(D, i, j, k), N = D.N, Nh = D.Nh;
  const int64_t hu = hx(D, i - 1, j, k), hv = hx(D, i, j - 1, k), hw = hx(D, i, j, k - 1);
  P.cent[q] = 0.0;
  P.cent[N + q] = 0.0;
  P.cent[2 * N + q] = 0.0;
  const double dx = P.dx[q], dy = P.dy[q], dz = P.dz[q], vf = P.vof[q];
  if (vf < 1.0 && vf > 0.0) {
    const double ph = P.phi[q], n1 = P.nx[q], n2 = P.ny[q], n3 = P.nz[q];
    double F[3];
    U.ee[hu] = face_geo(dy, dz, n2, n3, n1, 0.0, ph, F);
    U.fce[hu] = F[2] + dx / 2.0;
    U.fce[Nh + hu] = F[0];
    U.fce[2 * Nh + hu] = F[1];
    V.ne[hv] = face_geo(dz, dx, n3, n1, n2, 0.0, ph, F);
    V.fcn[Nh + hv] = F[2] + dy / 2.0;
    V.fcn[hv] = F[1];
    V.fcn[2 * Nh + hv] = F[0];
    W.te[hw] = face_geo(dx, dy, n1, n2, n3, 0.0, ph, F);
    W.fct[2 * Nh + hw] = F[2] + dz / 2.0;
    W.fct[hw] = F[0];
    W.fct[Nh + hw] = F[1];
  } else if (vf > 0.5) {
    U.fce[hu] = dx / 2.0; U.fce[Nh + hu] = 0.0; U.fce[2 * Nh + hu] = 0.0;
    U.ee[hu] = 1.0;
    V.fcn[Nh + hv] = dy / 2.0; V.fcn[hv] = 0.0; V.fcn[2 * Nh + hv] = 0.0;
    V.ne[hv] = 1.0;
    W.fct[2 * Nh + hw] = dz / 2.0; W.fct[hw] = 0.0; W.fct[Nh + hw] = 0.0;
    W.te[hw] = 1.0;
  } else {
    U.ee[hu] = 0.0; V.ne[hv] = 0.0; W.te[hw] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      U.fce[m * Nh + hu] = 0.0;
      V.fcn[m * Nh + hv] = 0.0;
      W.fct[m * Nh + hw] = 0.0;
    }
  }
}

// :225-323: the P face (E, N, T of cell (i,j,k)) through the centre of the
// U (V, W) cell (i,j,k), and that cell's centre.
__global__ void __launch_bounds__(256) k_geo_uvwcent(GeoDims D, GeoGrid P, GeoGrid U, GeoGrid V,
                                                     GeoGrid W) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > D.I || j > D.J) return;
  const int64_t q = nx_(D, i, j, k), h = hx(D, i, j, k), N = D.N, Nh = D.Nh;
  double F[3];
  // U: P%FCE = (F3 + dx/2, F1, F2), U%Cell_Cent = (F3, F1, F2)
  double vf = U.vof[q];
  if (vf < 1.0 && vf > 0.0) {
    face_geo(U.dy[q], U.dz[q], U.ny[q], U.nz[q], U.nx[q], 0.0, U.phi[q], F);
    P.fce[h] = F[2] + P.dx[q] / 2.0;
    P.fce[Nh + h] = F[0];
    P.fce[2 * Nh + h] = F[1];
    U.cent[q] = F[2]; U.cent[N + q] = F[0]; U.cent[2 * N + q] = F[1];
  } else if (vf > 0.5) {
    P.ee[h] = 1.0;
    P.fce[h] = P.dx[q] / 2.0; P.fce[Nh + h] = 0.0; P.fce[2 * Nh + h] = 0.0;
    U.cent[q] = 0.0; U.cent[N + q] = 0.0; U.cent[2 * N + q] = 0.0;
  } else {
    P.ee[h] = 0.0;
    P.fce[h] = 0.0; P.fce[Nh + h] = 0.0; P.fce[2 * Nh + h] = 0.0;
    U.cent[q] = 0.0; U.cent[N + q] = 0.0; U.cent[2 * N + q] = 0.0;
  }
  // V: P%FCN = (F2, F3 + dy/2, F1), V%Cell_Cent = (F2, F3, F1)
  vf = V.vof[q];
  if (vf < 1.0 && vf > 0.0) {
    face_geo(V.dz[q], V.dx[q], V.nz[q], V.nx[q], V.ny[q], 0.0, V.phi[q], F);
    P.fcn[h] = F[1];
    P.fcn[2 * Nh + h] = F[0];
    P.fcn[Nh + h] = F[2] + P.dy[q] / 2.0;
    V.cent[q] = F[1]; V.cent[N + q] = F[2]; V.cent[2 * N + q] = F[0];
  } else if (vf > 0.5) {
    P.ne[h] = 1.0;
    P.fcn[h] = 0.0; P.fcn[2 * Nh + h] = 0.0; P.fcn[Nh + h] = P.dy[q] / 2.0;
    V.cent[q] = 0.0; V.cent[N + q] = 0.0; V.cent[2 * N + q] = 0.0;
  } else {
    P.ne[h] = 0.0;
    P.fcn[h] = 0.0; P.fcn[Nh + h] = 0.0; P.fcn[2 * Nh + h] = 0.0;
    V.cent[q] = 0.0; V.cent[N + q] = 0.0; V.cent[2 * N + q] = 0.0;
  }
  // W: P%FCT = (F1, F2, F3 + dz/2), W%Cell_Cent = (F1, F2, F3)
  vf = W.vof[q];
  if (vf < 1.0 && vf > 0.0) {
    face_geo(W.dx[q], W.dy[q], W.nx[q], W.ny[q], W.nz[q], 0.0, W.phi[q], F);
    P.fct[h] = F[0];
    P.fct[Nh + h] = F[1];
    P.fct[2 * Nh + h] = F[2] + P.dz[q] / 2.0;
    W.cent[q] = F[0]; W.cent[N + q] = F[1]; W.cent[2 * N + q] = F[2];
  } else if (vf > 0.5) {
    P.te[h] = 1.0;
    P.fct[h] = 0.0; P.fct[Nh + h] = 0.0; P.fct[2 * Nh + h] = P.dz[q] / 2.0;
    W.cent[q] = 0.0; W.cent[N + q] = 0.0; W.cent[2 * N + q] = 0.0;
  } else {
    P.te[h] = 0.0;
    P.fct[h] = 0.0; P.fct[Nh + h] = 0.0; P.fct[2 * Nh + h] = 0.0;
    W.cent[q] = 0.0; W.cent[N + q] = 0.0; W.cent[2 * N + q] = 0.0;
  }
}

GeoGrid to_grid(const sp_cell_geom& g) {
  GeoGrid G;
  G.dx = g.dx; G.dy = g.dy; G.dz = g.dz;
  G.phi = g.phi; G.nx = g.nx; G.ny = g.ny; G.nz = g.nz;
  G.vof = g.vof; G.vofl = g.vof_l; G.type = g.cell_type;
  G.ee = g.ee_area; G.ne = g.ne_area; G.te = g.te_area;
  G.fce = g.fce; G.fcn = g.fcn; G.fct = g.fct;
  G.cent = g.cell_cent; G.wllh = g.wllh; G.delh = g.delh;
  return G;
}

}  // namespace

sp_status grid_preprocess(sp_ctx* c, const sp_cell_geom* const g[4]) {
  const Dims& Dm = c->dm;
  GeoDims D{Dm.I, Dm.J, Dm.K, Dm.SX, Dm.SY, Dm.N, Dm.Nh};
  GeoGrid G[4];
  for (int m = 0; m < 4; ++m) G[m] = to_grid(*g[m]);
  const dim3 blk(64, 4, 1);
  const int sab = std::max(std::max(D.SX, D.SY), D.K + 2);
  const dim3 halo_planes((unsigned)((sab + 63) / 64), (unsigned)((sab + 3) / 4), 6);
  const dim3 col_chunks((unsigned)((D.I + 63) / 64), (unsigned)((D.J + 3) / 4),
                        (unsigned)((D.K + kGeoChunk - 1) / kGeoChunk));
  const dim3 cell_grid((unsigned)((D.I + 63) / 64), (unsigned)((D.J + 3) / 4), (unsigned)D.K);
  const dim3 inner_grid((unsigned)((std::max(D.I - 1, 1) + 63) / 64),
                        (unsigned)((std::max(D.J - 1, 1) + 3) / 4),
                        (unsigned)std::max(D.K - 1, 1));
  const bool inner = D.I >= 2 && D.J >= 2 && D.K >= 2;
  auto plane = [&](int ax) {
    const int na = ax == 0 ? D.SY : D.SX, nb = ax == 2 ? D.SY : D.K + 2;
    return dim3((unsigned)((na + 63) / 64), (unsigned)((nb + 3) / 4), 1);
  };
  const double N = (double)D.N, Nh = (double)D.Nh;
  int slot;
  for (int m = 0; m < 4; ++m) {
    const GeoGrid& T = G[m];
    // per cell: 7 geometry doubles + vof, vofL, type read; vof, vofL, type,
    // 3 centre, 3 apertures, 9 face-centre components, WlLh written; halo
    // entries write 3 apertures
    timer_begin(c, "geo_cells", N * (9 * 8 + 4) + N * (15 * 8 + 8 + 4) + (Nh - N) * 24, &slot);
    hipLaunchKernelGGL(k_geo_halo, halo_planes, blk, 0, c->stream, D, T);
    hipLaunchKernelGGL(k_geo_cells, col_chunks, blk, 0, c->stream, D, T);
    timer_end(c, slot);
    timer_begin(c, "geo_faces", 0.0, &slot);
    if (m == 0 && inner) hipLaunchKernelGGL(k_geo_demote, inner_grid, blk, 0, c->stream, D, T);
    hipLaunchKernelGGL(k_geo_zero, cell_grid, blk, 0, c->stream, D, T);
    // the order of :601-640
    hipLaunchKernelGGL(k_geo_fc_copy, plane(0), blk, 0, c->stream, D, T, 0, 0, 1, 7);
    hipLaunchKernelGGL(k_geo_fc_copy, plane(0), blk, 0, c->stream, D, T, 0, D.I, D.I - 1, 1);
    hipLaunchKernelGGL(k_geo_fc_copy, plane(1), blk, 0, c->stream, D, T, 1, 0, 1, 7);
    hipLaunchKernelGGL(k_geo_fc_copy, plane(1), blk, 0, c->stream, D, T, 1, D.J, D.J - 1, 2);
    hipLaunchKernelGGL(k_geo_fc_copy, plane(2), blk, 0, c->stream, D, T, 2, 0, 1, 7);
    hipLaunchKernelGGL(k_geo_fc_copy, plane(2), blk, 0, c->stream, D, T, 2, D.K, D.K - 1, 4);
    if (inner) hipLaunchKernelGGL(k_geo_final, inner_grid, blk, 0, c->stream, D, T);
    timer_end(c, slot);
  }
  timer_begin(c, "geo_centres", 0.0, &slot);
  for (int ax = 0; ax < 3; ++ax)
    hipLaunchKernelGGL(k_geo_stretch, cell_grid, blk, 0, c->stream, D, G[1 + ax], G[0], ax);
  hipLaunchKernelGGL(k_geo_pcent, cell_grid, blk, 0, c->stream, D, G[0], G[1], G[2], G[3]);
  hipLaunchKernelGGL(k_geo_uvwcent, cell_grid, blk, 0, c->stream, D, G[0], G[1], G[2], G[3]);
  timer_end(c, slot);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_check(c, e, "grid preprocess launch");
  return SP_OK;
}

}  // namespace sp
