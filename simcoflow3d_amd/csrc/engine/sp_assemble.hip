// Assembly on the GPU: GFM face coefficients (ProjectionP.f90:123-265),
// the symmetric 7-point operator and the right-hand side
// (ProjectionP.f90:301-576).
//
// Every kernel here evaluates the reference's expressions in the reference's
// operation order with floating-point contraction disabled, so PoCoef and b
// are bit-identical to a serial evaluation of the Fortran code.

#include <hip/hip_runtime.h>

#include <cmath>

#include "sp_device.h"
#include "sp_geom.h"
#include "sp_internal.h"

namespace sp {
namespace {

// Cell classes of the GFM (ProjectionP.f90:170-174, 212-216).
constexpr int kNone = 0, kWet = 1, kDry = 2;

// The tests are evaluated without short-circuit (bitwise on the comparison
// results): they have no side effects, so the class is the same, and the
// kernels need no branch per test.
__device__ inline int gfm_class(double vof, double vofL, double phiL, double epsi) {
  const double one_m = 1.0 - epsi;
  const bool full = vof > one_m, cut = (vof <= one_m) & (vof > epsi);
  const bool wet = ((vofL >= 0.5) & full) | ((phiL < 0.0) & cut);
  const bool dry = ((vofL < 0.5) & full) | ((phiL >= 0.0) & cut);
  return wet ? kWet : (dry ? kDry : kNone);
}

struct GfmArgs {
  int I, J, K, SX, SY;
  int64_t N;
  const double *phiL, *vofL, *vof;
  const double* dp[3];     // Dp on U, V, W faces (halo)
  const double* dgrid[3];  // dxu, dyv, dzw
  double betaP, betaM, epsi;
  double* pocoef;
  int32_t* stale;          // (n, face, kind) triples
  uint32_t stale_cap;
  uint32_t* stale_count;
  int64_t* col_last;       // [12][I*J] loop position of the last assignment
  int64_t* col_prefix;     // [12][I*J]
  uint8_t* cls;            // GFM class of the interior cells (halo layout)
};

__device__ inline int64_t hstride(const GfmArgs& A, int dir) {
  return dir == 0 ? 1 : (dir == 1 ? (int64_t)A.SX : (int64_t)A.SX * A.SY);
}

// What the reference's loop body does for face f at 0-based cell (a,b,c):
//   returns kWet / kDry when the branch assigns BetaW / BetaD (value in beta),
//   -kWet / -kDry when the neighbour is unclassified (stale value reused),
//   kNone when the P cell is unclassified (coefficient stays 0).
__device__ inline int gfm_face(const GfmArgs& A, int f, int a, int b, int c, double& beta,
                               double& pref) {
#pragma clang fp contract(off)
  const int dir = f % 3;
  const int side = f < 3 ? -1 : 1;
  const int64_t h = (int64_t)(a + 1) + (int64_t)A.SX * ((int64_t)(b + 1) + (int64_t)A.SY * (c + 1));
  const int64_t s = hstride(A, dir);
  const int64_t q = h + side * s;
  const int coord = dir == 0 ? a : (dir == 1 ? b : c);
  // DGrid: TGrid%d(max(im,1)) with im = i-1 on the minus side (:151-153)
  const double dgrid = A.dgrid[dir][side < 0 ? (coord > 0 ? coord - 1 : 0) : coord];
  const double Dp = A.dp[dir][side < 0 ? h - s : h];
  const double fP = A.phiL[h], fQ = A.phiL[q];
  const double lam = fabs(fP) / (fabs(fP) + fabs(fQ) + 1.0e-30);  // :155
  const int cP = gfm_class(A.vof[h], A.vofL[h], fP, A.epsi);
  // prefactor Dp/DGrid*BetaP*BetaM of :194 / :235
  pref = Dp / dgrid * A.betaP * A.betaM;
  if (cP == kNone) return kNone;
  const int cQ = gfm_class(A.vof[q], A.vofL[q], fQ, A.epsi);
  if (cP == kWet) {
    if (cQ == kDry) { beta = lam * A.betaM + (1.0 - lam) * A.betaP; return kWet; }  // :176
    if (cQ == kWet) { beta = A.betaM; return kWet; }                                 // :181
    return -kWet;
  }
  if (cQ == kWet) { beta = lam * A.betaP + (1.0 - lam) * A.betaM; return kDry; }     // :218
  if (cQ == kDry) { beta = A.betaP; return kDry; }                                   // :223
  return -kDry;
}

// GFM coefficients of the six faces of every cell; KC cells of a column per
// thread.  The arithmetic is gfm_face's, expression for expression.  Every
// load of a cell (its own fields, the six neighbours', Dp on both faces of
// each direction) is issued before any class test, at addresses that are
// always valid (halo arrays), and the reference's branches select among the
// results: neighbour loads behind the class test would add a memory latency
// per cell.
template <int KC>
__global__ void __launch_bounds__(256) k_gfm(GfmArgs A) {
#pragma clang fp contract(off)
  const Col3 q = col_of<KC>(A.I, A.J, A.K);
  if (!q.in) return;
  const int a = q.a, b = q.b;
  const int64_t IJ = (int64_t)A.I * A.J, SXY = (int64_t)A.SX * A.SY;
  int64_t n = (int64_t)a + (int64_t)A.I * ((int64_t)b + (int64_t)A.J * q.c0);
  int64_t h = (int64_t)(a + 1) + (int64_t)A.SX * ((int64_t)(b + 1) + (int64_t)A.SY * (q.c0 + 1));
#pragma unroll
  for (int cc = 0; cc < KC; ++cc, n += IJ, h += SXY) {
    const int c = q.c0 + cc;
    if (c >= q.c1) break;
    const int64_t st[3] = {1, (int64_t)A.SX, SXY};
    const double fP = A.phiL[h], vP = A.vof[h], lP = A.vofL[h];
    double fQ[6], vQ[6], lQ[6], Dp[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int dir = f % 3;
      const int64_t hq = f < 3 ? h - st[dir] : h + st[dir];
      fQ[f] = A.phiL[hq];
      vQ[f] = A.vof[hq];
      lQ[f] = A.vofL[hq];
      Dp[f] = A.dp[dir][f < 3 ? h - st[dir] : h];
    }
    const int cP = gfm_class(vP, lP, fP, A.epsi);
    A.cls[h] = (uint8_t)cP;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int dir = f % 3;
      const int side = f < 3 ? -1 : 1;
      const int coord = dir == 0 ? a : (dir == 1 ? b : c);
      // DGrid: TGrid%d(max(im,1)) with im = i-1 on the minus side (:151-153)
      const double dgrid = A.dgrid[dir][side < 0 ? (coord > 0 ? coord - 1 : 0) : coord];
      const double pref = Dp[f] / dgrid * A.betaP * A.betaM;  // :194 / :235
      const int cQ = gfm_class(vQ[f], lQ[f], fQ[f], A.epsi);
      // wet P: dry Q lam*BetaM + (1-lam)*BetaP (:176), wet Q BetaM (:181);
      // dry P: wet Q lam*BetaP + (1-lam)*BetaM (:218), dry Q BetaP (:223).
      // Operands are selected, the arithmetic is the reference's.
      const bool pw = cP == kWet;
      const double b1 = pw ? A.betaM : A.betaP, b2 = pw ? A.betaP : A.betaM;
      const bool other = pw ? cQ == kDry : cQ == kWet;
      const bool same = cQ == cP;
      // lambda (:155) enters only on faces between a wet and a dry cell,
      // a thin layer: its division is skipped by waves that have none
      double beta = same ? b1 : 1.0;
      if (other) {
        const double lam = fabs(fP) / (fabs(fP) + fabs(fQ[f]) + 1.0e-30);  // :155
        beta = lam * b1 + (1.0 - lam) * b2;
      }
      const bool assigns = (cP != kNone) & (other | same);
      const double coef = assigns ? pref / beta : 0.0;
      // P classified, neighbour not: the previous beta in loop order
      if (cP != kNone && cQ == kNone) {
        const uint32_t slot = atomicAdd(A.stale_count, 1u);
        if (slot < A.stale_cap) {
          A.stale[3 * (int64_t)slot + 0] = (int32_t)n;
          A.stale[3 * (int64_t)slot + 1] = f;
          A.stale[3 * (int64_t)slot + 2] = cP;
        }
      }
      A.pocoef[(int64_t)f * A.N + n] = coef;
    }
  }
}
// cells per thread of k_gfm
constexpr int kGfmK = 1;

// Whether face f of 0-based cell (a,b,c) assigns BetaW (kWet) or BetaD
// (kDry) in the reference's loop, from the classes k_gfm stored: both the
// P cell and its neighbour classified (gfm_face's positive results).  A
// neighbour in the halo is classified from its fields.
__device__ inline int face_kind(const GfmArgs& A, int f, int a, int b, int c) {
  const int dir = f % 3;
  const int side = f < 3 ? -1 : 1;
  const int64_t h = (int64_t)(a + 1) + (int64_t)A.SX * ((int64_t)(b + 1) + (int64_t)A.SY * (c + 1));
  const int cP = A.cls[h];
  if (cP == kNone) return kNone;
  const int64_t q = h + side * hstride(A, dir);
  const int coord = (dir == 0 ? a : (dir == 1 ? b : c)) + side;
  const int lim = dir == 0 ? A.I : (dir == 1 ? A.J : A.K);
  const int cQ = coord >= 0 && coord < lim ? (int)A.cls[q]
                                           : gfm_class(A.vof[q], A.vofL[q], A.phiL[q], A.epsi);
  return cQ == kNone ? kNone : cP;
}

// Per (column (i,j), face): loop position of the last cell of the column that
// assigns BetaW (kind 1) and BetaD (kind 2), -1 if none.  Columns are the
// innermost k loop of the reference (ProjectionP.f90:142-144).  One thread
// per (face, column) walks the column's stored classes down from the top
// until it has found both.
__global__ void __launch_bounds__(256) k_gfm_col_last(GfmArgs A) {
  if (*A.stale_count == 0u) return;  // no face needs the previous beta
  const int64_t cols = (int64_t)A.I * A.J;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cols * 6) return;
  const int f = (int)(t / cols);
  // consecutive threads take consecutive i (coalesced loads); the column's
  // loop position is col = a*J + b (i outer, j inner)
  const int64_t u = t % cols;
  const int a = (int)(u % A.I), b = (int)(u / A.I);
  const int64_t col = (int64_t)a * A.J + b;
  int64_t lastW = -1, lastD = -1;
  // eight planes' classes in flight at a time
  for (int c0 = A.K - 1; c0 >= 0 && (lastW < 0 || lastD < 0); c0 -= 8) {
    int r[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = c0 - e >= 0 ? face_kind(A, f, a, b, c0 - e) : kNone;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t pos = col * A.K + (c0 - e);
      if (r[e] == kWet && lastW < 0) lastW = pos;
      if (r[e] == kDry && lastD < 0) lastD = pos;
    }
  }
  A.col_last[(int64_t)(2 * f + 0) * cols + col] = lastW;
  A.col_last[(int64_t)(2 * f + 1) * cols + col] = lastD;
}

// Exclusive prefix-max over columns (positions grow along the loop order, so
// "last assignment before column c" is the maximum over earlier columns).
// One workgroup of 1024 threads per (face, kind) walks the columns in tiles
// of 4096: coalesced loads into LDS, four consecutive columns per thread, a
// max-scan over the threads, the maximum of the earlier tiles.
__global__ void __launch_bounds__(1024) k_gfm_col_prefix(GfmArgs A) {
  if (*A.stale_count == 0u) return;
  constexpr int kTile = 4096;
  __shared__ int64_t tile[kTile];
  __shared__ int64_t sm[1024];
  const int64_t cols = (int64_t)A.I * A.J;
  const int64_t* src = A.col_last + (int64_t)blockIdx.x * cols;
  int64_t* dst = A.col_prefix + (int64_t)blockIdx.x * cols;
  const int t = threadIdx.x;
  int64_t carry = -1;
  for (int64_t base = 0; base < cols; base += kTile) {
    for (int e = 0; e < 4; ++e) {
      const int l = e * 1024 + t;
      tile[l] = base + l < cols ? src[base + l] : -1;
    }
    __syncthreads();
    int64_t m = -1;
    for (int e = 0; e < 4; ++e) m = tile[4 * t + e] > m ? tile[4 * t + e] : m;
    sm[t] = m;
    __syncthreads();
    // inclusive Hillis-Steele max-scan
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t v = t >= off ? sm[t - off] : -1;
      __syncthreads();
      sm[t] = v > sm[t] ? v : sm[t];
      __syncthreads();
    }
    int64_t run = t > 0 ? sm[t - 1] : -1;
    run = carry > run ? carry : run;
    for (int e = 0; e < 4; ++e) {
      const int l = 4 * t + e;
      const int64_t v = tile[l];
      tile[l] = run;
      run = v > run ? v : run;
    }
    __syncthreads();
    carry = sm[1023] > carry ? sm[1023] : carry;
    for (int e = 0; e < 4; ++e) {
      const int l = e * 1024 + t;
      if (base + l < cols) dst[base + l] = tile[l];
    }
    __syncthreads();
  }
}

// Resolve the faces whose neighbour is unclassified: the reference reuses the
// BetaW/BetaD left by the most recent assigning cell in its i-j-k loop.
__global__ void __launch_bounds__(256) k_gfm_fixup(GfmArgs A) {
#pragma clang fp contract(off)
  const uint32_t count = min(*A.stale_count, A.stale_cap);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count;
       t += gridDim.x * blockDim.x) {
    const int64_t n = A.stale[3 * (int64_t)t + 0];
    const int f = A.stale[3 * (int64_t)t + 1];
    const int kind = A.stale[3 * (int64_t)t + 2];
    const int a = (int)(n % A.I), b = (int)((n / A.I) % A.J), c = (int)(n / ((int64_t)A.I * A.J));
    const int64_t cols = (int64_t)A.I * A.J;
    const int64_t col = (int64_t)a * A.J + b;
    int64_t pos = -1;
    for (int cc = c - 1; cc >= 0; --cc)
      if (face_kind(A, f, a, b, cc) == kind) { pos = col * A.K + cc; break; }
    if (pos < 0) pos = A.col_prefix[(int64_t)(2 * f + kind - 1) * cols + col];
    double beta;
    if (pos >= 0) {
      const int64_t pc = pos / A.K;
      const int cc = (int)(pos % A.K);
      double pref2;
      gfm_face(A, f, (int)(pc / A.J), (int)(pc % A.J), cc, beta, pref2);
    } else {
      // no earlier assignment: the reference's local is undefined here; use the
      // P cell's own phase
      beta = kind == kWet ? A.betaM : A.betaP;
    }
    double b0, pref;
    gfm_face(A, f, a, b, c, b0, pref);
    A.pocoef[(int64_t)f * A.N + n] = pref / beta;
  }
}

// Boundary faces copy the first interior value (ProjectionP.f90:256-264).
__global__ void __launch_bounds__(256) k_gfm_boundary(GfmArgs A) {
  const int I = A.I, J = A.J, K = A.K;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int face = blockIdx.y;  // 0..5
  double* T = A.pocoef + (int64_t)face * A.N;
  auto idx = [&](int a, int b, int c) { return (int64_t)a + (int64_t)I * ((int64_t)b + (int64_t)J * c); };
  if (face == 0 || face == 3) {  // W / E: plane over (j,k)
    if (t >= (int64_t)J * K) return;
    const int b = (int)(t % J), c = (int)(t / J);
    if (face == 0) T[idx(0, b, c)] = T[idx(1, b, c)];
    else T[idx(I - 1, b, c)] = T[idx(I - 2, b, c)];
  } else if (face == 1 || face == 4) {
    if (t >= (int64_t)I * K) return;
    const int a = (int)(t % I), c = (int)(t / I);
    if (face == 1) T[idx(a, 0, c)] = T[idx(a, 1, c)];
    else T[idx(a, J - 1, c)] = T[idx(a, J - 2, c)];
  } else {
    if (t >= (int64_t)I * J) return;
    const int a = (int)(t % I), b = (int)(t / I);
    if (face == 2) T[idx(a, b, 0)] = T[idx(a, b, 1)];
    else T[idx(a, b, K - 1)] = T[idx(a, b, K - 2)];
  }
}

// Beta parametrisation: PoCoef(:,4..6) = Dp/delta * beta, W,S,B from the
// neighbour's E,N,T value (first plane uses its own).
struct BetaArgs {
  int I, J, K, SX, SY;
  int64_t N;
  const double* beta[3];
  const double* dp[3];
  const double* dgrid[3];
  double* pocoef;
};

__global__ void __launch_bounds__(256) k_beta_coef(BetaArgs A) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  const int c = blockIdx.z;
  if (a >= A.I || b >= A.J) return;
  const int64_t n = (int64_t)a + (int64_t)A.I * ((int64_t)b + (int64_t)A.J * c);
  const int64_t h = (int64_t)(a + 1) + (int64_t)A.SX * ((int64_t)(b + 1) + (int64_t)A.SY * (c + 1));
  const int coord[3] = {a, b, c};
  const int64_t ns[3] = {1, A.I, (int64_t)A.I * A.J};
  const int64_t hs[3] = {1, A.SX, (int64_t)A.SX * A.SY};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double kp = A.dp[d][h] / A.dgrid[d][coord[d]] * A.beta[d][n];
    A.pocoef[(int64_t)(3 + d) * A.N + n] = kp;
    const int64_t nm = coord[d] > 0 ? n - ns[d] : n;
    const int64_t hm = coord[d] > 0 ? h - hs[d] : h;
    const int cm = coord[d] > 0 ? coord[d] - 1 : coord[d];
    A.pocoef[(int64_t)d * A.N + n] = A.dp[d][hm] / A.dgrid[d][cm] * A.beta[d][nm];
  }
}

struct OpArgs {
  int I, J, K, SX, SY;
  int64_t N;
  const double* pocoef;
  const double *ee, *ne, *te;
  const int32_t* ct;
  int32_t fl[6];
  const double* bcv[6];
  const double *u, *v, *w, *vof, *vofO;
  const double *dxp, *dyp, *dzp;
  double dt;
  double *cx, *cy, *cz, *e, *b;
  float *cx32, *cy32, *cz32, *e32;
  uint8_t* active;
  float* dw;
  RedSite site;
  Scalars* S;
};

// Symmetric operator, diagonal-only terms and RHS.  One cell per thread
// (kOpK = 1) measured fastest: the body has many independent loads, and
// walking several cells per thread serialises them.
constexpr int kOpK = 1;
__global__ void __launch_bounds__(256) k_operator(OpArgs A) {
#pragma clang fp contract(off)
  const Col3 q = col_of<kOpK>(A.I, A.J, A.K);
  const int64_t IJ = (int64_t)A.I * A.J, SXY = (int64_t)A.SX * A.SY, N = A.N;
  // sum b, sum |e|, max |b|/V, active count, non-finite, sum d (as stored)
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (q.in) {
    const int a = q.a, b = q.b;
    const double dx = A.dxp[a], dy = A.dyp[b];
    int64_t n = (int64_t)a + (int64_t)A.I * ((int64_t)b + (int64_t)A.J * q.c0);
    int64_t h = (int64_t)(a + 1) + (int64_t)A.SX * ((int64_t)(b + 1) + (int64_t)A.SY * (q.c0 + 1));
#pragma unroll
    for (int cc = 0; cc < kOpK; ++cc, n += IJ, h += SXY) {
      const int c = q.c0 + cc;
      if (c >= q.c1) break;
      const double dz = A.dzp[c];
      // every load the row needs in every case, issued before any test at
      // addresses that are always valid (a neighbour past the domain edge
      // reads the cell itself); the reference's conditions then select
      const bool hW = a > 0, hS = b > 0, hB = c > 0;
      const bool hE = a < A.I - 1, hN = b < A.J - 1, hT = c < A.K - 1;
      const int64_t nW = hW ? n - 1 : n, nS = hS ? n - A.I : n, nB = hB ? n - IJ : n;
      const int64_t nE = hE ? n + 1 : n, nN = hN ? n + A.I : n, nT = hT ? n + IJ : n;
      const int ctP = A.ct[n];
      const int ctW = A.ct[nW], ctS = A.ct[nS], ctB = A.ct[nB];
      const int ctE = A.ct[nE], ctN = A.ct[nN], ctT = A.ct[nT];
      const double E = A.pocoef[3 * N + n], Nn = A.pocoef[4 * N + n], T = A.pocoef[5 * N + n];
      const double Ew = A.pocoef[3 * N + nW], Ns = A.pocoef[4 * N + nS], Tb = A.pocoef[5 * N + nB];
      const double eeP = A.ee[h], eeW = A.ee[h - 1], neP = A.ne[h], neS = A.ne[h - A.SX];
      const double teP = A.te[h], teB = A.te[h - SXY];
      const double uP = A.u[h], uW = A.u[h - 1], vP = A.v[h], vS = A.v[h - A.SX];
      const double wP = A.w[h], wB = A.w[h - SXY];
      const double vofP = A.vof[h], vofOP = A.vofO[h];
      double cxv = 0, cyv = 0, czv = 0, cW = 0, cS = 0, cB = 0, ex = 0, rhs = 0, dval = 0;
      bool act = false;
      if (ctP != 2) {
        // the W/S/B coefficients enter only at walls and faces towards solid
        // cells: loaded there alone (a third of the PoCoef traffic otherwise)
        const double* pcW = A.pocoef;
        const double* pcS = A.pocoef + N;
        const double* pcB = A.pocoef + 2 * N;
        const bool wS = hW ? ctW == 2 : A.fl[0] == 0;
        const bool sS = hS ? ctS == 2 : A.fl[2] == 0;
        const bool bS = hB ? ctB == 2 : A.fl[4] == 0;
        const bool eS = hE ? ctE == 2 : A.fl[1] == 0;
        const bool nS = hN ? ctN == 2 : A.fl[3] == 0;
        const bool tS = hT ? ctT == 2 : A.fl[5] == 0;
        // couplings to active neighbours: the lower cell's E/N/T coefficient
        cxv = (hE && ctE != 2) ? E * eeP * dy * dz : 0.0;
        cyv = (hN && ctN != 2) ? Nn * neP * dx * dz : 0.0;
        czv = (hT && ctT != 2) ? T * teP * dx * dy : 0.0;
        cW = (hW && ctW != 2) ? Ew * eeW * dy * dz : 0.0;
        cS = (hS && ctS != 2) ? Ns * neS * dx * dz : 0.0;
        cB = (hB && ctB != 2) ? Tb * teB * dx * dy : 0.0;
        // diagonal-only terms with the row's own coefficient (:379-430)
        if (wS) ex = ex + pcW[n] * eeW * dy * dz;
        if (sS) ex = ex + pcS[n] * neS * dx * dz;
        if (bS) ex = ex + pcB[n] * teB * dx * dy;
        if (eS) ex = ex + E * eeP * dy * dz;
        if (nS) ex = ex + Nn * neP * dx * dz;
        if (tS) ex = ex + T * teP * dx * dy;
        double d = cW;
        d = d + cS;
        d = d + cB;
        d = d + cxv;
        d = d + cyv;
        d = d + czv;
        d = d + ex;
        act = d > 0.0;
        dval = act ? d : 0.0;
        if (!is_finite(d)) v[4] = 1.0;
        // RHS (:521-546)
        rhs = -dy * dz * (eeP * uP - eeW * uW) -
              dx * dz * (neP * vP - neS * vS) -
              dx * dy * (teP * wP - teB * wB);
        rhs = rhs + (vofP - vofOP) * dx * dy * dz / A.dt;
        if (A.fl[0] == 0 && a == 0) rhs = rhs + pcW[n] * A.bcv[0][b + (int64_t)A.J * c];
        if (A.fl[1] == 0 && a == A.I - 1) rhs = rhs + E * A.bcv[1][b + (int64_t)A.J * c];
        if (A.fl[2] == 0 && b == 0) rhs = rhs + pcS[n] * A.bcv[2][a + (int64_t)A.I * c];
        if (A.fl[3] == 0 && b == A.J - 1) rhs = rhs + Nn * A.bcv[3][a + (int64_t)A.I * c];
        if (A.fl[4] == 0 && c == 0) rhs = rhs + pcB[n] * A.bcv[4][a + (int64_t)A.I * b];
        if (A.fl[5] == 0 && c == A.K - 1) rhs = rhs + T * A.bcv[5][a + (int64_t)A.I * b];
        if (act) {
          if (!is_finite(rhs)) v[4] = 1.0;
          v[0] += rhs;
          v[1] += fabs(ex);
          const double bv = fabs(rhs) / (dx * dy * dz);
          v[2] = (bv > v[2] || bv != bv) ? bv : v[2];
          v[3] += 1.0;
          v[5] += (double)(float)dval;
        }
      }
      A.cx[n] = cxv; A.cy[n] = cyv; A.cz[n] = czv; A.e[n] = ex; A.b[n] = rhs;
      A.active[n] = act ? 1 : 0;
      A.dw[n] = (float)dval;
      if (A.cx32) {
        A.cx32[n] = (float)cxv; A.cy32[n] = (float)cyv; A.cz32[n] = (float)czv;
        A.e32[n] = (float)ex;
      }
    }
  }
  // value 2 (max |b|/V) and 4 (non-finite flag) are maxima
  block_partials<kTileThreads, 6, 0x14u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_operator_fin(OpArgs A, uint32_t nb) {
  double t[6];
  finalize_partials<kFinThreads, 6, 0x14u>(A.site, nb, t);
  if (flat_tid() == 0) {
    Scalars* S = A.S;
    S->bsum = t[0];
    S->esum = t[1];
    S->bmax = t[2];
    S->active = (int64_t)t[3];
    S->dsum = t[5];
    S->singular = t[1] == 0.0 ? 1 : 0;
    S->bmean = (t[1] == 0.0 && t[3] > 0) ? t[0] / t[3] : 0.0;
    S->nonfinite = (t[4] > 0.0 || !is_finite(t[0]) || !is_finite(t[2])) ? 1 : 0;
  }
}

}  // namespace

sp_status assemble(sp_ctx* c, const sp_inputs& in, int64_t* stale_out) {
  const Dims& D = c->dm;
  const bool mixed = c->prm.precision == SP_PREC_MIXED;
  dim3 blk(64, 4, 1), grd((D.I + 63) / 64, (D.J + 3) / 4, D.K);
  int slot;
  if (in.beta_e) {
    BetaArgs B{};
    B.I = D.I; B.J = D.J; B.K = D.K; B.SX = D.SX; B.SY = D.SY; B.N = D.N;
    B.beta[0] = in.beta_e; B.beta[1] = in.beta_n; B.beta[2] = in.beta_t;
    B.dp[0] = in.dp_u; B.dp[1] = in.dp_v; B.dp[2] = in.dp_w;
    B.dgrid[0] = c->dxu; B.dgrid[1] = c->dyv; B.dgrid[2] = c->dzw;
    B.pocoef = c->pocoef;
    timer_begin(c, "beta_coef", (double)D.N * (3 * 8 + 6 * 8 + 3 * 8), &slot);
    hipLaunchKernelGGL(k_beta_coef, grd, blk, 0, c->stream, B);
    timer_end(c, slot);
  } else {
    GfmArgs G{};
    G.I = D.I; G.J = D.J; G.K = D.K; G.SX = D.SX; G.SY = D.SY; G.N = D.N;
    G.phiL = in.phi_l; G.vofL = in.vof_l; G.vof = in.vof;
    G.dp[0] = in.dp_u; G.dp[1] = in.dp_v; G.dp[2] = in.dp_w;
    G.dgrid[0] = c->dxu; G.dgrid[1] = c->dyv; G.dgrid[2] = c->dzw;
    G.betaP = 1.0 / (c->prm.rho_l / c->prm.rho_ref);  // ProjectionP.f90:140
    G.betaM = 1.0 / (c->prm.rho_g / c->prm.rho_ref);  // ProjectionP.f90:141
    G.epsi = c->prm.epsi;
    G.pocoef = c->pocoef;
    G.stale = c->stale_list;
    G.stale_cap = (uint32_t)c->stale_cap;
    G.stale_count = &c->S->stale_count;
    G.col_last = c->col_last;
    G.col_prefix = c->col_prefix;
    G.cls = c->gfm_cls;
    timer_begin(c, "gfm", (double)D.N * (3 * 8 + 3 * 8 + 6 * 8), &slot);
    hipLaunchKernelGGL(k_gfm<kGfmK>, col_grid_k<kGfmK>(D.I, D.J, D.K), blk, 0, c->stream, G);
    timer_end(c, slot);
    // faces next to unclassified cells take the previous beta in loop
    // order; the three kernels read the count on the device and return at
    // once when there are none (the count is checked against the list's
    // capacity after the assembly)
    {
      const int64_t cols = (int64_t)D.I * D.J;
      timer_begin(c, "gfm_col_last", 0, &slot);
      hipLaunchKernelGGL(k_gfm_col_last, dim3((unsigned)((cols * 6 + 255) / 256)), dim3(256), 0,
                         c->stream, G);
      timer_end(c, slot);
      timer_begin(c, "gfm_col_prefix", 0, &slot);
      hipLaunchKernelGGL(k_gfm_col_prefix, dim3(12), dim3(1024), 0, c->stream, G);
      timer_end(c, slot);
      timer_begin(c, "gfm_fixup", 0, &slot);
      const int64_t fb = std::min<int64_t>((c->stale_cap + 255) / 256, 2048);
      hipLaunchKernelGGL(k_gfm_fixup, dim3((unsigned)fb), dim3(256), 0, c->stream, G);
      timer_end(c, slot);
    }
    int64_t plane = std::max<int64_t>(std::max<int64_t>((int64_t)D.J * D.K, (int64_t)D.I * D.K),
                                      (int64_t)D.I * D.J);
    hipLaunchKernelGGL(k_gfm_boundary, dim3((unsigned)((plane + 255) / 256), 6), dim3(256), 0,
                       c->stream, G);
  }
  if (stale_out) *stale_out = 0;  // counted on the device (Scalars::stale_count)

  OpArgs O{};
  O.I = D.I; O.J = D.J; O.K = D.K; O.SX = D.SX; O.SY = D.SY; O.N = D.N;
  O.pocoef = c->pocoef;
  O.ee = in.ee_area; O.ne = in.ne_area; O.te = in.te_area;
  O.ct = in.cell_type_p;
  for (int f = 0; f < 6; ++f) {
    O.fl[f] = in.bc_flag[f];
    O.bcv[f] = c->bcv[f];
  }
  O.u = in.u_star; O.v = in.v_star; O.w = in.w_star; O.vof = in.vof; O.vofO = in.vof_old;
  O.dxp = c->dxp; O.dyp = c->dyp; O.dzp = c->dzp;
  O.dt = in.dt;
  O.cx = c->cx; O.cy = c->cy; O.cz = c->cz; O.e = c->ee; O.b = c->b;
  if (mixed && c->lf[0].cx) {
    O.cx32 = c->lf[0].cx; O.cy32 = c->lf[0].cy; O.cz32 = c->lf[0].cz; O.e32 = c->lf[0].e_store;
  }
  O.active = c->active;
  O.dw = c->dw;
  O.site = red_site(c, kSiteOperator);
  O.S = c->S;
  if (flat_blocks(D.I, D.J, D.K) > c->red_max_blocks)
    return fail(c, SP_ERR_INVALID_ARG, "grid too large for the reduction buffers");
  // per cell: PoCoef E,N,T (W,S,B only at walls and solid faces), apertures,
  // cell type, u*, v*, w*, vof, vofO -> cx, cy, cz, e, b, mask, diagonal
  // weight (and the FP32 couplings)
  timer_begin(c, "operator_rhs",
              (double)D.N * (3 * 8 + 3 * 8 + 4 + 5 * 8 + 4 * 8 + 8 + 1 + 4 + (mixed ? 16 : 0)), &slot);
  hipLaunchKernelGGL(k_operator, col_grid_k<kOpK>(D.I, D.J, D.K), dim3(64, 4, 1), 0, c->stream, O);
  timer_end(c, slot);
  {
    // one partial per 256 cells: fold them on many workgroups first
    const uint32_t nb = col_blocks_k<kOpK>(D.I, D.J, D.K), chunk = 512;
    const uint32_t nf = (nb + chunk - 1) / chunk;
    const RedSite fold = red_site(c, kSiteFold);
    hipLaunchKernelGGL((k_fold_partials<6, 0x14u>), dim3(nf), dim3(256), 0, c->stream, O.site, fold,
                       nb, chunk);
    O.site = fold;
    hipLaunchKernelGGL(k_operator_fin, dim3(1), dim3(kFinThreads), 0, c->stream, O, nf);
  }
  return SP_OK;
}

}  // namespace sp
