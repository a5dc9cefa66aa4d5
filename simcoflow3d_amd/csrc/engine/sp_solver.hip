// Multigrid-preconditioned flexible CG on the GPU.
//
// Replaces HYPRE PCG + BoomerAMG (ProjectionP.f90:95-116, 267-299).  The
// outer CG, its dot products and the residual are FP64; the V-cycle runs in
// T = float (mixed) or double (FP64).  CG scalars stay in device memory: the
// host only polls a convergence flag the finalize kernel writes into pinned
// host memory, one iteration behind.
//
// Multigrid: cell-centred coarsening by 2, piecewise-constant transfers
// (restriction = sum over children, prolongation = injection), coarse face
// coefficients = scale * sum of the fine face coefficients they cover (the
// Galerkin product for these transfers, scaled), red-black Gauss-Seidel
// (pre: red,black; post: black,red -- a symmetric V-cycle).  Levels holding
// at most kTailCells cells run in one single-workgroup kernel.
//
// Grid kernels run one cell per thread on 64x4 workgroups whose order is
// remapped so that each XCD owns a contiguous slab (stencil halos stay in one
// L2); reductions store one partial per workgroup and a one-workgroup
// finalize kernel folds them in a fixed order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "sp_device.h"
#include "sp_geom.h"
#include "sp_internal.h"
#include "sp_smooth_fused.h"

namespace sp {
namespace {

template <typename T>
struct LevelView {
  int nx, ny, nz;
  const T *cx, *cy, *cz, *e;
  T* x;
  T* b;
  const uint8_t* act;  // active mask (level 0 only), else nullptr
};

template <typename T>
__device__ inline T diag_of(const LevelView<T>& L, int i, int j, int k, int64_t n) {
  const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
  T d = i > 0 ? L.cx[n - 1] : T(0);
  d = d + (j > 0 ? L.cy[n - sy] : T(0));
  d = d + (k > 0 ? L.cz[n - sz] : T(0));
  d = d + L.cx[n];
  d = d + L.cy[n];
  d = d + L.cz[n];
  if (L.e) d = d + L.e[n];
  return d;
}

// sum_nb c_f x_nb
template <typename T>
__device__ inline T offdiag_sum(const LevelView<T>& L, int i, int j, int k, int64_t n) {
  const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
  T s = T(0);
  if (i > 0) s += L.cx[n - 1] * L.x[n - 1];
  if (j > 0) s += L.cy[n - sy] * L.x[n - sy];
  if (k > 0) s += L.cz[n - sz] * L.x[n - sz];
  if (i < L.nx - 1) s += L.cx[n] * L.x[n + 1];
  if (j < L.ny - 1) s += L.cy[n] * L.x[n + sy];
  if (k < L.nz - 1) s += L.cz[n] * L.x[n + sz];
  return s;
}

// One colour of Gauss-Seidel; one thread per cell pair along i (the pair
// holds one cell of each colour).  zero_init: x starts at 0 (the other colour
// is written as 0).  FINE only distinguishes the finest level in profiles.
template <typename T, bool FINE>
__global__ void __launch_bounds__(256) k_gs(LevelView<T> L, int color, int zero_init, T omega,
                                            const Scalars* S) {
  if (S->done) return;
  const int nxh = (L.nx + 1) >> 1;
  const Cell3 q = cell_of(nxh, L.ny, L.nz);
  if (!q.in) return;
  const int j = q.b, k = q.c;
  const int par = (j + k + color) & 1;
  const int i = 2 * q.a + par;
  const int64_t row = (int64_t)L.nx * (j + (int64_t)L.ny * k);
  if (zero_init) {
    const int io = 2 * q.a + (par ^ 1);
    if (io < L.nx) L.x[row + io] = T(0);
  }
  if (i >= L.nx) return;
  const int64_t n = row + i;
  const T d = diag_of(L, i, j, k, n);
  if (!(d > T(0))) {
    if (zero_init) L.x[n] = T(0);
    return;
  }
  T s = L.b[n];
  if (!zero_init) s += offdiag_sum(L, i, j, k, n);
  // same rounding as the fused smoother (FP32: times 1/d, FP64: divided by d),
  // over-relaxed by omega
  const T v = sizeof(T) == 4 ? s * smoother_rcp(d) : s / d;
  const T x0 = zero_init ? T(0) : L.x[n];
  L.x[n] = x0 + omega * (v - x0);
}

// One step of the Chebyshev semi-iteration for D^-1 A (Saad, Iterative
// Methods for Sparse Linear Systems, 2nd ed., Algorithm 12.1):
//   d <- c1 d + c2 D^-1 (b - A x_in),  x_out = x_in + d
// (c1 unused on the first step; x_in == nullptr stands for x_in = 0).
// x_out is another buffer than x_in: every lane reads its neighbours' x_in.
template <typename T, bool FINE>
__global__ void __launch_bounds__(256) k_cheb(LevelView<T> L, const T* __restrict__ xin,
                                              T* __restrict__ xout, T* __restrict__ dv, T c1,
                                              T c2, int first, const Scalars* S) {
  if (S->done) return;
  const Cell3 q = cell_of(L.nx, L.ny, L.nz);
  if (!q.in) return;
  const int i = q.a, j = q.b, k = q.c;
  const int64_t n = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
  const T d = diag_of(L, i, j, k, n);
  if (!(d > T(0))) {
    xout[n] = T(0);
    dv[n] = T(0);
    return;
  }
  T r = L.b[n];
  T x0 = T(0);
  if (xin) {
    LevelView<T> V = L;
    V.x = const_cast<T*>(xin);
    x0 = xin[n];
    r = r - (d * x0 - offdiag_sum(V, i, j, k, n));
  }
  const T z = sizeof(T) == 4 ? r * smoother_rcp(d) : r / d;
  const T dn = first ? c2 * z : c1 * dv[n] + c2 * z;
  dv[n] = dn;
  xout[n] = x0 + dn;
}

// Coarse b = sum over children of the fine residual.
template <typename T, bool FINE>
__global__ void __launch_bounds__(256) k_resid_restrict(LevelView<T> F, LevelView<T> C,
                                                        const Scalars* S) {
  if (S->done) return;
  const Cell3 q = cell_of(C.nx, C.ny, C.nz);
  if (!q.in) return;
  T s = T(0);
#pragma unroll
  for (int cc = 0; cc < 2; ++cc)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int aa = 0; aa < 2; ++aa) {
        const int fi = 2 * q.a + aa, fj = 2 * q.b + bb, fk = 2 * q.c + cc;
        if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
        const int64_t m = fi + (int64_t)F.nx * (fj + (int64_t)F.ny * fk);
        const T d = diag_of(F, fi, fj, fk, m);
        if (d > T(0)) s += F.b[m] - (d * F.x[m] - offdiag_sum(F, fi, fj, fk, m));
      }
  C.b[q.a + (int64_t)C.nx * (q.b + (int64_t)C.ny * q.c)] = s;
}

template <typename T, bool FINE>
__global__ void __launch_bounds__(256) k_prolong(LevelView<T> F, LevelView<T> C, const Scalars* S) {
  if (S->done) return;
  const Cell3 q = cell_of(F.nx, F.ny, F.nz);
  if (!q.in) return;
  const int64_t m = q.a + (int64_t)F.nx * (q.b + (int64_t)F.ny * q.c);
  const bool act = F.act ? F.act[m] != 0 : diag_of(F, q.a, q.b, q.c, m) > T(0);
  if (act) F.x[m] += C.x[(q.a >> 1) + (int64_t)C.nx * ((q.b >> 1) + (int64_t)C.ny * (q.c >> 1))];
}

// ---- single-workgroup V-cycle over the small levels ----

// ---- coarse levels in one workgroup, resident in LDS ----

// One tail level in LDS: couplings, diagonal and its reciprocal, b and x.
template <typename T>
struct TailLevel {
  int nx, ny, nz;
  T *cx, *cy, *cz, *d, *rd, *b, *x;
};
constexpr int kTailArrays = 7;

template <typename T>
struct TailArgs {
  int first, last;  // levels [first, last)
  int pre, post, coarse_sweeps, singular, has_e;
  T omega, omega_c;  // over-relaxation of the Gauss-Seidel updates (level 0, coarser)
  const LevelView<T>* lv;   // the levels in global memory
  int off[kMaxLevels];      // element offset of each tail level in LDS
};

template <typename T>
__device__ inline TailLevel<T> tail_level(T* lds, const TailArgs<T>& A, int l) {
  const LevelView<T>& G = A.lv[l];
  const int n = G.nx * G.ny * G.nz;
  T* base = lds + A.off[l];
  return TailLevel<T>{G.nx, G.ny, G.nz, base, base + n, base + 2 * n, base + 3 * n,
                      base + 4 * n, base + 5 * n, base + 6 * n};
}

// (i, j, k) of cell n.  Tail levels hold far fewer than 2^24 cells, so the
// quotients come from a float reciprocal, corrected by one step (the float
// estimate is within 1 of the integer quotient).
struct Ijk {
  int i, j, k;
};
__device__ inline unsigned small_div(unsigned n, unsigned d) {
  unsigned q = (unsigned)((float)n * __builtin_amdgcn_rcpf((float)d));
  if (q * d > n) --q;
  else if ((q + 1) * d <= n) ++q;
  return q;
}
__device__ inline Ijk ijk_of(unsigned n, unsigned nx, unsigned nxy) {
  const unsigned k = small_div(n, nxy), r = n - k * nxy, j = small_div(r, nx);
  return Ijk{(int)(r - j * nx), (int)j, (int)k};
}

template <typename T>
__device__ inline T tail_offdiag(const TailLevel<T>& L, int i, int j, int k, int n) {
  const int sy = L.nx, sz = L.nx * L.ny;
  T s = T(0);
  if (i > 0) s += L.cx[n - 1] * L.x[n - 1];
  if (j > 0) s += L.cy[n - sy] * L.x[n - sy];
  if (k > 0) s += L.cz[n - sz] * L.x[n - sz];
  if (i < L.nx - 1) s += L.cx[n] * L.x[n + 1];
  if (j < L.ny - 1) s += L.cy[n] * L.x[n + sy];
  if (k < L.nz - 1) s += L.cz[n] * L.x[n + sz];
  return s;
}

// The threads that work on a tail level: the whole workgroup, or wave 0
// alone on the smallest levels, which then need no workgroup barrier (one
// wave's LDS operations complete in issue order; the fences keep the
// compiler from moving LDS accesses across the step boundary).
template <bool WAVE>
struct TailTeam {
  __device__ static unsigned id() { return WAVE ? (threadIdx.x & 63u) : threadIdx.x; }
  __device__ static unsigned size() { return WAVE ? 64u : blockDim.x; }
  __device__ static void sync() {
    if (WAVE) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  }
};
constexpr int kTailWaveCells = 64;  // levels this small run on one wave

// One colour of Gauss-Seidel (one thread per cell pair along i); zero_init:
// x starts at 0 and the other colour is written as 0.  FP32 multiplies by
// 1/d, FP64 divides by d, as the other smoothers do.
template <bool WAVE, typename T>
__device__ void tail_gs(const TailLevel<T>& L, int color, int zero_init, T omega) {
  using G = TailTeam<WAVE>;
  const unsigned nxh = (unsigned)(L.nx + 1) / 2, nxy = nxh * (unsigned)L.ny;
  const unsigned half = nxy * (unsigned)L.nz;
  for (unsigned t = G::id(); t < half; t += G::size()) {
    const Ijk q = ijk_of(t, nxh, nxy);
    const int par = (q.j + q.k + color) & 1;
    const int i = 2 * q.i + par;
    const int row = L.nx * (q.j + L.ny * q.k);
    if (zero_init) {
      const int io = 2 * q.i + (par ^ 1);
      if (io < L.nx) L.x[row + io] = T(0);
    }
    if (i >= L.nx) continue;
    const int n = row + i;
    const T d = L.d[n];
    if (!(d > T(0))) {
      if (zero_init) L.x[n] = T(0);
      continue;
    }
    T s = L.b[n];
    if (!zero_init) s += tail_offdiag(L, i, q.j, q.k, n);
    const T v = sizeof(T) == 4 ? s * L.rd[n] : s / d;
    const T x0 = zero_init ? T(0) : L.x[n];
    L.x[n] = x0 + omega * (v - x0);
  }
  G::sync();
}

// v -= mean of v over the cells with d > 0
template <bool WAVE, typename T>
__device__ void tail_project_mean(const TailLevel<T>& L, T* v) {
  using G = TailTeam<WAVE>;
  __shared__ double red[2][16];
  double s = 0.0, c = 0.0;
  const int nn = L.nx * L.ny * L.nz;
  for (int n = (int)G::id(); n < nn; n += (int)G::size()) {
    if (L.d[n] > T(0)) { s += (double)v[n]; c += 1.0; }
  }
  s = wave_sum(s);
  c = wave_sum(c);
  double ts = 0.0, tc = 0.0;
  if (WAVE) {
    ts = __shfl(s, 0, 64);
    tc = __shfl(c, 0, 64);
  } else {
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = s; red[1][threadIdx.x >> 6] = c; }
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { ts += red[0][w]; tc += red[1][w]; }
  }
  const T mean = tc > 0 ? (T)(ts / tc) : T(0);
  G::sync();
  for (int n = (int)G::id(); n < nn; n += (int)G::size()) {
    if (L.d[n] > T(0)) v[n] -= mean;
  }
  G::sync();
}

// coarse b = sum over the 2x2x2 children of the fine residual
template <bool WAVE, typename T>
__device__ void tail_restrict(const TailLevel<T>& F, const TailLevel<T>& C) {
  using G = TailTeam<WAVE>;
  const unsigned nxy = (unsigned)C.nx * C.ny, nc = nxy * (unsigned)C.nz;
  for (unsigned t = G::id(); t < nc; t += G::size()) {
    const Ijk q = ijk_of(t, (unsigned)C.nx, nxy);
    T s = T(0);
    for (int cc = 0; cc < 2; ++cc)
      for (int bb = 0; bb < 2; ++bb)
        for (int aa = 0; aa < 2; ++aa) {
          const int fi = 2 * q.i + aa, fj = 2 * q.j + bb, fk = 2 * q.k + cc;
          if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
          const int m = fi + F.nx * (fj + F.ny * fk);
          const T d = F.d[m];
          if (d > T(0)) s += F.b[m] - (d * F.x[m] - tail_offdiag(F, fi, fj, fk, m));
        }
    C.b[t] = s;
  }
  G::sync();
}

template <bool WAVE, typename T>
__device__ void tail_prolong(const TailLevel<T>& F, const TailLevel<T>& C) {
  using G = TailTeam<WAVE>;
  const unsigned nxy = (unsigned)F.nx * F.ny, nf = nxy * (unsigned)F.nz;
  for (unsigned m = G::id(); m < nf; m += G::size()) {
    const Ijk q = ijk_of(m, (unsigned)F.nx, nxy);
    if (F.d[m] > T(0))
      F.x[m] += C.x[(q.i >> 1) + C.nx * ((q.j >> 1) + C.ny * (q.k >> 1))];
  }
  G::sync();
}

template <bool WAVE, typename T>
__device__ void tail_pre(T* lds, const TailArgs<T>& A, int l) {
  const TailLevel<T> F = tail_level(lds, A, l);
  const T om = l == 0 ? A.omega : A.omega_c;
  tail_gs<WAVE>(F, 0, 1, om);
  tail_gs<WAVE>(F, 1, 0, om);
  for (int s = 1; s < A.pre; ++s) { tail_gs<WAVE>(F, 0, 0, om); tail_gs<WAVE>(F, 1, 0, om); }
  tail_restrict<WAVE>(F, tail_level(lds, A, l + 1));
}

template <bool WAVE, typename T>
__device__ void tail_coarsest(T* lds, const TailArgs<T>& A) {
  const TailLevel<T> C = tail_level(lds, A, A.last - 1);
  const T om = A.last - 1 == 0 ? A.omega : A.omega_c;
  if (A.singular) tail_project_mean<WAVE>(C, C.b);
  tail_gs<WAVE>(C, 0, 1, om);
  tail_gs<WAVE>(C, 1, 0, om);
  for (int s = 1; s < A.coarse_sweeps; ++s) { tail_gs<WAVE>(C, 0, 0, om); tail_gs<WAVE>(C, 1, 0, om); }
  for (int s = 0; s < A.coarse_sweeps; ++s) { tail_gs<WAVE>(C, 1, 0, om); tail_gs<WAVE>(C, 0, 0, om); }
  if (A.singular) tail_project_mean<WAVE>(C, C.x);
}

template <bool WAVE, typename T>
__device__ void tail_post(T* lds, const TailArgs<T>& A, int l) {
  const TailLevel<T> F = tail_level(lds, A, l);
  const T om = l == 0 ? A.omega : A.omega_c;
  tail_prolong<WAVE>(F, tail_level(lds, A, l + 1));
  for (int s = 0; s < A.post; ++s) { tail_gs<WAVE>(F, 1, 0, om); tail_gs<WAVE>(F, 0, 0, om); }
}

// The coarse levels of the V-cycle in one workgroup, resident in LDS: the
// couplings of every tail level and the right-hand side of the first are
// copied in (the diagonal and 1/d are formed once), the whole sub-cycle
// runs on LDS, and x of the first tail level is written back for the
// prolongation above it.
template <typename T>
__global__ void __launch_bounds__(1024) k_tail(TailArgs<T> A, const Scalars* S) {
  if (S->done) return;
  __shared__ double lds_raw[kTailLdsBytes / sizeof(double)];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const int L = A.last;
  for (int l = A.first; l < L; ++l) {
    const LevelView<T> G = A.lv[l];
    const TailLevel<T> V = tail_level(lds, A, l);
    const int n = G.nx * G.ny * G.nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      V.cx[q] = G.cx[q];
      V.cy[q] = G.cy[q];
      V.cz[q] = G.cz[q];
      V.d[q] = A.has_e ? G.e[q] : T(0);  // diagonal-only term, completed below
      if (l == A.first) V.b[q] = G.b[q];
    }
  }
  __syncthreads();
  for (int l = A.first; l < L; ++l) {
    const TailLevel<T> V = tail_level(lds, A, l);
    const unsigned nxy = (unsigned)V.nx * V.ny, n = nxy * (unsigned)V.nz;
    for (unsigned m = threadIdx.x; m < n; m += blockDim.x) {
      const Ijk q = ijk_of(m, (unsigned)V.nx, nxy);
      // same order as the other kernels: W, S, B, E, N, T, then e
      T d = q.i > 0 ? V.cx[m - 1] : T(0);
      d = d + (q.j > 0 ? V.cy[m - V.nx] : T(0));
      d = d + (q.k > 0 ? V.cz[m - nxy] : T(0));
      d = d + V.cx[m];
      d = d + V.cy[m];
      d = d + V.cz[m];
      d = d + V.d[m];
      V.d[m] = d;
      V.rd[m] = d > T(0) ? T(1) / d : T(0);
    }
  }
  __syncthreads();
  // levels [first, lw) on the whole workgroup, [lw, L) on wave 0 alone
  int lw = L;
  for (int l = A.first; l < L; ++l) {
    const LevelView<T>& G = A.lv[l];
    if (G.nx * G.ny * G.nz <= kTailWaveCells) { lw = l; break; }
  }
  for (int l = A.first; l < lw && l < L - 1; ++l) tail_pre<false>(lds, A, l);
  if (lw >= L) {
    tail_coarsest<false>(lds, A);
  } else {
    if (threadIdx.x < 64) {
      for (int l = lw; l < L - 1; ++l) tail_pre<true>(lds, A, l);
      tail_coarsest<true>(lds, A);
      for (int l = L - 2; l >= lw; --l) tail_post<true>(lds, A, l);
    }
    __syncthreads();
  }
  for (int l = (lw >= L ? L - 2 : lw - 1); l >= A.first; --l) tail_post<false>(lds, A, l);
  {
    const LevelView<T> G = A.lv[A.first];
    const TailLevel<T> V = tail_level(lds, A, A.first);
    const int n = G.nx * G.ny * G.nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) G.x[q] = V.x[q];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_coarsen(LevelView<T> F, LevelView<T> C, T* ccx, T* ccy,
                                                 T* ccz, T* ce, T scale) {
  const Cell3 q = cell_of(C.nx, C.ny, C.nz);
  if (!q.in) return;
  T sx = 0, sy = 0, sz = 0, se = 0;
  for (int cc = 0; cc < 2; ++cc)
    for (int bb = 0; bb < 2; ++bb)
      for (int aa = 0; aa < 2; ++aa) {
        const int fi = 2 * q.a + aa, fj = 2 * q.b + bb, fk = 2 * q.c + cc;
        if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
        const int64_t m = fi + (int64_t)F.nx * (fj + (int64_t)F.ny * fk);
        if (F.e) se += F.e[m];
        if (aa == 1) sx += F.cx[m];
        if (bb == 1) sy += F.cy[m];
        if (cc == 1) sz += F.cz[m];
      }
  const int64_t n = q.a + (int64_t)C.nx * (q.b + (int64_t)C.ny * q.c);
  ccx[n] = scale * sx;
  ccy[n] = scale * sy;
  ccz[n] = scale * sz;
  if (ce) ce[n] = scale * se;
}

// ---- outer CG (FP64) ----

struct CgArgs {
  int I, J, K;
  const double *cx, *cy, *cz, *e;
  const uint8_t* active;
  const float* dw;
  const double *dxp, *dyp, *dzp;
  const double* b;
  double *x, *r, *q;
  double* p;  // search direction
  double* z64;  // written by the Jacobi preconditioner only
  const float* z32;
  float* r32;
  Scalars* S;
  int32_t* hflag;  // pinned host flag (device view)
  double tol;
  RedSite site;
};

__device__ inline double zval(const CgArgs& A, int64_t n) {
  return A.z32 ? (double)A.z32[n] : A.z64[n];
}

__device__ inline double apply_row(const CgArgs& A, const double* v, int a, int b, int c,
                                   int64_t n) {
  // (A v)_n in difference form: sum of c_f (v_n - v_nb) plus the
  // diagonal-only terms.  Neighbouring values of an iterate with a large
  // level (a pressure fixed by a Dirichlet face) are close, so their
  // differences are exact and the result is accurate to the size of the
  // fluxes rather than to that of the level.
  const int64_t IJ = (int64_t)A.I * A.J;
  const double vn = v[n];
  double s = A.e ? A.e[n] * vn : 0.0;
  if (a > 0) s += A.cx[n - 1] * (vn - v[n - 1]);
  if (b > 0) s += A.cy[n - A.I] * (vn - v[n - A.I]);
  if (c > 0) s += A.cz[n - IJ] * (vn - v[n - IJ]);
  s += a < A.I - 1 ? A.cx[n] * (vn - v[n + 1]) : A.cx[n] * vn;
  s += b < A.J - 1 ? A.cy[n] * (vn - v[n + A.I]) : A.cy[n] * vn;
  s += c < A.K - 1 ? A.cz[n] * (vn - v[n + IJ]) : A.cz[n] * vn;
  return s;
}

// Residual of the refined pressure x + xd of a non-singular system:
// rt = b - A x - A xd on active cells, in difference form; max |rt|/V.
__global__ void __launch_bounds__(256) k_refine_residual(CgArgs A, const double* xd, double* rt) {
  const Col3 q = col_of(A.I, A.J, A.K);
  double v[1] = {0.0};
  if (q.in) {
    const int64_t IJ = (int64_t)A.I * A.J;
    const double dxy = A.dxp[q.a] * A.dyp[q.b];
    int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
    for (int c = q.c0; c < q.c1; ++c, n += IJ) {
      double rv = 0.0;
      if (A.active[n]) {
        rv = A.b[n] - apply_row(A, A.x, q.a, q.b, c, n);
        rv -= apply_row(A, xd, q.a, q.b, c, n);
        const double rr = fabs(rv) / (dxy * A.dzp[c]);
        v[0] = (rr > v[0] || rr != rr) ? rr : v[0];
      }
      rt[n] = rv;
    }
  }
  block_partials<kTileThreads, 1, 1u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_refine_residual_fin(CgArgs A, uint32_t nb) {
  double out[1];
  finalize_partials<kFinThreads, 1, 1u>(A.site, nb, out);
  if (flat_tid() == 0) A.S->npart[0] = out[0];
}

// y += x
__global__ void __launch_bounds__(256) k_refine_add(CgArgs A, const double* x, double* y) {
  const Col3 q = col_of(A.I, A.J, A.K);
  if (!q.in) return;
  const int64_t IJ = (int64_t)A.I * A.J;
  int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
  for (int c = q.c0; c < q.c1; ++c, n += IJ) y[n] += x[n];
}

// Jacobi preconditioner (SP_PRECOND_JACOBI): z = r / diagonal on active cells
__global__ void __launch_bounds__(256) k_cg_jacobi(CgArgs A) {
  if (A.S->done) return;
  const Col3 q = col_of(A.I, A.J, A.K);
  if (!q.in) return;
  const int64_t IJ = (int64_t)A.I * A.J;
  int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
  for (int c = q.c0; c < q.c1; ++c, n += IJ) {
    double d = q.a > 0 ? A.cx[n - 1] : 0.0;
    d = d + (q.b > 0 ? A.cy[n - A.I] : 0.0);
    d = d + (c > 0 ? A.cz[n - IJ] : 0.0);
    d = d + A.cx[n];
    d = d + A.cy[n];
    d = d + A.cz[n];
    if (A.e) d = d + A.e[n];
    A.z64[n] = (A.active[n] && d > 0.0) ? A.r[n] / d : 0.0;
  }
}

// r = b - mean(b) - A x  (x = 0 when from_zero); resets p and q.  KC cells
// of a column per thread, the mask and b of all of them loaded first (cell
// index clamped into the column).
template <int KC>
__global__ void __launch_bounds__(256) k_cg_residual(CgArgs A, int from_zero) {
  const Col3 q = col_of<KC>(A.I, A.J, A.K);
  Scalars* S = A.S;
  const double bmean = S->bmean;
  double v[3] = {0.0, 0.0, 0.0};  // max |r|/V, non-finite, sum r
  if (q.in) {
    const int64_t IJ = (int64_t)A.I * A.J;
    const double dxy = A.dxp[q.a] * A.dyp[q.b];
    const int64_t n0 = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
    bool act[KC];
    double bv[KC], dz[KC];
#pragma unroll
    for (int cc = 0; cc < KC; ++cc) {
      const int c = min(q.c0 + cc, q.c1 - 1);
      const int64_t n = n0 + IJ * (c - q.c0);
      act[cc] = A.active[n] != 0;
      bv[cc] = A.b[n];
      dz[cc] = A.dzp[c];
    }
#pragma unroll
    for (int cc = 0; cc < KC; ++cc) {
      const int c = q.c0 + cc;
      if (c >= q.c1) break;
      const int64_t n = n0 + IJ * cc;
      double rv = 0.0;
      if (act[cc]) {
        rv = bv[cc] - bmean;
        if (!from_zero) rv -= apply_row(A, A.x, q.a, q.b, c, n);
        const double rr = fabs(rv) / (dxy * dz[cc]);
        v[0] = (rr > v[0] || rr != rr) ? rr : v[0];
        if (!is_finite(rv)) v[1] = 1.0;
        v[2] += rv;
      }
      A.r[n] = rv;
      if (A.r32) A.r32[n] = (float)rv;
      if (from_zero) A.x[n] = 0.0;
      A.q[n] = 0.0;
      A.p[n] = 0.0;
    }
  }
  block_partials<kTileThreads, 3, 3u>(A.site, v);
}
// cells per thread of k_cg_residual
constexpr int kResK = 4;

__global__ void __launch_bounds__(kFinThreads) k_cg_residual_fin(CgArgs A, uint32_t nb) {
  double out[3];
  finalize_partials<kFinThreads, 3, 3u>(A.site, nb, out);
  if (flat_tid() == 0) {
    Scalars* S = A.S;
    S->rmax = out[0];
    S->rsum = out[2];
    S->qsum = 0.0;  // q was reset
    S->nonfinite = S->nonfinite || out[1] > 0.0 || !is_finite(out[0]);
    S->done = (out[0] <= A.tol || S->nonfinite) ? 1 : 0;
    S->first = 1;
    S->alpha = 0.0;
    S->beta = 0.0;
    S->rz = 0.0;
    *A.hflag = S->done;
  }
}

// rz = (r, z), zq = (z, q);  beta = -alpha zq / rz_old (flexible CG).
// Singular systems: z is shifted by its diagonal-weighted mean m so the
// iterate keeps the pressure of the stiff (high-coefficient) phase near zero;
// (r, z-m) and (z-m, q) follow from the raw sums.
__global__ void __launch_bounds__(256) k_cg_dots(CgArgs A) {
  Scalars* S = A.S;
  if (S->done) return;
  const Col3 q = col_of(A.I, A.J, A.K);
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (q.in) {
    const int64_t IJ = (int64_t)A.I * A.J;
    int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
#pragma unroll 4
    for (int c = q.c0; c < q.c1; ++c, n += IJ) {
      if (A.active[n]) {
        const double z = zval(A, n), r = A.r[n], qq = A.q[n];
        v[0] += r * z;
        v[1] += z * qq;
        v[2] += (double)A.dw[n] * z;
        v[3] += r;
        v[4] += qq;
      }
    }
  }
  block_partials<kTileThreads, 5, 0u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_cg_dots_fin(CgArgs A, uint32_t nb) {
  Scalars* S = A.S;
  if (S->done) return;
  double t[5];
  finalize_partials<kFinThreads, 5, 0u>(A.site, nb, t);
  if (flat_tid() == 0) {
    const double m = (S->singular && S->dsum > 0.0) ? t[2] / S->dsum : 0.0;
    const double rzm = t[0] - m * t[3], zqm = t[1] - m * t[4];
    S->zshift = m;
    S->beta = S->first ? 0.0 : -S->alpha * zqm / S->rz;
    S->rz = rzm;
    S->zq = zqm;
    S->first = 0;
    if (!is_finite(rzm) || !is_finite(S->beta)) { S->nonfinite = 1; S->done = 1; }
  }
}

// Same as k_cg_dots_fin with (r,z), (z,q), (dw,z) from the finest
// post-smoother's partials and the sums of r and q from the update and the
// SpMV of the previous iteration.
__global__ void __launch_bounds__(kFinThreads) k_cg_dots3_fin(CgArgs A, uint32_t nb) {
  Scalars* S = A.S;
  if (S->done) return;
  double t[3];
  finalize_partials<kFinThreads, 3, 0u>(A.site, nb, t);
  if (flat_tid() == 0) {
    const double m = (S->singular && S->dsum > 0.0) ? t[2] / S->dsum : 0.0;
    const double rzm = t[0] - m * S->rsum, zqm = t[1] - m * S->qsum;
    S->zshift = m;
    S->beta = S->first ? 0.0 : -S->alpha * zqm / S->rz;
    S->rz = rzm;
    S->zq = zqm;
    S->first = 0;
    if (!is_finite(rzm) || !is_finite(S->beta)) { S->nonfinite = 1; S->done = 1; }
  }
}

// The search direction and its image in one pass:
//   x += alpha p_old            (the x half of the previous CG step)
//   p = (z - m) + beta p_old,   q = A p = A z - m e + beta q_old
// on active cells (0 elsewhere), with (p, q) and sum q for alpha.  A 1 is the
// diagonal-only term e on active rows (the face couplings of a row sum to its
// diagonal minus e), so the stencil is applied to z, which is complete in
// memory, and p and q need no neighbours.  Each thread walks its column
// keeping z and the z-coupling of planes k-1 / k / k+1 in registers.  After
// the last CG step k_cg_xflush applies its x half.
template <typename Z>
__global__ void __launch_bounds__(256) k_cg_pspmv(CgArgs A, const Z* zv) {
  Scalars* S = A.S;
  if (S->done) return;
  const Col3 q = col_of(A.I, A.J, A.K);
  const double alpha = S->alpha;
  double v[2] = {0.0, 0.0};  // (p, q), sum q
  if (q.in) {
    const double beta = S->beta, m = S->zshift;
    const int64_t I = A.I, IJ = (int64_t)A.I * A.J;
    const int a = q.a, b = q.b;
    const int64_t dW = a > 0 ? 1 : 0, dE = a < A.I - 1 ? 1 : 0;
    const int64_t dS = b > 0 ? I : 0, dN = b < A.J - 1 ? I : 0;
    int64_t n = a + I * (b + (int64_t)A.J * q.c0);
    double zm = 0.0, czm = 0.0;
    if (q.c0 > 0) {
      zm = (double)zv[n - IJ];
      czm = A.cz[n - IJ];
    }
    double zc = (double)zv[n];
    for (int c = q.c0; c < q.c1; ++c, n += IJ) {
      const double zp = c < A.K - 1 ? (double)zv[n + IJ] : 0.0;
      const double cxm = a > 0 ? A.cx[n - dW] : 0.0;
      const double cym = b > 0 ? A.cy[n - dS] : 0.0;
      const double cx = A.cx[n], cy = A.cy[n], cz = A.cz[n];
      const double zw = (double)zv[n - dW], ze = (double)zv[n + dE];
      const double zs = (double)zv[n - dS], zn = (double)zv[n + dN];
      const double en = A.e ? A.e[n] : 0.0;
      double d = cxm + cym;
      d = d + czm;
      d = d + cx;
      d = d + cy;
      d = d + cz;
      d = d + en;
      double az = d * zc;
      az -= cxm * zw;
      az -= cym * zs;
      az -= czm * zm;
      az -= cx * ze;
      az -= cy * zn;
      az -= cz * zp;
      const bool act = A.active[n] != 0;
      const double po = A.p[n];
      A.x[n] += alpha * po;
      const double pv = act ? (zc - m) + beta * po : 0.0;
      const double qv = act ? (az - m * en) + beta * A.q[n] : 0.0;
      A.p[n] = pv;
      A.q[n] = qv;
      v[0] += pv * qv;
      v[1] += qv;
      zm = zc;
      zc = zp;
      czm = cz;
    }
  }
  block_partials<kTileThreads, 2, 0u>(A.site, v);
}

__global__ void __launch_bounds__(kFinThreads) k_cg_spmv_fin(CgArgs A, uint32_t nb) {
  Scalars* S = A.S;
  if (S->done) return;
  double t[2];
  finalize_partials<kFinThreads, 2, 0u>(A.site, nb, t);
  if (flat_tid() == 0) {
    S->pq = t[0];
    S->qsum = t[1];
    S->alpha = S->rz / t[0];
    if (!is_finite(S->alpha)) { S->nonfinite = 1; S->done = 1; }
  }
}

// r -= alpha q ; rmax = max |r|/V (x += alpha p follows in the next
// k_cg_pspmv, or in k_cg_xflush after the last step)
__global__ void __launch_bounds__(256) k_cg_update(CgArgs A) {
  Scalars* S = A.S;
  if (S->done) return;
  const Col3 q = col_of(A.I, A.J, A.K);
  const double alpha = S->alpha;
  double v[2] = {0.0, 0.0};  // max |r|/V, sum r
  if (q.in) {
    const int64_t IJ = (int64_t)A.I * A.J;
    const double dxy = A.dxp[q.a] * A.dyp[q.b];
    int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
#pragma unroll 4
    for (int c = q.c0; c < q.c1; ++c, n += IJ) {
      const double rv = A.r[n] - alpha * A.q[n];
      A.r[n] = rv;
      if (A.r32) A.r32[n] = (float)rv;
      const double rr = fabs(rv) / (dxy * A.dzp[c]);
      v[0] = (rr > v[0] || rr != rr) ? rr : v[0];  // a NaN sticks
      v[1] += rv;
    }
  }
  block_partials<kTileThreads, 2, 1u>(A.site, v);
}

// x += alpha p of the last CG step (its r update is done; the k_cg_pspmv
// that would apply it is never launched, or returns once the loop is done).
// alpha = 0 and p = 0 after a residual that already meets the tolerance.
__global__ void __launch_bounds__(256) k_cg_xflush(CgArgs A) {
  const Scalars* S = A.S;
  const Col3 q = col_of(A.I, A.J, A.K);
  if (!q.in) return;
  const double alpha = S->alpha;
  const int64_t IJ = (int64_t)A.I * A.J;
  int64_t n = q.a + (int64_t)A.I * (q.b + (int64_t)A.J * q.c0);
#pragma unroll 4
  for (int c = q.c0; c < q.c1; ++c, n += IJ) A.x[n] += alpha * A.p[n];
}

__global__ void __launch_bounds__(kFinThreads) k_cg_update_fin(CgArgs A, uint32_t nb) {
  Scalars* S = A.S;
  if (S->done) return;
  double t[2];
  finalize_partials<kFinThreads, 2, 1u>(A.site, nb, t);
  if (flat_tid() == 0) {
    S->rmax = t[0];
    S->rsum = t[1];
    S->iter += 1;
    if (!is_finite(t[0])) S->nonfinite = 1;
    S->done = (t[0] <= A.tol || S->nonfinite) ? 1 : 0;
    *A.hflag = S->done;
  }
}

template <typename T>
LevelView<T> view_of(const Level<T>& L, const uint8_t* act = nullptr) {
  return LevelView<T>{L.nx, L.ny, L.nz, L.cx, L.cy, L.cz, L.e, L.x, L.b, act};
}

template <typename T>
double level_bytes(const Level<T>& L, double per_cell_words) {
  return (double)L.n * sizeof(T) * per_cell_words;
}

template <typename T>
void launch_gs(sp_ctx* c, const Level<T>& L, int color, int zero_init, bool fine) {
  const int nxh = (L.nx + 1) / 2;
  dim3 blk(64, 4, 1), grd = flat_grid(nxh, L.ny, L.nz);
  int slot;
  // algorithmic words per cell: b for one colour, the three face arrays,
  // x read (all) and written (one colour); a zero-init sweep writes all of x
  const double words = (zero_init ? 0.5 + 3.0 + 1.0 : 1.0 + 0.5 + 3.0 + 0.5) + (L.e ? 1.0 : 0.0);
  timer_begin(c, fine ? (zero_init ? "mg_gs0_fine" : "mg_gs_fine") : "mg_gs_coarse",
              level_bytes(L, words), &slot);
  if (fine)
    hipLaunchKernelGGL((k_gs<T, true>), grd, blk, 0, c->stream, view_of(L), color, zero_init,
                       (T)(fine ? c->prm.mg_omega : c->prm.mg_omega_coarse),
                       (const Scalars*)c->S);
  else
    hipLaunchKernelGGL((k_gs<T, false>), grd, blk, 0, c->stream, view_of(L), color, zero_init,
                       (T)(fine ? c->prm.mg_omega : c->prm.mg_omega_coarse),
                       (const Scalars*)c->S);
  timer_end(c, slot);
}

// m steps of the Chebyshev smoother on level L, from zero (pre) or from L.x
// (post); the result ends in L.x (L.xs is the other iterate buffer).
template <typename T>
void launch_cheb(sp_ctx* c, const Level<T>& L, int m, bool from_zero, bool fine) {
  constexpr double lmax = 2.0, lmin = lmax / 8.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  dim3 blk(64, 4, 1), grd = flat_grid(L.nx, L.ny, L.nz);
  const LevelView<T> V = view_of(L);
  // the last step writes L.x; a step never writes the buffer it reads
  T* buf[2] = {L.x, L.xs};
  int w = from_zero ? (m - 1) % 2 : 1;  // buffer of the first step's output
  const T* in = from_zero ? nullptr : L.x;
  double rho = 1.0 / sigma;
  int slot;
  for (int s = 0; s < m; ++s) {
    double c1 = 0.0, c2 = 1.0 / theta;
    if (s > 0) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      c1 = rn * rho;
      c2 = 2.0 * rn / delta;
      rho = rn;
    }
    // words per cell: b and the three face arrays read, x in read (not on a
    // step from zero), d read after the first step, e read where present;
    // x out and d written
    const double words = 6.0 + (in ? 1.0 : 0.0) + (s > 0 ? 1.0 : 0.0) + (L.e ? 1.0 : 0.0);
    timer_begin(c, fine ? "mg_cheb_fine" : "mg_cheb_coarse", level_bytes(L, words), &slot);
    if (fine)
      hipLaunchKernelGGL((k_cheb<T, true>), grd, blk, 0, c->stream, V, in, buf[w], L.cd, (T)c1,
                         (T)c2, s == 0 ? 1 : 0, (const Scalars*)c->S);
    else
      hipLaunchKernelGGL((k_cheb<T, false>), grd, blk, 0, c->stream, V, in, buf[w], L.cd, (T)c1,
                         (T)c2, s == 0 ? 1 : 0, (const Scalars*)c->S);
    timer_end(c, slot);
    in = buf[w];
    w ^= 1;
  }
  if (in != L.x)  // post with an odd m: the result is in L.xs
    hipMemcpyAsync(L.x, in, sizeof(T) * (size_t)L.n, hipMemcpyDeviceToDevice, c->stream);
}

// One smoother leg with extended tiles of EX x EY threads; per_cu: how many
// such workgroups a CU holds at once (registers and LDS).
template <typename T, int NU, int EX, int EY>
void launch_fused_tile(sp_ctx* c, const Level<T>& L, FusedArgs<T>& A, bool fine, int mode,
                       int per_cu) {
  using Sh = FusedShape<NU, EX, EY>;
  const int gx = (L.nx + Sh::TX - 1) / Sh::TX, gy = (L.ny + Sh::TY - 1) / Sh::TY;
  // k chunk: even (coarse cells never straddle chunks).  A workgroup marches
  // kchunk + 2H planes one after the other, so pick the chunk that minimises
  // rounds-of-workgroups x planes marched.
  {
    const int64_t tiles = (int64_t)gx * gy, slots = (int64_t)std::max(1, c->num_cu) * per_cu;
    const int kmax = (L.nz + 1) & ~1;
    int best = kmax;
    int64_t best_cost = -1;
    for (int kc = 2; kc <= kmax; kc += 2) {
      const int64_t blocks = tiles * ((L.nz + kc - 1) / kc);
      const int64_t cost = ((blocks + slots - 1) / slots) * (kc + 2 * Sh::H);
      if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = kc; }
    }
    A.kchunk = best;
  }
  const dim3 grd(gx, gy, (L.nz + A.kchunk - 1) / A.kchunk), blk(EX, EY, 1);
  int slot;
  if (A.dots) c->dots_nb = grd.x * grd.y * grd.z;
  // algorithmic words per cell: pre b + 3 faces -> x (+ coarse b);
  // post x_pre + 3 faces + b -> x (+ coarse x); the dot products add q
  // (FP64)
  const double words = mode == 0 ? 5.0 + 0.125 : 6.0 + 0.125;
  timer_begin(c, fine ? (mode == 0 ? "mg_pre_fine" : "mg_post_fine")
                      : (mode == 0 ? "mg_pre_coarse" : "mg_post_coarse"),
              level_bytes(L, words + (L.e ? 1.0 : 0.0)) + (A.dots ? 8.0 * (double)L.n : 0.0),
              &slot);
  const Scalars* S = c->S;
#define SP_FUSED_LAUNCH(MODE_, FINE_)                                                         \
  do {                                                                                        \
    if (A.e)                                                                                  \
      hipLaunchKernelGGL((k_smooth_fused<T, NU, MODE_, FINE_, true, EX, EY>), grd, blk, 0,    \
                         c->stream, A, S);                                                    \
    else                                                                                      \
      hipLaunchKernelGGL((k_smooth_fused<T, NU, MODE_, FINE_, false, EX, EY>), grd, blk, 0,   \
                         c->stream, A, S);                                                    \
  } while (0)
  if (mode == 0) {
    if (fine) SP_FUSED_LAUNCH(0, true);
    else SP_FUSED_LAUNCH(0, false);
  } else {
    if (fine) SP_FUSED_LAUNCH(1, true);
    else SP_FUSED_LAUNCH(1, false);
  }
#undef SP_FUSED_LAUNCH
  timer_end(c, slot);
}

template <typename T, int NU>
void launch_fused(sp_ctx* c, Level<T>* lv, int l, int mode) {
  const Level<T>& L = lv[l];
  const Level<T>& C = lv[l + 1];
  FusedArgs<T> A{};
  A.nx = L.nx; A.ny = L.ny; A.nz = L.nz;
  A.cx = L.cx; A.cy = L.cy; A.cz = L.cz; A.e = L.e;
  A.b = L.b;
  A.cnx = C.nx; A.cny = C.ny;
  A.omega = (T)(l == 0 ? c->prm.mg_omega : c->prm.mg_omega_coarse);
  if (mode == 0) { A.xout = L.xs; A.bc = C.b; }
  else { A.xin = L.xs; A.xc = C.x; A.xout = L.x; }
  if (mode == 1 && l == 0 && c->fuse_dots) {
    A.dots = 1;
    A.dq = c->q;
    A.dsite = red_site(c, kSiteDots);
  }
  const bool fine = l == 0;
  // a level of at most 32 x 32 columns has two 32 x 16 tiles: its 10-step
  // marches (2 planes + the 8-plane pipeline) run on 32 CUs.  Small tiles put
  // a workgroup on every CU (32^3: 10.5 -> 7.8 us per leg); at 64 x 64 they
  // measured slower than the large ones (13.0 against 11.7 us)
  if (!fine && (int64_t)L.nx * L.ny <= 32 * 32)
    launch_fused_tile<T, NU, kFEXs, kFEYs>(c, L, A, fine, mode, 4);
  else
    launch_fused_tile<T, NU, kFEX, kFEY>(c, L, A, fine, mode, 1);
}

// The V-cycle uses the temporally blocked smoother legs
bool fused_vcycle(const sp_ctx* c) {
  const int pre = std::max(1, c->prm.mg_pre), post = c->prm.mg_post;
  return pre == post && (pre == 1 || pre == 2) &&
         !(c->prm.flags & (SP_FLAG_UNFUSED_SMOOTHER | SP_FLAG_CHEBYSHEV));
}

template <typename T>
void vcycle(sp_ctx* c, Level<T>* lv) {
  const int L = c->levels, Lt = c->tail_level;
  // levels with a pre-smoothing leg and a restriction: those above the tail,
  // or every level but the coarsest when no level fits the tail
  const int Lp = Lt < L ? Lt : L - 1;
  const int pre = std::max(1, c->prm.mg_pre), post = c->prm.mg_post;
  const bool fused = fused_vcycle(c);
  const bool cheb = (c->prm.flags & SP_FLAG_CHEBYSHEV) != 0;
  dim3 blk(64, 4, 1);
  int slot;
  for (int l = 0; l < Lp; ++l) {
    const bool fine = l == 0;
    if (fused) {
      if (pre == 2) launch_fused<T, 2>(c, lv, l, 0);
      else launch_fused<T, 1>(c, lv, l, 0);
      continue;
    }
    if (cheb) {
      launch_cheb(c, lv[l], pre, true, fine);
    } else {
      launch_gs(c, lv[l], 0, 1, fine);
      launch_gs(c, lv[l], 1, 0, fine);
      for (int s = 1; s < pre; ++s) { launch_gs(c, lv[l], 0, 0, fine); launch_gs(c, lv[l], 1, 0, fine); }
    }
    dim3 grd = flat_grid(lv[l + 1].nx, lv[l + 1].ny, lv[l + 1].nz);
    timer_begin(c, fine ? "mg_restrict_fine" : "mg_restrict_coarse",
                level_bytes(lv[l], 5.0 + (lv[l].e ? 1.0 : 0.0)) + level_bytes(lv[l + 1], 1.0), &slot);
    if (fine)
      hipLaunchKernelGGL((k_resid_restrict<T, true>), grd, blk, 0, c->stream, view_of(lv[l]),
                         view_of(lv[l + 1]), (const Scalars*)c->S);
    else
      hipLaunchKernelGGL((k_resid_restrict<T, false>), grd, blk, 0, c->stream, view_of(lv[l]),
                         view_of(lv[l + 1]), (const Scalars*)c->S);
    timer_end(c, slot);
  }
  if (Lt >= L) {
    // the coarsest level does not fit the tail's LDS: the tail's coarsest
    // sweeps (red-black pairs from x = 0, then as many black-red pairs)
    // with the global-memory kernels
    const int Lc = L - 1;
    const bool fine = Lc == 0;
    const int cs = std::max(1, c->prm.mg_coarse_sweeps);
    launch_gs(c, lv[Lc], 0, 1, fine);
    launch_gs(c, lv[Lc], 1, 0, fine);
    for (int s = 1; s < cs; ++s) { launch_gs(c, lv[Lc], 0, 0, fine); launch_gs(c, lv[Lc], 1, 0, fine); }
    for (int s = 0; s < cs; ++s) { launch_gs(c, lv[Lc], 1, 0, fine); launch_gs(c, lv[Lc], 0, 0, fine); }
  } else {
    TailArgs<T> A;
    A.first = Lt;
    A.last = L;
    A.pre = pre;
    A.post = post;
    A.coarse_sweeps = std::max(1, c->prm.mg_coarse_sweeps);
    A.omega = (T)c->prm.mg_omega;
    A.omega_c = (T)c->prm.mg_omega_coarse;
    A.singular = c->S_host->singular;
    A.has_e = lv[Lt].e ? 1 : 0;
    A.lv = reinterpret_cast<const LevelView<T>*>(c->mg_desc);
    int64_t off = 0;
    for (int l = 0; l < kMaxLevels; ++l) {
      A.off[l] = (int)off;
      if (l >= Lt && l < L) off += lv[l].n * kTailArrays;
    }
    timer_begin(c, "mg_tail", 0, &slot);
    hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(1024), 0, c->stream, A, (const Scalars*)c->S);
    timer_end(c, slot);
  }
  for (int l = Lp - 1; l >= 0; --l) {
    const bool fine = l == 0;
    if (fused) {
      if (pre == 2) launch_fused<T, 2>(c, lv, l, 1);
      else launch_fused<T, 1>(c, lv, l, 1);
      continue;
    }
    dim3 grd = flat_grid(lv[l].nx, lv[l].ny, lv[l].nz);
    const uint8_t* act = fine ? c->active : nullptr;
    timer_begin(c, fine ? "mg_prolong_fine" : "mg_prolong_coarse",
                fine ? level_bytes(lv[l], 2.0) + (double)lv[l].n + level_bytes(lv[l + 1], 1.0)
                     : level_bytes(lv[l], 5.0 + (lv[l].e ? 1.0 : 0.0)) + level_bytes(lv[l + 1], 1.0),
                &slot);
    if (fine)
      hipLaunchKernelGGL((k_prolong<T, true>), grd, blk, 0, c->stream, view_of(lv[l], act),
                         view_of(lv[l + 1]), (const Scalars*)c->S);
    else
      hipLaunchKernelGGL((k_prolong<T, false>), grd, blk, 0, c->stream, view_of(lv[l], act),
                         view_of(lv[l + 1]), (const Scalars*)c->S);
    timer_end(c, slot);
    if (cheb) {
      if (post > 0) launch_cheb(c, lv[l], post, false, fine);
    } else {
      for (int s = 0; s < post; ++s) { launch_gs(c, lv[l], 1, 0, fine); launch_gs(c, lv[l], 0, 0, fine); }
    }
  }
}

template <typename T>
sp_status setup_levels(sp_ctx* c, Level<T>* lv) {
  const int L = c->levels;
  dim3 blk(64, 4, 1);
  for (int l = 0; l < L; ++l) lv[l].e = c->has_e ? lv[l].e_store : nullptr;
  for (int l = 1; l < L; ++l) {
    Level<T>& C = lv[l];
    hipLaunchKernelGGL(k_coarsen<T>, flat_grid(C.nx, C.ny, C.nz), blk, 0, c->stream,
                       view_of(lv[l - 1]), view_of(C), C.cx, C.cy, C.cz, C.e,
                       (T)c->prm.mg_coarse_scale);
  }
  // level table for the tail kernel: it changes only with the buffers and
  // with the presence of diagonal-only terms
  if (c->desc_state != (c->has_e ? 1 : 0)) {
    std::vector<LevelView<T>> desc(kMaxLevels);
    for (int l = 0; l < L; ++l) desc[l] = view_of(lv[l]);
    hipError_t e = hipMemcpyAsync(c->mg_desc, desc.data(), sizeof(LevelView<T>) * kMaxLevels,
                                  hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(mg_desc)");
    // the host table must stay valid until the copy has been consumed
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "mg setup");
    c->desc_state = c->has_e ? 1 : 0;
  }
  return SP_OK;
}

}  // namespace

sp_status mg_setup(sp_ctx* c) {
  if (c->prm.precond != SP_PRECOND_MG) return SP_OK;
  if (c->prm.precision == SP_PREC_MIXED) return setup_levels<float>(c, c->lf);
  return setup_levels<double>(c, c->ld);
}

// Solve A x = rhs - mean (singular) with the multigrid-preconditioned
// flexible CG until max |r_i| / V_i <= tol.  The operator arrays and c->S
// (bmean, singular) come from the assembly.  Writes x into c->x.
//   check_true: after convergence of the recursive residual, form the true
//               residual and restart from it (at most 3 times) if it is
//               above tol.  The forward projection instead measures the
//               divergence of the corrected field, which is the same
//               quantity, and resumes when needed.
//   resume:     continue from the current x (true residual, iteration count
//               kept) instead of starting from x = 0.
sp_status solve(sp_ctx* c, const double* rhs, double tol, sp_stats* st, bool check_true,
                bool resume) {
  const Dims& D = c->dm;
  const bool mixed = c->prm.precision == SP_PREC_MIXED;
  const bool mg = c->prm.precond == SP_PRECOND_MG;
  const dim3 blk(64, 4, 1);
  if (flat_blocks(D.I, D.J, D.K) > c->red_max_blocks)
    return fail(c, SP_ERR_INVALID_ARG, "grid too large for the reduction buffers");
  CgArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K;
  A.cx = c->cx; A.cy = c->cy; A.cz = c->cz; A.e = c->has_e ? c->ee : nullptr;
  A.active = c->active;
  A.dw = c->dw;
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp;
  A.b = rhs;
  A.x = c->x; A.r = c->r; A.q = c->q; A.p = c->p0;
  A.S = c->S;
  A.hflag = c->hflags_dev;
  A.tol = tol;
  if (mixed && mg) {
    A.z32 = c->z32;
    A.r32 = c->r32;
  } else {
    A.z64 = c->z64;
  }
  int slot;
  const double NB = (double)D.N;
  const double eb = c->has_e ? 8.0 : 0.0;
  const double zb = (mixed && mg) ? 4.0 : 8.0;

  const uint32_t ncol = col_blocks(D.I, D.J, D.K);
  const dim3 cgrd = col_grid(D.I, D.J, D.K);
  auto launch_col = [&](auto kern, int site, const char* name, double bytes) {
    A.site = red_site(c, site);
    timer_begin(c, name, bytes, &slot);
    hipLaunchKernelGGL(kern, cgrd, blk, 0, c->stream, A);
    timer_end(c, slot);
  };
  auto finalize = [&](auto kern, int site, uint32_t nb) {
    A.site = red_site(c, site);
    timer_begin(c, "cg_finalize", 0, &slot);
    hipLaunchKernelGGL(kern, dim3(1), dim3(kFinThreads), 0, c->stream, A, nb);
    timer_end(c, slot);
  };
  auto enqueue_residual = [&](int from_zero) {
    A.site = red_site(c, kSiteResidual);
    timer_begin(c, "cg_residual", NB * (from_zero ? 8 + 1 + 8 * 4 : 8 * 5 + eb + 1 + 8 * 4),
                &slot);
    hipLaunchKernelGGL(k_cg_residual<kResK>, col_grid_k<kResK>(D.I, D.J, D.K), blk, 0, c->stream,
                       A, from_zero);
    timer_end(c, slot);
    finalize(k_cg_residual_fin, kSiteResidual, col_blocks_k<kResK>(D.I, D.J, D.K));
  };
  auto precondition = [&]() {
    if (mg) {
      if (mixed) vcycle<float>(c, c->lf);
      else vcycle<double>(c, c->ld);
    } else {
      launch_col(k_cg_jacobi, kSiteDots, "cg_jacobi", NB * (8 * 4 + eb + 1 + 8));
    }
  };

  hipError_t e;
  volatile int32_t* hflag = c->hflags;
  // the finest post-smoother forms the CG dot products when the V-cycle has
  // a fused leg on the finest level (a level above the tail, or above the
  // coarsest when no level fits the tail) and an FP32 V-cycle (the FP64
  // smoother has no registers to spare for the extra inputs)
  const int legs = c->tail_level < c->levels ? c->tail_level : c->levels - 1;
  c->fuse_dots = mg && mixed && fused_vcycle(c) && legs > 0;
  c->dots_nb = 0;
  hipEvent_t evs[2];
  hipEventCreateWithFlags(&evs[0], hipEventDisableTiming);
  hipEventCreateWithFlags(&evs[1], hipEventDisableTiming);
  int it = 0;
  if (resume) {
    it = c->S_host->iter;
    hipError_t e0 = hipMemsetAsync(&c->S->done, 0, sizeof(int32_t), c->stream);
    if (e0 != hipSuccess) return hip_check(c, e0, "hipMemsetAsync");
    enqueue_residual(0);
  } else {
    enqueue_residual(1);
  }
  int restarts = 0;
  const int max_iter = std::max(1, c->prm.max_iter);
  bool finished = false;
  while (!finished) {
    // one iteration in flight: enqueue iteration `it`, then wait for the
    // flag written at the end of iteration `it - 1`
    bool prev_pending = false;
    int cur = 0;
    while (it < max_iter) {
      precondition();
      if (c->fuse_dots) {
        finalize(k_cg_dots3_fin, kSiteDots, c->dots_nb);
      } else {
        launch_col(k_cg_dots, kSiteDots, "cg_dots", NB * (8 + zb + 8 + 4 + 1));
        finalize(k_cg_dots_fin, kSiteDots, ncol);
      }
      // z (stencil), p, q, three couplings, e, mask -> p, q
      {
        A.site = red_site(c, kSiteSpmv);
        timer_begin(c, "cg_pspmv", NB * (zb + 8 + 8 + 8 + 24 + eb + 1 + 8 + 8 + 8), &slot);
        if (A.z32)
          hipLaunchKernelGGL(k_cg_pspmv<float>, cgrd, blk, 0, c->stream, A, (const float*)A.z32);
        else
          hipLaunchKernelGGL(k_cg_pspmv<double>, cgrd, blk, 0, c->stream, A, A.z64);
        timer_end(c, slot);
      }
      finalize(k_cg_spmv_fin, kSiteSpmv, ncol);
      launch_col(k_cg_update, kSiteUpdate, "cg_update", NB * (8 * 2 + 8 + (A.r32 ? 4 : 0)));
      finalize(k_cg_update_fin, kSiteUpdate, ncol);
      hipEventRecord(evs[cur], c->stream);
      ++it;
      if (prev_pending) {
        e = hipEventSynchronize(evs[cur ^ 1]);
        if (e != hipSuccess) { hipEventDestroy(evs[0]); hipEventDestroy(evs[1]); return hip_check(c, e, "solve"); }
        if (*hflag) break;
      }
      prev_pending = true;
      cur ^= 1;
    }
    // x += alpha p of the last step
    launch_col(k_cg_xflush, kSiteUpdate, "cg_xflush", NB * 24);
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { hipEventDestroy(evs[0]); hipEventDestroy(evs[1]); return hip_check(c, e, "solve"); }
    e = hipMemcpy(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { hipEventDestroy(evs[0]); hipEventDestroy(evs[1]); return hip_check(c, e, "solve"); }
    const int iters = c->S_host->iter;
    if (c->S_host->nonfinite || !check_true) break;
    // true residual of the final iterate
    hipMemsetAsync(&c->S->done, 0, sizeof(int32_t), c->stream);
    enqueue_residual(0);
    e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { hipEventDestroy(evs[0]); hipEventDestroy(evs[1]); return hip_check(c, e, "solve"); }
    if (c->S_host->done || iters >= max_iter || restarts >= 3) {
      finished = true;
    } else {
      ++restarts;  // recursive residual drifted: restart from the true residual
      it = iters;
    }
  }
  hipEventDestroy(evs[0]);
  hipEventDestroy(evs[1]);
  if (st) {
    st->iterations = c->S_host->iter;
    st->converged = c->S_host->done && !c->S_host->nonfinite;
    st->residual_div = c->S_host->rmax;
    st->levels = mg ? c->levels : 1;
  }
  if (c->S_host->nonfinite) return fail(c, SP_ERR_NONFINITE, "non-finite value in the solve");
  return SP_OK;
}

// Iterative refinement of a non-singular system's pressure, kept as
// x + xd: the residual of x + xd in difference form, then, when it is above
// tol, a solve of A e = residual from e = 0 and xd += e.  The corrected
// velocity takes the differences of x and of xd separately, so it is not
// limited by the rounding of a pressure level that a Dirichlet face fixes.
// *rmax: max |residual|/V before the solve; *iters: its CG iterations.
sp_status refine(sp_ctx* c, double tol, double* rmax, int* iters) {
  const Dims& D = c->dm;
  CgArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K;
  A.cx = c->cx; A.cy = c->cy; A.cz = c->cz; A.e = c->has_e ? c->ee : nullptr;
  A.active = c->active;
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp;
  A.b = c->b;
  A.x = c->x;
  A.S = c->S;
  const dim3 blk(64, 4, 1);
  const dim3 cgrd = col_grid(D.I, D.J, D.K);
  const uint32_t ncol = col_blocks(D.I, D.J, D.K);
  hipError_t e;
  if (!c->xd_active) {
    e = hipMemsetAsync(c->xd, 0, sizeof(double) * D.N, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemsetAsync(xd)");
    c->xd_active = true;
  }
  int slot;
  A.site = red_site(c, kSiteResidual);
  const double eb = c->has_e ? 8.0 : 0.0;
  timer_begin(c, "refine_residual", (double)D.N * (8 * 4 + 24 + eb + 1 + 8), &slot);
  hipLaunchKernelGGL(k_refine_residual, cgrd, blk, 0, c->stream, A, (const double*)c->xd, c->p1);
  hipLaunchKernelGGL(k_refine_residual_fin, dim3(1), dim3(kFinThreads), 0, c->stream, A, ncol);
  timer_end(c, slot);
  e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "refine");
  *rmax = c->S_host->npart[0];
  *iters = 0;
  if (!(*rmax > tol)) return SP_OK;
  const int32_t iter_before = c->S_host->iter;  // the device count runs on across solves
  // solve A e = rt into the spare iterate, then xd += e
  std::swap(c->x, c->xe);
  sp_stats st{};
  sp_status s = solve(c, c->p1, tol, &st, false, false);
  std::swap(c->x, c->xe);
  if (s != SP_OK) return s;
  *iters = st.iterations - iter_before;
  timer_begin(c, "refine_add", (double)D.N * 24, &slot);
  hipLaunchKernelGGL(k_refine_add, cgrd, blk, 0, c->stream, A, (const double*)c->xe, c->xd);
  timer_end(c, slot);
  return SP_OK;
}

// x += xd (the refined pressure, rounded, for the adjoint's checkpoint).
sp_status fold_refinement(sp_ctx* c) {
  const Dims& D = c->dm;
  CgArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K;
  hipLaunchKernelGGL(k_refine_add, col_grid(D.I, D.J, D.K), dim3(64, 4, 1), 0, c->stream, A,
                     (const double*)c->xd, c->x);
  c->xd_active = false;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "fold_refinement");
}

}  // namespace sp
