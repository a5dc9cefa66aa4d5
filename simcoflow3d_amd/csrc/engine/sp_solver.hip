// Multigrid-preconditioned flexible CG on the GPU.
//
// Replaces HYPRE PCG + BoomerAMG (ProjectionP.f90:95-116, 267-299).  The
// outer CG, its dot products and the residual are FP64; the V-cycle runs in
// T = float (mixed) or double (FP64).  CG scalars stay in device memory: the
// host only polls the convergence flag, one iteration behind.
//
// Multigrid: cell-centred coarsening by 2, piecewise-constant transfers
// (restriction = sum over children, prolongation = injection), coarse face
// coefficients = scale * sum of the fine face coefficients they cover (the
// Galerkin product for these transfers, scaled), red-black Gauss-Seidel
// (pre: red,black; post: black,red -- a symmetric V-cycle).  Levels holding
// at most kTailCells cells run in one single-workgroup kernel.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "sp_device.h"
#include "sp_geom.h"
#include "sp_internal.h"

namespace sp {

namespace {

template <typename T>
struct LevelView {
  int nx, ny, nz;
  const T *cx, *cy, *cz, *e;
  T* x;
  T* b;
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

// One colour of Gauss-Seidel.  zero_init: x starts at 0 (the other colour
// is written as 0).
template <typename T>
__global__ void __launch_bounds__(256) k_gs(LevelView<T> L, int color, int zero_init,
                                            const Scalars* S) {
  if (S->done) return;
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (j >= L.ny) return;
  const int par = (j + k + color) & 1;
  const int i = 2 * ia + par;
  if (zero_init) {
    const int io = 2 * ia + (par ^ 1);
    if (io < L.nx) L.x[io + (int64_t)L.nx * (j + (int64_t)L.ny * k)] = T(0);
  }
  if (i >= L.nx) return;
  const int64_t n = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
  const T d = diag_of(L, i, j, k, n);
  if (!(d > T(0))) {
    if (zero_init) L.x[n] = T(0);
    return;
  }
  T s = L.b[n];
  if (!zero_init) s += offdiag_sum(L, i, j, k, n);
  L.x[n] = s / d;
}

// Coarse b = sum over children of the fine residual.
template <typename T>
__global__ void __launch_bounds__(256) k_resid_restrict(LevelView<T> F, LevelView<T> C,
                                                        const Scalars* S) {
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= C.nx || j >= C.ny) return;
  T s = T(0);
#pragma unroll
  for (int cc = 0; cc < 2; ++cc)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int aa = 0; aa < 2; ++aa) {
        const int fi = 2 * i + aa, fj = 2 * j + bb, fk = 2 * k + cc;
        if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
        const int64_t m = fi + (int64_t)F.nx * (fj + (int64_t)F.ny * fk);
        const T d = diag_of(F, fi, fj, fk, m);
        if (d > T(0)) s += F.b[m] - (d * F.x[m] - offdiag_sum(F, fi, fj, fk, m));
      }
  C.b[i + (int64_t)C.nx * (j + (int64_t)C.ny * k)] = s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_prolong(LevelView<T> F, LevelView<T> C, const Scalars* S) {
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= F.nx || j >= F.ny) return;
  const int64_t m = i + (int64_t)F.nx * (j + (int64_t)F.ny * k);
  if (diag_of(F, i, j, k, m) > T(0))
    F.x[m] += C.x[(i >> 1) + (int64_t)C.nx * ((j >> 1) + (int64_t)C.ny * (k >> 1))];
}

// ---- single-workgroup V-cycle over the small levels ----

template <typename T>
__device__ void tail_gs(const LevelView<T>& L, int color, int zero_init) {
  const int64_t half = ((int64_t)(L.nx + 1) / 2) * L.ny * L.nz;
  const int nxh = (L.nx + 1) / 2;
  for (int64_t t = threadIdx.x; t < half; t += blockDim.x) {
    const int ia = (int)(t % nxh);
    const int j = (int)((t / nxh) % L.ny);
    const int k = (int)(t / ((int64_t)nxh * L.ny));
    const int par = (j + k + color) & 1;
    const int i = 2 * ia + par;
    if (zero_init) {
      const int io = 2 * ia + (par ^ 1);
      if (io < L.nx) L.x[io + (int64_t)L.nx * (j + (int64_t)L.ny * k)] = T(0);
    }
    if (i >= L.nx) continue;
    const int64_t n = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
    const T d = diag_of(L, i, j, k, n);
    if (!(d > T(0))) {
      if (zero_init) L.x[n] = T(0);
      continue;
    }
    T s = L.b[n];
    if (!zero_init) s += offdiag_sum(L, i, j, k, n);
    L.x[n] = s / d;
  }
  __syncthreads();
}

template <typename T>
__device__ void tail_project_mean(const LevelView<T>& L, T* v) {
  __shared__ double red[2][16];
  double s = 0.0, c = 0.0;
  for (int64_t n = threadIdx.x; n < (int64_t)L.nx * L.ny * L.nz; n += blockDim.x) {
    const int i = (int)(n % L.nx), j = (int)((n / L.nx) % L.ny), k = (int)(n / ((int64_t)L.nx * L.ny));
    if (diag_of(L, i, j, k, n) > T(0)) { s += (double)v[n]; c += 1.0; }
  }
  s = wave_sum(s);
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = s; red[1][threadIdx.x >> 6] = c; }
  __syncthreads();
  double ts = 0.0, tc = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { ts += red[0][w]; tc += red[1][w]; }
  const T mean = tc > 0 ? (T)(ts / tc) : T(0);
  __syncthreads();
  for (int64_t n = threadIdx.x; n < (int64_t)L.nx * L.ny * L.nz; n += blockDim.x) {
    const int i = (int)(n % L.nx), j = (int)((n / L.nx) % L.ny), k = (int)(n / ((int64_t)L.nx * L.ny));
    if (diag_of(L, i, j, k, n) > T(0)) v[n] -= mean;
  }
  __syncthreads();
}

template <typename T>
__device__ void tail_restrict(const LevelView<T>& F, const LevelView<T>& C) {
  const int64_t nc = (int64_t)C.nx * C.ny * C.nz;
  for (int64_t t = threadIdx.x; t < nc; t += blockDim.x) {
    const int i = (int)(t % C.nx), j = (int)((t / C.nx) % C.ny), k = (int)(t / ((int64_t)C.nx * C.ny));
    T s = T(0);
    for (int cc = 0; cc < 2; ++cc)
      for (int bb = 0; bb < 2; ++bb)
        for (int aa = 0; aa < 2; ++aa) {
          const int fi = 2 * i + aa, fj = 2 * j + bb, fk = 2 * k + cc;
          if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
          const int64_t m = fi + (int64_t)F.nx * (fj + (int64_t)F.ny * fk);
          const T d = diag_of(F, fi, fj, fk, m);
          if (d > T(0)) s += F.b[m] - (d * F.x[m] - offdiag_sum(F, fi, fj, fk, m));
        }
    C.b[t] = s;
  }
  __syncthreads();
}

template <typename T>
__device__ void tail_prolong(const LevelView<T>& F, const LevelView<T>& C) {
  const int64_t nf = (int64_t)F.nx * F.ny * F.nz;
  for (int64_t m = threadIdx.x; m < nf; m += blockDim.x) {
    const int i = (int)(m % F.nx), j = (int)((m / F.nx) % F.ny), k = (int)(m / ((int64_t)F.nx * F.ny));
    if (diag_of(F, i, j, k, m) > T(0))
      F.x[m] += C.x[(i >> 1) + (int64_t)C.nx * ((j >> 1) + (int64_t)C.ny * (k >> 1))];
  }
  __syncthreads();
}

template <typename T>
struct TailArgs {
  int first, last;  // levels [first, last)
  int pre, post, coarse_sweeps, singular;
  const LevelView<T>* lv;
};

template <typename T>
__global__ void __launch_bounds__(1024) k_tail(TailArgs<T> A, const Scalars* S) {
  if (S->done) return;
  const int L = A.last;
  for (int l = A.first; l < L - 1; ++l) {
    const LevelView<T> F = A.lv[l];
    tail_gs(F, 0, 1);
    tail_gs(F, 1, 0);
    for (int s = 1; s < A.pre; ++s) { tail_gs(F, 0, 0); tail_gs(F, 1, 0); }
    tail_restrict(F, A.lv[l + 1]);
  }
  {
    const LevelView<T> C = A.lv[L - 1];
    if (A.singular) tail_project_mean(C, C.b);
    tail_gs(C, 0, 1);
    tail_gs(C, 1, 0);
    for (int s = 1; s < A.coarse_sweeps; ++s) { tail_gs(C, 0, 0); tail_gs(C, 1, 0); }
    for (int s = 0; s < A.coarse_sweeps; ++s) { tail_gs(C, 1, 0); tail_gs(C, 0, 0); }
    if (A.singular) tail_project_mean(C, C.x);
  }
  for (int l = L - 2; l >= A.first; --l) {
    const LevelView<T> F = A.lv[l];
    tail_prolong(F, A.lv[l + 1]);
    for (int s = 0; s < A.post; ++s) { tail_gs(F, 1, 0); tail_gs(F, 0, 0); }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_coarsen(LevelView<T> F, LevelView<T> C, T* ccx, T* ccy,
                                                 T* ccz, T* ce, T scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= C.nx || j >= C.ny) return;
  T sx = 0, sy = 0, sz = 0, se = 0;
  for (int cc = 0; cc < 2; ++cc)
    for (int bb = 0; bb < 2; ++bb)
      for (int aa = 0; aa < 2; ++aa) {
        const int fi = 2 * i + aa, fj = 2 * j + bb, fk = 2 * k + cc;
        if (fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
        const int64_t m = fi + (int64_t)F.nx * (fj + (int64_t)F.ny * fk);
        if (F.e) se += F.e[m];
        if (aa == 1) sx += F.cx[m];
        if (bb == 1) sy += F.cy[m];
        if (cc == 1) sz += F.cz[m];
      }
  const int64_t n = i + (int64_t)C.nx * (j + (int64_t)C.ny * k);
  ccx[n] = scale * sx;
  ccy[n] = scale * sy;
  ccz[n] = scale * sz;
  if (ce) ce[n] = scale * se;
}

// ---- outer CG (FP64) ----

struct CgArgs {
  int I, J, K, kper;
  const double *cx, *cy, *cz, *e;
  const uint8_t* active;
  const float* dw;
  const double *dxp, *dyp, *dzp;
  const double* b;
  double *x, *r, *q, *pold, *pnew;
  const double* z64;
  const float* z32;
  float* r32;
  double* partials;
  Scalars* S;
  double tol;
};

__device__ inline double zval(const CgArgs& A, int64_t n) {
  return A.z32 ? (double)A.z32[n] : A.z64[n];
}

// r = b - mean(b) - A x  (x = 0 on the first call)
__global__ void __launch_bounds__(256) k_cg_residual(CgArgs A, int from_zero) {
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  const int a = bx * kTX + threadIdx.x, b = by * kTY + threadIdx.y;
  const int c0 = bz * A.kper, c1 = min(c0 + A.kper, A.K);
  const int64_t IJ = (int64_t)A.I * A.J;
  Scalars* S = A.S;
  const double bmean = S->bmean;
  double rmax = 0.0, bad = 0.0;
  if (a < A.I && b < A.J) {
    for (int c = c0; c < c1; ++c) {
      const int64_t n = a + (int64_t)A.I * (b + (int64_t)A.J * c);
      double rv = 0.0;
      if (A.active[n]) {
        rv = A.b[n] - bmean;
        if (!from_zero) {
          double d = a > 0 ? A.cx[n - 1] : 0.0;
          d = d + (b > 0 ? A.cy[n - A.I] : 0.0);
          d = d + (c > 0 ? A.cz[n - IJ] : 0.0);
          d = d + A.cx[n];
          d = d + A.cy[n];
          d = d + A.cz[n];
          if (A.e) d = d + A.e[n];
          double ax = d * A.x[n];
          if (a > 0) ax -= A.cx[n - 1] * A.x[n - 1];
          if (b > 0) ax -= A.cy[n - A.I] * A.x[n - A.I];
          if (c > 0) ax -= A.cz[n - IJ] * A.x[n - IJ];
          if (a < A.I - 1) ax -= A.cx[n] * A.x[n + 1];
          if (b < A.J - 1) ax -= A.cy[n] * A.x[n + A.I];
          if (c < A.K - 1) ax -= A.cz[n] * A.x[n + IJ];
          rv -= ax;
        }
        const double v = fabs(rv) / (A.dxp[a] * A.dyp[b] * A.dzp[c]);
        rmax = v > rmax ? v : rmax;
        if (!is_finite(rv)) bad = 1.0;
      }
      A.r[n] = rv;
      if (A.r32) A.r32[n] = (float)rv;
      if (from_zero) A.x[n] = 0.0;
      A.q[n] = 0.0;
      A.pold[n] = 0.0;
    }
  }
  const double m = block_max<kTileThreads>(rmax);
  const double f = block_max<kTileThreads>(bad);
  if (flat_tid() == 0) {
    store_partial(A.partials, flat_block(), m);
    store_partial(A.partials + kMaxPartials, flat_block(), f);
  }
  if (last_block(&S->counter[kCtrInit])) {
    const double mm = reduce_partials_max<kTileThreads>(A.partials, num_blocks());
    const double ff = reduce_partials_max<kTileThreads>(A.partials + kMaxPartials, num_blocks());
    if (flat_tid() == 0) {
      S->rmax = mm;
      S->nonfinite = S->nonfinite || ff > 0.0 || !is_finite(mm);
      S->done = (mm <= A.tol || S->nonfinite) ? 1 : 0;
      S->first = 1;
      S->alpha = 0.0;
      S->beta = 0.0;
      S->rz = 0.0;
    }
    reset_counter(&S->counter[kCtrInit]);
  }
}

// rz = (r, z), zq = (z, q);  beta = -alpha zq / rz_old (flexible CG).
// Singular systems: z is shifted by its diagonal-weighted mean m so the
// iterate keeps the pressure of the stiff (high-coefficient) phase near zero;
// (r, z-m) and (z-m, q) follow from the raw sums.
__global__ void __launch_bounds__(256) k_cg_dots(CgArgs A) {
  Scalars* S = A.S;
  if (S->done) return;
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  const int a = bx * kTX + threadIdx.x, b = by * kTY + threadIdx.y;
  const int c0 = bz * A.kper, c1 = min(c0 + A.kper, A.K);
  double rz = 0.0, zq = 0.0, dz = 0.0, rs = 0.0, qs = 0.0;
  if (a < A.I && b < A.J) {
    for (int c = c0; c < c1; ++c) {
      const int64_t n = a + (int64_t)A.I * (b + (int64_t)A.J * c);
      if (A.active[n]) {
        const double z = zval(A, n), r = A.r[n], q = A.q[n];
        rz += r * z;
        zq += z * q;
        dz += (double)A.dw[n] * z;
        rs += r;
        qs += q;
      }
    }
  }
  const double s0 = block_sum<kTileThreads>(rz);
  const double s1 = block_sum<kTileThreads>(zq);
  const double s2 = block_sum<kTileThreads>(dz);
  const double s3 = block_sum<kTileThreads>(rs);
  const double s4 = block_sum<kTileThreads>(qs);
  if (flat_tid() == 0) {
    const uint32_t fb = flat_block();
    store_partial(A.partials, fb, s0);
    store_partial(A.partials + kMaxPartials, fb, s1);
    store_partial(A.partials + 2 * (int64_t)kMaxPartials, fb, s2);
    store_partial(A.partials + 3 * (int64_t)kMaxPartials, fb, s3);
    store_partial(A.partials + 4 * (int64_t)kMaxPartials, fb, s4);
  }
  if (last_block(&S->counter[kCtrDots])) {
    const uint32_t nb = num_blocks();
    const double t0 = reduce_partials_sum<kTileThreads>(A.partials, nb);
    const double t1 = reduce_partials_sum<kTileThreads>(A.partials + kMaxPartials, nb);
    const double t2 = reduce_partials_sum<kTileThreads>(A.partials + 2 * (int64_t)kMaxPartials, nb);
    const double t3 = reduce_partials_sum<kTileThreads>(A.partials + 3 * (int64_t)kMaxPartials, nb);
    const double t4 = reduce_partials_sum<kTileThreads>(A.partials + 4 * (int64_t)kMaxPartials, nb);
    if (flat_tid() == 0) {
      const double m = (S->singular && S->dsum > 0.0) ? t2 / S->dsum : 0.0;
      const double rzm = t0 - m * t3, zqm = t1 - m * t4;
      S->zshift = m;
      S->beta = S->first ? 0.0 : -S->alpha * zqm / S->rz;
      S->rz = rzm;
      S->zq = zqm;
      S->first = 0;
      if (!is_finite(rzm) || !is_finite(S->beta)) { S->nonfinite = 1; S->done = 1; }
    }
    reset_counter(&S->counter[kCtrDots]);
  }
}

// p_new = z + beta p_old ; q = A p_new ; pq = (p_new, q) ; alpha = rz / pq
__global__ void __launch_bounds__(256) k_cg_spmv(CgArgs A) {
  Scalars* S = A.S;
  if (S->done) return;
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  const int a = bx * kTX + threadIdx.x, b = by * kTY + threadIdx.y;
  const int c0 = bz * A.kper, c1 = min(c0 + A.kper, A.K);
  const int64_t IJ = (int64_t)A.I * A.J;
  const double beta = S->beta;
  const double m = S->zshift;
  double pq = 0.0;
  if (a < A.I && b < A.J) {
    for (int c = c0; c < c1; ++c) {
      const int64_t n = a + (int64_t)A.I * (b + (int64_t)A.J * c);
      const double pc = (zval(A, n) - m) + beta * A.pold[n];
      double qv = 0.0;
      if (A.active[n]) {
        double d = a > 0 ? A.cx[n - 1] : 0.0;
        d = d + (b > 0 ? A.cy[n - A.I] : 0.0);
        d = d + (c > 0 ? A.cz[n - IJ] : 0.0);
        d = d + A.cx[n];
        d = d + A.cy[n];
        d = d + A.cz[n];
        if (A.e) d = d + A.e[n];
        qv = d * pc;
        if (a > 0) qv -= A.cx[n - 1] * ((zval(A, n - 1) - m) + beta * A.pold[n - 1]);
        if (b > 0) qv -= A.cy[n - A.I] * ((zval(A, n - A.I) - m) + beta * A.pold[n - A.I]);
        if (c > 0) qv -= A.cz[n - IJ] * ((zval(A, n - IJ) - m) + beta * A.pold[n - IJ]);
        if (a < A.I - 1) qv -= A.cx[n] * ((zval(A, n + 1) - m) + beta * A.pold[n + 1]);
        if (b < A.J - 1) qv -= A.cy[n] * ((zval(A, n + A.I) - m) + beta * A.pold[n + A.I]);
        if (c < A.K - 1) qv -= A.cz[n] * ((zval(A, n + IJ) - m) + beta * A.pold[n + IJ]);
        pq += pc * qv;
      }
      A.pnew[n] = A.active[n] ? pc : 0.0;
      A.q[n] = qv;
    }
  }
  const double s0 = block_sum<kTileThreads>(pq);
  if (flat_tid() == 0) store_partial(A.partials, flat_block(), s0);
  if (last_block(&S->counter[kCtrSpmv])) {
    const double t0 = reduce_partials_sum<kTileThreads>(A.partials, num_blocks());
    if (flat_tid() == 0) {
      S->pq = t0;
      S->alpha = S->rz / t0;
      if (!is_finite(S->alpha)) { S->nonfinite = 1; S->done = 1; }
    }
    reset_counter(&S->counter[kCtrSpmv]);
  }
}

// x += alpha p ; r -= alpha q ; rmax = max |r|/V
__global__ void __launch_bounds__(256) k_cg_update(CgArgs A) {
  Scalars* S = A.S;
  if (S->done) return;
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  const int a = bx * kTX + threadIdx.x, b = by * kTY + threadIdx.y;
  const int c0 = bz * A.kper, c1 = min(c0 + A.kper, A.K);
  const double alpha = S->alpha;
  double rmax = 0.0;
  if (a < A.I && b < A.J) {
    const double dxy = A.dxp[a] * A.dyp[b];
    for (int c = c0; c < c1; ++c) {
      const int64_t n = a + (int64_t)A.I * (b + (int64_t)A.J * c);
      A.x[n] += alpha * A.pnew[n];
      const double rv = A.r[n] - alpha * A.q[n];
      A.r[n] = rv;
      if (A.r32) A.r32[n] = (float)rv;
      const double v = fabs(rv) / (dxy * A.dzp[c]);
      rmax = v > rmax ? v : rmax;
    }
  }
  const double m = block_max<kTileThreads>(rmax);
  if (flat_tid() == 0) store_partial(A.partials, flat_block(), m);
  if (last_block(&S->counter[kCtrUpdate])) {
    const double mm = reduce_partials_max<kTileThreads>(A.partials, num_blocks());
    if (flat_tid() == 0) {
      S->rmax = mm;
      S->iter += 1;
      if (!is_finite(mm)) S->nonfinite = 1;
      S->done = (mm <= A.tol || S->nonfinite) ? 1 : 0;
    }
    reset_counter(&S->counter[kCtrUpdate]);
  }
}

template <typename T>
LevelView<T> view_of(const Level<T>& L) {
  return LevelView<T>{L.nx, L.ny, L.nz, L.cx, L.cy, L.cz, L.e, L.x, L.b};
}

template <typename T>
double level_bytes(const Level<T>& L, double per_cell_words) {
  return (double)L.n * sizeof(T) * per_cell_words;
}

template <typename T>
void launch_gs(sp_ctx* c, const Level<T>& L, int color, int zero_init) {
  const int nxh = (L.nx + 1) / 2;
  dim3 blk(64, 4, 1), grd((nxh + 63) / 64, (L.ny + 3) / 4, L.nz);
  int slot;
  // half the cells updated; every array streamed once
  const double words = (zero_init ? 1.0 : 2.0) + 3.0 + (L.e ? 1.0 : 0.0) + 0.5;
  timer_begin(c, L.n == c->dm.N ? (zero_init ? "mg_gs0_fine" : "mg_gs_fine") : "mg_gs_coarse",
              level_bytes(L, words), &slot);
  hipLaunchKernelGGL(k_gs<T>, grd, blk, 0, c->stream, view_of(L), color, zero_init,
                     (const Scalars*)c->S);
  timer_end(c, slot);
}

template <typename T>
void vcycle(sp_ctx* c, Level<T>* lv) {
  const int L = c->levels, Lt = c->tail_level;
  const int pre = std::max(1, c->prm.mg_pre), post = c->prm.mg_post;
  dim3 blk(64, 4, 1);
  int slot;
  for (int l = 0; l < Lt; ++l) {
    launch_gs(c, lv[l], 0, 1);
    launch_gs(c, lv[l], 1, 0);
    for (int s = 1; s < pre; ++s) { launch_gs(c, lv[l], 0, 0); launch_gs(c, lv[l], 1, 0); }
    dim3 grd((lv[l + 1].nx + 63) / 64, (lv[l + 1].ny + 3) / 4, lv[l + 1].nz);
    timer_begin(c, l == 0 ? "mg_restrict_fine" : "mg_restrict_coarse",
                level_bytes(lv[l], 5.0 + (lv[l].e ? 1.0 : 0.0)) + level_bytes(lv[l + 1], 1.0), &slot);
    hipLaunchKernelGGL(k_resid_restrict<T>, grd, blk, 0, c->stream, view_of(lv[l]),
                       view_of(lv[l + 1]), (const Scalars*)c->S);
    timer_end(c, slot);
  }
  {
    TailArgs<T> A;
    A.first = Lt;
    A.last = L;
    A.pre = pre;
    A.post = post;
    A.coarse_sweeps = std::max(1, c->prm.mg_coarse_sweeps);
    A.singular = c->S_host->singular;
    A.lv = reinterpret_cast<const LevelView<T>*>(c->mg_desc);
    timer_begin(c, "mg_tail", 0, &slot);
    hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(1024), 0, c->stream, A, (const Scalars*)c->S);
    timer_end(c, slot);
  }
  for (int l = Lt - 1; l >= 0; --l) {
    dim3 grd((lv[l].nx + 63) / 64, (lv[l].ny + 3) / 4, lv[l].nz);
    timer_begin(c, l == 0 ? "mg_prolong_fine" : "mg_prolong_coarse",
                level_bytes(lv[l], 5.0 + (lv[l].e ? 1.0 : 0.0)) + level_bytes(lv[l + 1], 1.0), &slot);
    hipLaunchKernelGGL(k_prolong<T>, grd, blk, 0, c->stream, view_of(lv[l]), view_of(lv[l + 1]),
                       (const Scalars*)c->S);
    timer_end(c, slot);
    for (int s = 0; s < post; ++s) { launch_gs(c, lv[l], 1, 0); launch_gs(c, lv[l], 0, 0); }
  }
}

template <typename T>
sp_status setup_levels(sp_ctx* c, Level<T>* lv) {
  const int L = c->levels;
  dim3 blk(64, 4, 1);
  for (int l = 0; l < L; ++l) lv[l].e = c->has_e ? lv[l].e_store : nullptr;
  for (int l = 1; l < L; ++l) {
    Level<T>& C = lv[l];
    dim3 grd((C.nx + 63) / 64, (C.ny + 3) / 4, C.nz);
    hipLaunchKernelGGL(k_coarsen<T>, grd, blk, 0, c->stream, view_of(lv[l - 1]), view_of(C), C.cx,
                       C.cy, C.cz, C.e, (T)c->prm.mg_coarse_scale);
  }
  // level table for the tail kernel
  std::vector<LevelView<T>> desc(kMaxLevels);
  for (int l = 0; l < L; ++l) desc[l] = view_of(lv[l]);
  hipError_t e = hipMemcpyAsync(c->mg_desc, desc.data(), sizeof(LevelView<T>) * kMaxLevels,
                                hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(mg_desc)");
  return SP_OK;
}

}  // namespace

sp_status mg_setup(sp_ctx* c) {
  if (c->prm.precond != SP_PRECOND_MG) return SP_OK;
  if (c->prm.precision == SP_PREC_MIXED) return setup_levels<float>(c, c->lf);
  return setup_levels<double>(c, c->ld);
}

// Solve A x = rhs - mean (singular) with the multigrid-preconditioned
// flexible CG.  The operator arrays and c->S (bmean, singular) come from the
// assembly.  Writes x into c->x.
sp_status solve(sp_ctx* c, const double* rhs, double* /*xout*/, bool /*forward*/, sp_stats* st) {
  const Dims& D = c->dm;
  const bool mixed = c->prm.precision == SP_PREC_MIXED;
  const bool mg = c->prm.precond == SP_PRECOND_MG;
  Geom g = make_geom(D.I, D.J, D.K, 2048);
  if (g.blocks() > (uint32_t)kMaxPartials) return fail(c, SP_ERR_INVALID_ARG, "grid too large");
  CgArgs A{};
  A.I = D.I; A.J = D.J; A.K = D.K; A.kper = g.kper;
  A.cx = c->cx; A.cy = c->cy; A.cz = c->cz; A.e = c->has_e ? c->ee : nullptr;
  A.active = c->active;
  A.dw = c->dw;
  A.dxp = c->dxp; A.dyp = c->dyp; A.dzp = c->dzp;
  A.b = rhs;
  A.x = c->x; A.r = c->r; A.q = c->q;
  A.partials = c->partials;
  A.S = c->S;
  A.tol = c->prm.div_tol;
  if (mixed && mg) {
    A.z32 = c->z32;
    A.r32 = c->r32;
  } else {
    A.z64 = c->z64;
  }
  double* pbuf[2] = {c->p0, c->p1};
  int slot;
  const double NB = (double)D.N;
  const double eb = c->has_e ? 8.0 : 0.0;
  const double zb = (mixed && mg) ? 4.0 : 8.0;

  auto enqueue_residual = [&](int from_zero) {
    A.pold = pbuf[0];
    A.pnew = pbuf[1];
    timer_begin(c, "cg_residual", NB * (from_zero ? 8 + 1 + 8 * 4 : 8 * 5 + eb + 1 + 8 * 4), &slot);
    hipLaunchKernelGGL(k_cg_residual, g.grid, g.block, 0, c->stream, A, from_zero);
    timer_end(c, slot);
  };
  auto precondition = [&]() {
    if (mg) {
      if (mixed) vcycle<float>(c, c->lf);
      else vcycle<double>(c, c->ld);
    } else {
      // Jacobi: z = r / d through a zero-init smoothing step of level 0
      Level<double> L0{};
      L0.nx = D.I; L0.ny = D.J; L0.nz = D.K; L0.n = D.N;
      L0.cx = c->cx; L0.cy = c->cy; L0.cz = c->cz; L0.e = c->has_e ? c->ee : nullptr;
      L0.x = c->z64; L0.b = c->r;
      LevelView<double> V = view_of(L0);
      const int nxh = (D.I + 1) / 2;
      dim3 blk(64, 4, 1), grd((nxh + 63) / 64, (D.J + 3) / 4, D.K);
      timer_begin(c, "jacobi", NB * 8 * 6, &slot);
      hipLaunchKernelGGL(k_gs<double>, grd, blk, 0, c->stream, V, 0, 1, (const Scalars*)c->S);
      hipLaunchKernelGGL(k_gs<double>, grd, blk, 0, c->stream, V, 1, 1, (const Scalars*)c->S);
      timer_end(c, slot);
    }
  };

  hipError_t e;
  int32_t* hflag = c->hflags;
  int pp = 0;  // parity of the p double buffer since the last (re)start
  enqueue_residual(1);
  int it = 0, cur = 0;
  e = hipMemcpyAsync(&hflag[cur], &c->S->done, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync");
  hipEvent_t evs[2];
  hipEventCreateWithFlags(&evs[0], hipEventDisableTiming);
  hipEventCreateWithFlags(&evs[1], hipEventDisableTiming);
  hipEventRecord(evs[cur], c->stream);
  int restarts = 0;
  const int max_iter = std::max(1, c->prm.max_iter);
  bool finished = false;
  while (!finished) {
    while (it < max_iter) {
      precondition();
      timer_begin(c, "cg_dots", NB * (8 + zb + 8 + 4 + 1), &slot);
      hipLaunchKernelGGL(k_cg_dots, g.grid, g.block, 0, c->stream, A);
      timer_end(c, slot);
      A.pold = pbuf[pp & 1];
      A.pnew = pbuf[(pp + 1) & 1];
      ++pp;
      timer_begin(c, "cg_spmv", NB * (zb + 8 + 24 + eb + 1 + 8 + 8), &slot);
      hipLaunchKernelGGL(k_cg_spmv, g.grid, g.block, 0, c->stream, A);
      timer_end(c, slot);
      timer_begin(c, "cg_update", NB * (8 * 4 + 8 * 2 + (A.r32 ? 4 : 0)), &slot);
      hipLaunchKernelGGL(k_cg_update, g.grid, g.block, 0, c->stream, A);
      timer_end(c, slot);
      const int nxt = cur ^ 1;
      hipMemcpyAsync(&hflag[nxt], &c->S->done, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
      hipEventRecord(evs[nxt], c->stream);
      // wait for the flag of the previous iteration (one iteration in flight)
      e = hipEventSynchronize(evs[cur]);
      if (e != hipSuccess) { hipEventDestroy(evs[0]); hipEventDestroy(evs[1]); return hip_check(c, e, "solve"); }
      cur = nxt;
      ++it;
      if (hflag[cur ^ 1]) break;
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { hipEventDestroy(evs[0]); hipEventDestroy(evs[1]); return hip_check(c, e, "solve"); }
    // true residual of the final iterate
    e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    const int iters = c->S_host->iter;
    if (c->S_host->nonfinite) break;
    hipMemsetAsync(&c->S->done, 0, sizeof(int32_t), c->stream);
    pp = 0;
    enqueue_residual(0);
    e = hipMemcpyAsync(c->S_host, c->S, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    c->S_host->iter = iters;
    if (c->S_host->done || iters >= max_iter || restarts >= 3) {
      finished = true;
    } else {
      ++restarts;  // recursive residual drifted: restart from the true residual
      it = iters;
      hflag[2] = iters;
      hipMemcpyAsync(&c->S->iter, &hflag[2], sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
      hipMemcpyAsync(&hflag[cur], &c->S->done, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
      hipEventRecord(evs[cur], c->stream);
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

}  // namespace sp
