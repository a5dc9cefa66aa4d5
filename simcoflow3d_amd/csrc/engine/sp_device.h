// Device-side helpers shared by the engine kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sp_internal.h"

namespace sp {

// Workgroups are dispatched round-robin over the 8 XCDs of MI355X.  Remap
// the hardware workgroup id so that each XCD receives one contiguous chunk
// of the logical grid: neighbouring tiles (and their stencil halos) then
// share one L2.
__device__ inline uint32_t xcd_logical_block(uint32_t hw, uint32_t total) {
  if ((total & 7u) != 0u) return hw;
  return (hw & 7u) * (total >> 3) + (hw >> 3);
}

// Logical (bx, by, bz) of a 3-D launch after the XCD remap.
__device__ inline void xcd_block3(uint32_t& bx, uint32_t& by, uint32_t& bz) {
  const uint32_t gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const uint32_t hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const uint32_t lb = xcd_logical_block(hw, gx * gy * gz);
  bx = lb % gx;
  by = (lb / gx) % gy;
  bz = lb / (gx * gy);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Deterministic block reductions (result valid in thread 0).
template <int NT, typename T>
__device__ inline T block_sum(T v) {
  __shared__ T sm[NT / 64];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_sum(v);
  __syncthreads();
  if ((tid & 63) == 0) sm[tid >> 6] = v;
  __syncthreads();
  if (tid < 64) {
    v = tid < NT / 64 ? sm[tid] : T(0);
    v = wave_sum(v);
  }
  return v;
}

template <int NT, typename T>
__device__ inline T block_max(T v) {
  __shared__ T sm[NT / 64];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_max(v);
  __syncthreads();
  if ((tid & 63) == 0) sm[tid >> 6] = v;
  __syncthreads();
  if (tid < 64) {
    v = tid < NT / 64 ? sm[tid] : T(0);
    v = wave_max(v);
  }
  return v;
}

__device__ inline int flat_tid() {
  return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ inline uint32_t flat_block() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ inline uint32_t num_blocks() { return gridDim.x * gridDim.y * gridDim.z; }

__device__ inline void store_partial(double* slot, uint32_t b, double v) {
  __hip_atomic_store(slot + b, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_partial(const double* slot, uint32_t b) {
  return __hip_atomic_load(const_cast<double*>(slot) + b, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Returns true in every thread of the last block to finish.  Every block must
// call it exactly once after publishing its partials.
__device__ inline bool last_block(uint32_t* counter) {
  __shared__ int is_last;
  __syncthreads();
  if (flat_tid() == 0) {
    __threadfence();
    const uint32_t prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == num_blocks() - 1u);
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last != 0;
}

__device__ inline void reset_counter(uint32_t* counter) {
  if (flat_tid() == 0)
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum / max of `nb` partials in a fixed order by the calling block.
template <int NT>
__device__ inline double reduce_partials_sum(const double* slot, uint32_t nb) {
  double s = 0.0;
  for (uint32_t q = flat_tid(); q < nb; q += NT) s += load_partial(slot, q);
  return block_sum<NT>(s);
}
template <int NT>
__device__ inline double reduce_partials_max(const double* slot, uint32_t nb) {
  double s = 0.0;
  for (uint32_t q = flat_tid(); q < nb; q += NT) {
    double v = load_partial(slot, q);
    s = v > s ? v : s;
  }
  return block_max<NT>(s);
}

__device__ inline bool is_finite(double v) { return v - v == 0.0; }

// ---- deterministic grid reductions ----
//
// Every workgroup folds its values (fixed tree) and stores one partial per
// value with a plain store; a one-workgroup finalize kernel launched next on
// the same stream folds the partials in workgroup order.  The kernel boundary
// makes the partials visible, so no device-scope fences are needed, and the
// summation order depends only on the launch shape (bitwise reproducible).
// Layout of one reduction site: [kSiteValues][max_blocks] partials.
struct RedSite {
  double* part;
  uint32_t max_blocks;
};

// Reduce NV values over the workgroup with one barrier; results valid in
// thread 0 (fixed order: lanes, then waves).
template <int NT, int NV, unsigned MAXMASK>
__device__ inline void block_reduce_multi(double (&v)[NV]) {
  __shared__ double sm[NV][NT / 64];
  const int tid = flat_tid();
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = ((MAXMASK >> q) & 1u) ? wave_max(v[q]) : wave_sum(v[q]);
  __syncthreads();
  if ((tid & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q][tid >> 6] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const bool mx = (MAXMASK >> q) & 1u;
      double s = sm[q][0];
      for (int w = 1; w < NT / 64; ++w) s = mx ? (sm[q][w] > s ? sm[q][w] : s) : s + sm[q][w];
      v[q] = s;
    }
  }
}

template <int NT, int NV, unsigned MAXMASK>
__device__ inline void block_partials(const RedSite& R, double (&v)[NV]) {
  const uint32_t b = flat_block();
  block_reduce_multi<NT, NV, MAXMASK>(v);
  if (flat_tid() == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) R.part[(size_t)q * R.max_blocks + b] = v[q];
  }
}

// Executed by one workgroup of NT threads: folds nb partials per value.
// Each thread keeps 8 independent accumulators so loads stay in flight.
template <int NT, int NV, unsigned MAXMASK>
__device__ inline void finalize_partials(const RedSite& R, uint32_t nb, double (&out)[NV]) {
  __shared__ double tot[NV];
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const bool mx = (MAXMASK >> q) & 1u;
    const double* p = R.part + (size_t)q * R.max_blocks;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t k = flat_tid();
    for (; k + 7 * NT < nb; k += 8 * NT) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double x = p[k + u * NT];
        s[u] = mx ? (x > s[u] ? x : s[u]) : s[u] + x;
      }
    }
    for (int u = 0; k < nb; k += NT, ++u) {
      const double x = p[k];
      s[u] = mx ? (x > s[u] ? x : s[u]) : s[u] + x;
    }
    double t = s[0];
#pragma unroll
    for (int u = 1; u < 8; ++u) t = mx ? (s[u] > t ? s[u] : t) : t + s[u];
    acc[q] = t;
  }
  block_reduce_multi<NT, NV, MAXMASK>(acc);
  if (flat_tid() == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = acc[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = tot[q];
}

// First stage of a reduction over many partials: workgroup g folds the
// partials [g*chunk, min((g+1)*chunk, nb)) of every value into slot g of
// `out` (fixed order, so the result stays bitwise reproducible).  The
// one-workgroup finalize then reads gridDim.x partials instead of nb.
template <int NV, unsigned MAXMASK>
__global__ void __launch_bounds__(256) k_fold_partials(RedSite in, RedSite out, uint32_t nb,
                                                       uint32_t chunk) {
  const uint32_t g = blockIdx.x;
  const uint32_t lo = g * chunk, hi = lo + chunk < nb ? lo + chunk : nb;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (uint32_t k = lo + threadIdx.x; k < hi; k += 256) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double x = in.part[(size_t)q * in.max_blocks + k];
      acc[q] = ((MAXMASK >> q) & 1u) ? (x > acc[q] ? x : acc[q]) : acc[q] + x;
    }
  }
  block_reduce_multi<256, NV, MAXMASK>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) out.part[(size_t)q * out.max_blocks + g] = acc[q];
  }
}

inline RedSite red_site(const sp_ctx* c, int id) {
  RedSite r;
  r.part = c->red_part + (size_t)kSiteValues * c->red_max_blocks * id;
  r.max_blocks = c->red_max_blocks;
  return r;
}

constexpr int kFinThreads = 256;

// Flat launch: one cell per thread, 64x4 workgroups.
inline dim3 flat_grid(int nx, int ny, int nz) {
  return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz);
}
inline uint32_t flat_blocks(int nx, int ny, int nz) {
  return (uint32_t)((nx + 63) / 64) * (uint32_t)((ny + 3) / 4) * (uint32_t)nz;
}

// One cell per thread over a 3-D grid of 64x4 workgroups, workgroup order
// remapped so that every XCD owns a contiguous slab.
struct Cell3 {
  int a, b, c;
  bool in;
};
__device__ inline Cell3 cell_of(int nx, int ny, int nz) {
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  Cell3 r;
  r.a = (int)(bx * blockDim.x + threadIdx.x);
  r.b = (int)(by * blockDim.y + threadIdx.y);
  r.c = (int)bz;
  r.in = r.a < nx && r.b < ny && r.c < nz;
  return r;
}

// Column launch: 64x4 workgroups, every thread walks kColK consecutive cells
// up k at one (i, j).  Block reductions are paid once per kColK cells, and a
// stencil keeps its k-1 / k+1 values in registers.
constexpr int kColK = 16;
inline dim3 col_grid(int nx, int ny, int nz) {
  return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4),
              (unsigned)((nz + kColK - 1) / kColK));
}
inline uint32_t col_blocks(int nx, int ny, int nz) {
  return (uint32_t)((nx + 63) / 64) * (uint32_t)((ny + 3) / 4) *
         (uint32_t)((nz + kColK - 1) / kColK);
}
struct Col3 {
  int a, b, c0, c1;  // cells (a, b, c) for c0 <= c < c1
  bool in;
};
template <int KC = kColK>
__device__ inline Col3 col_of(int nx, int ny, int nz) {
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  Col3 r;
  r.a = (int)(bx * blockDim.x + threadIdx.x);
  r.b = (int)(by * blockDim.y + threadIdx.y);
  r.c0 = (int)bz * KC;
  r.c1 = r.c0 + KC < nz ? r.c0 + KC : nz;
  r.in = r.a < nx && r.b < ny;
  return r;
}
// Columns of KC cells (launch and partial count)
template <int KC>
inline dim3 col_grid_k(int nx, int ny, int nz) {
  return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)((nz + KC - 1) / KC));
}
template <int KC>
inline uint32_t col_blocks_k(int nx, int ny, int nz) {
  return (uint32_t)((nx + 63) / 64) * (uint32_t)((ny + 3) / 4) * (uint32_t)((nz + KC - 1) / KC);
}

}  // namespace sp
