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

}  // namespace sp
