// Launch geometry of the grid kernels.
//
// Grid kernels use 64x4 workgroups (one wavefront per i-row of 64 cells, four
// j-rows per workgroup) and march over a chunk of k planes, so the k-1/k/k+1
// planes of a column stay in registers and the workgroup count stays near a
// few per CU (few reduction partials, full occupancy of 256 CUs).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace sp {

constexpr int kTX = 64;
constexpr int kTY = 4;
constexpr int kTileThreads = kTX * kTY;

struct Geom {
  dim3 grid, block;
  int kper;  // k planes per workgroup
  uint32_t blocks() const { return grid.x * grid.y * grid.z; }
};

// `target` = desired number of workgroups.
inline Geom make_geom(int nx, int ny, int nz, int target = 2048) {
  Geom g;
  g.block = dim3(kTX, kTY, 1);
  const int gx = (nx + kTX - 1) / kTX, gy = (ny + kTY - 1) / kTY;
  const int64_t cols = (int64_t)gx * gy;
  int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(nz, (target + cols - 1) / cols));
  int kper = (nz + chunks - 1) / chunks;
  chunks = (nz + kper - 1) / kper;
  g.grid = dim3(gx, gy, chunks);
  g.kper = kper;
  return g;
}

}  // namespace sp
