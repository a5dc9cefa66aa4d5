// Cell bookkeeping that follows GridPreprocess when a solid moves
// (SolidBody.f90:817-822, Main.f90:207-214):
//
//   DefineMomentumExchangeCell (CutCell.f90:673-751): cut U/V/W cells next
//     to a solid P cell are flagged in MoExCell;
//   NumberExternalCell (CutCell.f90:753-786): the non-solid cells numbered
//     from 0 in the reference's loop order (i outermost, k innermost).
//
// The numbering is a prefix count in an order where each (i, j) column of k
// is contiguous: one pass counts pieces of 64 planes of every column, a
// two-level scan (workgroups, then their totals) numbers the pieces, and a
// last pass writes each piece's numbers.

#include <hip/hip_runtime.h>

#include "sp_device.h"
#include "sp_internal.h"

namespace sp {
namespace {

constexpr int kScanThreads = 1024;
constexpr int kNumChunkPlanes = 64;

int number_chunks(int K2) { return (K2 + kNumChunkPlanes - 1) / kNumChunkPlanes; }

__global__ void __launch_bounds__(256) k_moex(int I, int J, int K, const int32_t* ctp,
                                              const int32_t* ctu, const int32_t* ctv,
                                              const int32_t* ctw, int32_t* mu, int32_t* mv,
                                              int32_t* mw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > I || j > J) return;
  const int64_t q = (int64_t)(i - 1) + (int64_t)I * ((j - 1) + (int64_t)J * (k - 1));
  const bool sp = ctp[q] == 2;
  if (ctu[q] == 1 && (sp || (i < I && ctp[q + 1] == 2))) mu[q] = 1;
  if (ctv[q] == 1 && (sp || (j < J && ctp[q + I] == 2))) mv[q] = 1;
  if (ctw[q] == 1 && (sp || (k < K && ctp[q + (int64_t)I * J] == 2))) mw[q] = 1;
}

// Numbering: the scan order is (i, j, k) with k innermost, so column (i, j)
// is contiguous in it and a chunk of kNumChunk planes of a column is a
// contiguous piece.  Piece p = ((i-1)*J2 + (j-1))*nc + chunk.

constexpr int kNumChunk = kNumChunkPlanes;
constexpr int kScanPer = 8;                          // elements per scan thread
constexpr int kScanBlock = kScanThreads * kScanPer;  // elements per scan workgroup

// Non-solid cells of each piece.
__global__ void __launch_bounds__(256) k_number_count(int I, int J, int I2, int J2, int K2, int nc,
                                                      const int32_t* ct, int32_t* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int c = blockIdx.z;
  if (i > I2 || j > J2) return;
  const int k0 = c * kNumChunk + 1, k1 = min(K2, k0 + kNumChunk - 1);
  const int64_t IJ = (int64_t)I * J;
  int64_t q = (int64_t)(i - 1) + (int64_t)I * (j - 1) + IJ * (k0 - 1);
  int32_t n = 0;
  for (int k = k0; k <= k1; ++k, q += IJ) n += ct[q] != 2;
  cnt[((int64_t)(i - 1) * J2 + (j - 1)) * nc + c] = n;
}

// Workgroup-level exclusive scan of kScanBlock piece counts in place; the
// workgroup's total goes to bsum[blockIdx.x].
__global__ void __launch_bounds__(kScanThreads) k_number_scan(int64_t n, int32_t* cnt,
                                                             int32_t* bsum) {
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  const int64_t lo = (int64_t)blockIdx.x * kScanBlock + (int64_t)t * kScanPer;
  int32_t v[kScanPer];
  int32_t s = 0;
#pragma unroll
  for (int e = 0; e < kScanPer; ++e) {
    v[e] = lo + e < n ? cnt[lo + e] : 0;
    s += v[e];
  }
  part[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the thread sums
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int32_t w = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += w;
    __syncthreads();
  }
  int32_t run = part[t] - s;
#pragma unroll
  for (int e = 0; e < kScanPer; ++e) {
    if (lo + e < n) cnt[lo + e] = run;
    run += v[e];
  }
  if (t == kScanThreads - 1) bsum[blockIdx.x] = part[t];
}

// Exclusive scan of the nb workgroup totals (one workgroup, a thread per
// contiguous segment); the grand total - 1 (ExtCell) into *ext.
__global__ void __launch_bounds__(kScanThreads) k_number_scan_top(int64_t nb, int32_t* bsum,
                                                                 int32_t* ext) {
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  const int64_t per = (nb + kScanThreads - 1) / kScanThreads;
  const int64_t lo = per * t < nb ? per * t : nb;
  const int64_t hi = lo + per < nb ? lo + per : nb;
  int32_t s = 0;
  for (int64_t x = lo; x < hi; ++x) s += bsum[x];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int32_t w = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += w;
    __syncthreads();
  }
  int32_t run = part[t] - s;
  for (int64_t x = lo; x < hi; ++x) {
    const int32_t c = bsum[x];
    bsum[x] = run;
    run += c;
  }
  if (t == kScanThreads - 1) ext[0] = part[t] - 1;
}

__global__ void __launch_bounds__(256) k_number_write(int I, int J, int I2, int J2, int K2, int nc,
                                                      const int32_t* ct, const int32_t* off,
                                                      const int32_t* bsum, int32_t* pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int c = blockIdx.z;
  if (i > I2 || j > J2) return;
  const int k0 = c * kNumChunk + 1, k1 = min(K2, k0 + kNumChunk - 1);
  const int64_t IJ = (int64_t)I * J, IJ2 = (int64_t)I2 * J2;
  int64_t q = (int64_t)(i - 1) + (int64_t)I * (j - 1) + IJ * (k0 - 1);
  int64_t p = (int64_t)(i - 1) + (int64_t)I2 * (j - 1) + IJ2 * (k0 - 1);
  const int64_t x = ((int64_t)(i - 1) * J2 + (j - 1)) * nc + c;
  int32_t r = off[x] + bsum[x / kScanBlock];
  for (int k = k0; k <= k1; ++k, q += IJ, p += IJ2) {
    if (ct[q] != 2) pos[p] = r++;
    else pos[p] = -1;
  }
}

}  // namespace

sp_status define_moex(sp_ctx* c, const int32_t* ctp, const int32_t* ctu, const int32_t* ctv,
                      const int32_t* ctw, int32_t* mu, int32_t* mv, int32_t* mw) {
  const Dims& D = c->dm;
  const dim3 blk(64, 4, 1);
  const dim3 grid((unsigned)((D.I + 63) / 64), (unsigned)((D.J + 3) / 4), (unsigned)D.K);
  int slot;
  timer_begin(c, "moex_cells", (double)D.N * (4 * 4 + 3 * 4), &slot);
  hipLaunchKernelGGL(k_moex, grid, blk, 0, c->stream, D.I, D.J, D.K, ctp, ctu, ctv, ctw, mu,
                     mv, mw);
  timer_end(c, slot);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "MoExCell launch");
}

sp_status number_external(sp_ctx* c, int iu, int iv, int iw, const int32_t* ct, int32_t* pos,
                          int32_t* cnt, int32_t* ext, bool timed) {
  const Dims& D = c->dm;
  const int I2 = D.I - iu, J2 = D.J - iv, K2 = D.K - iw;
  const int nc = number_chunks(K2);
  const int64_t n = (int64_t)I2 * J2 * nc;
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  int32_t* bsum = cnt + n;
  const dim3 blk(64, 4, 1);
  const dim3 grid((unsigned)((I2 + 63) / 64), (unsigned)((J2 + 3) / 4), (unsigned)nc);
  int slot = -1;
  if (timed) timer_begin(c, "number_cells", (double)I2 * J2 * K2 * 8, &slot);
  hipLaunchKernelGGL(k_number_count, grid, blk, 0, c->stream, D.I, D.J, I2, J2, K2, nc, ct, cnt);
  hipLaunchKernelGGL(k_number_scan, dim3((unsigned)nb), dim3(kScanThreads), 0, c->stream, n, cnt,
                     bsum);
  hipLaunchKernelGGL(k_number_scan_top, dim3(1), dim3(kScanThreads), 0, c->stream, nb, bsum, ext);
  hipLaunchKernelGGL(k_number_write, grid, blk, 0, c->stream, D.I, D.J, I2, J2, K2, nc, ct, cnt,
                     bsum, pos);
  timer_end(c, slot);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "NumberExternalCell launch");
}

size_t number_scratch(int I2, int J2, int K2) {
  const int64_t n = (int64_t)I2 * J2 * number_chunks(K2);
  return (size_t)(n + (n + kScanBlock - 1) / kScanBlock);
}

}  // namespace sp
