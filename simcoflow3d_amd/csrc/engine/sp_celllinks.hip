// Cell bookkeeping that follows GridPreprocess when a solid moves
// (SolidBody.f90:817-822, Main.f90:207-214):
//
//   DefineMomentumExchangeCell (CutCell.f90:673-751): cut U/V/W cells next
//     to a solid P cell are flagged in MoExCell;
//   NumberExternalCell (CutCell.f90:753-786): the non-solid cells numbered
//     from 0 in the reference's loop order (i outermost, k innermost).
//
// The numbering is a prefix count in an order where each (i, j) column of k
// is contiguous: one pass counts the column totals, one workgroup scans the
// totals in (i, j) order, and a last pass writes each column's numbers.

#include <hip/hip_runtime.h>

#include "sp_device.h"
#include "sp_internal.h"

namespace sp {
namespace {

constexpr int kScanThreads = 1024;

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

// Non-solid cells of column (i, j), k = 1..K2, stored at the column's
// position i-major in the scan order: (i-1)*J2 + (j-1).
__global__ void __launch_bounds__(256) k_number_count(int I, int J, int I2, int J2, int K2,
                                                      const int32_t* ct, int32_t* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > I2 || j > J2) return;
  int64_t q = (int64_t)(i - 1) + (int64_t)I * (j - 1);
  int32_t n = 0;
  for (int k = 1; k <= K2; ++k, q += (int64_t)I * J) n += ct[q] != 2;
  cnt[(int64_t)(i - 1) * J2 + (j - 1)] = n;
}

// Exclusive scan of n column counts in one workgroup; total - 1 (ExtCell)
// into *ext.  Each thread scans a contiguous segment.
__global__ void __launch_bounds__(kScanThreads) k_number_scan(int64_t n, int32_t* cnt,
                                                             int32_t* ext) {
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  const int64_t per = (n + kScanThreads - 1) / kScanThreads;
  const int64_t lo = per * t < n ? per * t : n;
  const int64_t hi = lo + per < n ? lo + per : n;
  int32_t s = 0;
  for (int64_t x = lo; x < hi; ++x) s += cnt[x];
  part[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the thread sums
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = part[t] - s;
  for (int64_t x = lo; x < hi; ++x) {
    const int32_t c = cnt[x];
    cnt[x] = run;
    run += c;
  }
  if (t == kScanThreads - 1) ext[0] = part[t] - 1;
}

__global__ void __launch_bounds__(256) k_number_write(int I, int J, int I2, int J2, int K2,
                                                      const int32_t* ct, const int32_t* off,
                                                      int32_t* pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > I2 || j > J2) return;
  int64_t q = (int64_t)(i - 1) + (int64_t)I * (j - 1);
  int64_t p = (int64_t)(i - 1) + (int64_t)I2 * (j - 1);
  int32_t r = off[(int64_t)(i - 1) * J2 + (j - 1)];
  for (int k = 1; k <= K2; ++k, q += (int64_t)I * J, p += (int64_t)I2 * J2) {
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
                          int32_t* cnt, int32_t* ext) {
  const Dims& D = c->dm;
  const int I2 = D.I - iu, J2 = D.J - iv, K2 = D.K - iw;
  const dim3 blk(64, 4, 1);
  const dim3 grid((unsigned)((I2 + 63) / 64), (unsigned)((J2 + 3) / 4), 1);
  int slot;
  timer_begin(c, "number_cells", (double)I2 * J2 * K2 * 8, &slot);
  hipLaunchKernelGGL(k_number_count, grid, blk, 0, c->stream, D.I, D.J, I2, J2, K2, ct, cnt);
  hipLaunchKernelGGL(k_number_scan, dim3(1), dim3(kScanThreads), 0, c->stream,
                     (int64_t)I2 * J2, cnt, ext);
  hipLaunchKernelGGL(k_number_write, grid, blk, 0, c->stream, D.I, D.J, I2, J2, K2, ct, cnt,
                     pos);
  timer_end(c, slot);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "NumberExternalCell launch");
}

}  // namespace sp
