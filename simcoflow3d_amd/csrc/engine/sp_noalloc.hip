// Allocation-free NumberExternalCell (sp_number_external_cell_noalloc).
//
// The numbering's piece counts, workgroup totals and scan (sp_celllinks.hip)
// need number_scratch(Imax-iu, Jmax-iv, Kmax-iw) int32 of device scratch.
// sp_number_external_cell takes it from the context's staging buffers, which
// hipMalloc on first use; this variant takes it from `scratch`, a static
// device array that every GPU the library runs on holds from module load to
// process end, so the call allocates nothing.  Its device address on each
// GPU is looked up once (noalloc_init, from sp_create with that device
// current) and used instead of the array's name in host code.
//
// Every call on one GPU shares that GPU's array, whatever the context:
// calls that overlap on one GPU (two host threads, or two contexts' streams
// running at once) race on it.  Every scratch element a call reads is
// written earlier in the same call (the count pass writes every piece, the
// workgroup scan every total), so nothing needs clearing between calls.

#include <hip/hip_runtime.h>

#include <mutex>

#include "sp_internal.h"

namespace sp {

// 2^22 int32 = 16 MiB per GPU.
constexpr size_t kNoAllocScratchInts = size_t(1) << 22;
static __device__ __align__(256) int32_t scratch[kNoAllocScratchInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[kMaxDevices];  // scratch's address on each device
}  // namespace

hipError_t noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<int32_t*>(p);
  return hipSuccess;
}

static int32_t* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace sp

using namespace sp;

extern "C" sp_status sp_number_external_cell_noalloc(sp_ctx* c, int iu, int iv, int iw,
                                                     const int32_t* ct, int32_t* posnu,
                                                     int32_t* ext_cell) {
  if (!c) return SP_ERR_INVALID_ARG;
  if (!ct || !posnu || !ext_cell)
    return fail(c, SP_ERR_INVALID_ARG, "null Cell_Type, PosNu or ExtCell pointer");
  if ((iu != 0 && iu != 1) || (iv != 0 && iv != 1) || (iw != 0 && iw != 1))
    return fail(c, SP_ERR_INVALID_ARG, "iu, iv, iw must be 0 or 1");
  const Dims& D = c->dm;
  const size_t need = number_scratch(D.I - iu, D.J - iv, D.K - iw);
  if (need > kNoAllocScratchInts)
    return fail(c, SP_ERR_INVALID_ARG,
                "grid too large for sp_number_external_cell_noalloc: it needs more than the "
                "2^22 int32 of static scratch (use sp_number_external_cell)");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_check(c, e, "hipSetDevice");
  int32_t* s = scratch_ptr(c->device);
  if (!s) return fail(c, SP_ERR_HIP, "static scratch not set up on this device");
  return number_external(c, iu, iv, iw, ct, posnu, s, ext_cell, false);
}
