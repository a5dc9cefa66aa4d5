// Copies between the caller's pageable host arrays and device memory for the
// host-array entry points (sp_project, sp_adjoint, sp_wall_step).
//
// A Fortran host hands over ordinary (pageable) arrays, which the DMA engine
// cannot read directly.  The arrays go through a ring of pinned buffers: a
// few host threads copy chunk n + 1 into its pinned buffer while the DMA
// engine moves chunk n, so the memory copy and the PCIe transfer overlap
// and the memory copy itself runs on several cores.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "sp_internal.h"

namespace sp {

// Rows of a halo array (0:I+1, 0:J+1, 0:K+1) of doubles: a download with a
// RowSpec copies only the entries (1:I, 1:J, 1:K) into the caller's array and
// leaves the others as the caller had them.
struct RowSpec {
  size_t sx, sy;  // elements per row, rows per plane
  int I, J, K;
};

// The interior entries of the element range [off + lo, off + hi) of an array
// (byte offsets, multiples of 8) from `chunk`, which holds bytes [off, ...).
static void copy_interior(char* dst, const char* chunk, size_t off, size_t lo, size_t hi,
                          const RowSpec& r) {
  const size_t e0 = (off + lo) / 8, e1 = (off + hi) / 8;
  if (e1 <= e0) return;
  for (size_t R = e0 / r.sx; R <= (e1 - 1) / r.sx; ++R) {
    const size_t j = R % r.sy, k = R / r.sy;
    if (j < 1 || j > (size_t)r.J || k < 1 || k > (size_t)r.K) continue;
    const size_t a = std::max(e0, R * r.sx + 1), b = std::min(e1, R * r.sx + 1 + (size_t)r.I);
    if (a < b) std::memcpy(dst + a * 8, chunk + (a * 8 - off), (b - a) * 8);
  }
}

struct HostPipe {
  static constexpr int kBufs = 4;
  static constexpr size_t kChunk = size_t(32) << 20;
  char* buf[kBufs] = {};
  hipEvent_t ev[kBufs] = {};
  bool pending[kBufs] = {};
  int next = 0;

  // worker threads for the memory copies
  std::vector<std::thread> workers;
  std::mutex m;
  std::condition_variable cv_job, cv_done;
  uint64_t generation = 0;
  int remaining = 0;
  bool stop = false;
  char* job_dst = nullptr;
  const char* job_src = nullptr;
  size_t job_len = 0;
  size_t job_off = 0;
  const RowSpec* job_rows = nullptr;  // null: plain copy of [0, len)

  // part p of n of [0, len), on 4 KiB boundaries
  static void part(size_t len, int p, int n, size_t* lo, size_t* hi) {
    const size_t pages = (len + 4095) / 4096;
    *lo = std::min(len, pages * (size_t)p / (size_t)n * 4096);
    *hi = std::min(len, pages * (size_t)(p + 1) / (size_t)n * 4096);
  }

  void worker(int id) {
    uint64_t seen = 0;
    for (;;) {
      char* dst;
      const char* src;
      size_t len, off;
      const RowSpec* rows;
      {
        std::unique_lock<std::mutex> lk(m);
        cv_job.wait(lk, [&] { return stop || generation != seen; });
        if (stop) return;
        seen = generation;
        dst = job_dst; src = job_src; len = job_len; off = job_off; rows = job_rows;
      }
      size_t lo, hi;
      part(len, id + 1, (int)workers.size() + 1, &lo, &hi);
      if (hi > lo) run(dst, src, off, lo, hi, rows);
      std::lock_guard<std::mutex> lk(m);
      if (--remaining == 0) cv_done.notify_one();
    }
  }

  // plain: dst[lo, hi) = src[lo, hi); rows: the interior entries of bytes
  // [off + lo, off + hi) of the array at dst from the chunk src
  static void run(char* dst, const char* src, size_t off, size_t lo, size_t hi,
                  const RowSpec* rows) {
    if (rows) copy_interior(dst, src, off, lo, hi, *rows);
    else std::memcpy(dst + lo, src + lo, hi - lo);
  }

  // Copy `len` bytes from src to dst on the calling thread and the workers.
  // With `rows`, dst is the whole array, src holds its bytes [off, off + len)
  // and only interior entries are written.
  void copy(char* dst, const char* src, size_t len, size_t off = 0,
            const RowSpec* rows = nullptr) {
    if (workers.empty() || len < (size_t(1) << 20)) {
      run(dst, src, off, 0, len, rows);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m);
      job_dst = dst; job_src = src; job_len = len; job_off = off; job_rows = rows;
      remaining = (int)workers.size();
      ++generation;
    }
    cv_job.notify_all();
    size_t lo, hi;
    part(len, 0, (int)workers.size() + 1, &lo, &hi);
    if (hi > lo) run(dst, src, off, lo, hi, rows);
    std::unique_lock<std::mutex> lk(m);
    cv_done.wait(lk, [&] { return remaining == 0; });
  }

  ~HostPipe() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv_job.notify_all();
    for (auto& t : workers) t.join();
  }
};

static sp_status pipe_init(sp_ctx* c) {
  if (c->pipe) return SP_OK;
  HostPipe* p = new (std::nothrow) HostPipe();
  if (!p) return fail(c, SP_ERR_OOM, "host copy pipeline");
  for (int b = 0; b < HostPipe::kBufs; ++b) {
    hipError_t e = hipHostMalloc(&p->buf[b], HostPipe::kChunk, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev[b], hipEventDisableTiming);
    if (e != hipSuccess) {
      c->pipe = p;
      pipe_destroy(c);
      return hip_check(c, e, "pinned staging buffers");
    }
  }
  const unsigned hw = std::thread::hardware_concurrency();
  const int nw = (int)std::min(7u, hw > 1 ? hw - 1 : 0u);
  for (int i = 0; i < nw; ++i) p->workers.emplace_back(&HostPipe::worker, p, i);
  c->pipe = p;
  return SP_OK;
}

void pipe_destroy(sp_ctx* c) {
  HostPipe* p = c->pipe;
  if (!p) return;
  for (int b = 0; b < HostPipe::kBufs; ++b) {
    if (p->ev[b]) {
      hipEventSynchronize(p->ev[b]);
      hipEventDestroy(p->ev[b]);
    }
    if (p->buf[b]) hipHostFree(p->buf[b]);
  }
  delete p;
  c->pipe = nullptr;
}

sp_status pipe_upload(sp_ctx* c, void* dst, const void* src, size_t bytes) {
  sp_status s = pipe_init(c);
  if (s) return s;
  HostPipe& P = *c->pipe;
  for (size_t off = 0; off < bytes; off += HostPipe::kChunk) {
    const size_t len = std::min(HostPipe::kChunk, bytes - off);
    const int b = P.next;
    P.next = (P.next + 1) % HostPipe::kBufs;
    hipError_t e = hipSuccess;
    if (P.pending[b]) e = hipEventSynchronize(P.ev[b]);  // its last transfer has finished
    if (e != hipSuccess) return hip_check(c, e, "pinned staging");
    P.copy(P.buf[b], static_cast<const char*>(src) + off, len);
    e = hipMemcpyAsync(static_cast<char*>(dst) + off, P.buf[b], len, hipMemcpyHostToDevice,
                       c->stream);
    if (e == hipSuccess) e = hipEventRecord(P.ev[b], c->stream);
    if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(H2D)");
    P.pending[b] = true;
  }
  return SP_OK;
}

static sp_status download(sp_ctx* c, void* dst, const void* src, size_t bytes,
                          const RowSpec* rows) {
  sp_status s = pipe_init(c);
  if (s) return s;
  HostPipe& P = *c->pipe;
  const size_t nchunk = (bytes + HostPipe::kChunk - 1) / HostPipe::kChunk;
  std::vector<int> slot(nchunk);
  // keep up to kBufs - 1 transfers in flight ahead of the chunk being copied out
  size_t issued = 0;
  auto issue = [&](size_t q) -> hipError_t {
    const size_t off = q * HostPipe::kChunk, len = std::min(HostPipe::kChunk, bytes - off);
    const int b = P.next;
    P.next = (P.next + 1) % HostPipe::kBufs;
    hipError_t e = hipSuccess;
    if (P.pending[b]) e = hipEventSynchronize(P.ev[b]);
    if (e == hipSuccess)
      e = hipMemcpyAsync(P.buf[b], static_cast<const char*>(src) + off, len,
                         hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipEventRecord(P.ev[b], c->stream);
    P.pending[b] = true;
    slot[q] = b;
    return e;
  };
  for (size_t q = 0; q < nchunk; ++q) {
    while (issued < nchunk && issued < q + HostPipe::kBufs - 1) {
      hipError_t e = issue(issued++);
      if (e != hipSuccess) return hip_check(c, e, "hipMemcpyAsync(D2H)");
    }
    const int b = slot[q];
    hipError_t e = hipEventSynchronize(P.ev[b]);
    if (e != hipSuccess) return hip_check(c, e, "pinned staging");
    P.pending[b] = false;
    const size_t off = q * HostPipe::kChunk, len = std::min(HostPipe::kChunk, bytes - off);
    if (rows) P.copy(static_cast<char*>(dst), P.buf[b], len, off, rows);
    else P.copy(static_cast<char*>(dst) + off, P.buf[b], len);
  }
  return SP_OK;
}

sp_status pipe_download(sp_ctx* c, void* dst, const void* src, size_t bytes) {
  return download(c, dst, src, bytes, nullptr);
}

sp_status pipe_download_interior(sp_ctx* c, double* dst, const double* src) {
  const Dims& D = c->dm;
  const RowSpec r{(size_t)D.SX, (size_t)D.SY, D.I, D.J, D.K};
  return download(c, dst, src, (size_t)D.Nh * sizeof(double), &r);
}

}  // namespace sp
