// VTK RectilinearGrid output of the fields PrintResultVTR3D writes
// (PrintResult.f90:97-147, through VTR_open_file / VTR_write_mesh /
// VTR_write_var / VTR_close_file of VTR_mod.f90).
//
// The document has the reference's elements, attributes, array names and
// order, and every value is the field rounded to float32 (real(., kind=4)).
// The numbers are formatted on the GPU: each value becomes a fixed 17-byte
// field " -d.dddddddddE+XX" (10 significant digits), five to a line, so the
// text position of value q is 17 q + q / 5 and every value is formatted by
// its own thread with no scan.  Ten digits computed in double arithmetic are
// within 5e-10 relative of the float32 value, far inside half its spacing
// (3e-8 relative), so a reader gets back exactly that float32.  The
// reference's list-directed output (fmt=*) has a compiler-dependent layout;
// only the values are fixed, and they are the same.

#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sp_internal.h"

namespace sp {
namespace {

constexpr int kVal = 17;  // bytes per formatted value
constexpr int kPerLine = 5;
constexpr int64_t kChunk = 5 << 20;  // values per formatting launch (multiple of 5)

// powers of ten 1e-30 .. 1e55 (decimal literals: correctly rounded)
__constant__ double kP10[86] = {
    1e-30, 1e-29, 1e-28, 1e-27, 1e-26, 1e-25, 1e-24, 1e-23, 1e-22, 1e-21, 1e-20, 1e-19, 1e-18,
    1e-17, 1e-16, 1e-15, 1e-14, 1e-13, 1e-12, 1e-11, 1e-10, 1e-9,  1e-8,  1e-7,  1e-6,  1e-5,
    1e-4,  1e-3,  1e-2,  1e-1,  1e0,   1e1,   1e2,   1e3,   1e4,   1e5,   1e6,   1e7,   1e8,
    1e9,   1e10,  1e11,  1e12,  1e13,  1e14,  1e15,  1e16,  1e17,  1e18,  1e19,  1e20,  1e21,
    1e22,  1e23,  1e24,  1e25,  1e26,  1e27,  1e28,  1e29,  1e30,  1e31,  1e32,  1e33,  1e34,
    1e35,  1e36,  1e37,  1e38,  1e39,  1e40,  1e41,  1e42,  1e43,  1e44,  1e45,  1e46,  1e47,
    1e48,  1e49,  1e50,  1e51,  1e52,  1e53,  1e54,  1e55};

__device__ inline double p10(int k) { return kP10[k + 30]; }

// " -d.dddddddddE+XX" for the float32 value f
__device__ void format_value(float f, char* o) {
  char s[kVal];
  s[0] = ' ';
  s[1] = signbit(f) ? '-' : ' ';
  const double a = fabs((double)f);
  if (!(a <= 3.5e38)) {  // Inf or NaN, right-aligned
    for (int q = 1; q < kVal; ++q) s[q] = ' ';
    const bool nan = a != a;
    const char* t = nan ? "NaN" : "Infinity";
    const int n = nan ? 3 : 8;
    for (int q = 0; q < n; ++q) s[kVal - n + q] = t[q];
    if (!nan && signbit(f)) s[kVal - n - 1] = '-';
  } else {
    uint64_t m = 0;
    int E = 0;
    if (a > 0.0) {
      E = (int)floor(log10(a));
      double sc = a * p10(9 - E);
      if (sc >= 9999999999.5) { E += 1; sc = a * p10(9 - E); }
      else if (sc < 999999999.5) { E -= 1; sc = a * p10(9 - E); }
      m = (uint64_t)llrint(sc);
      if (m >= 10000000000ull) { m /= 10; E += 1; }
    }
    // ten digits of m, most significant first
    char d[10];
#pragma unroll
    for (int q = 9; q >= 0; --q) { d[q] = (char)('0' + (int)(m % 10)); m /= 10; }
    s[2] = d[0];
    s[3] = '.';
#pragma unroll
    for (int q = 1; q < 10; ++q) s[3 + q] = d[q];
    s[13] = 'E';
    s[14] = E < 0 ? '-' : '+';
    const int ae = E < 0 ? -E : E;
    s[15] = (char)('0' + ae / 10);
    s[16] = (char)('0' + ae % 10);
  }
#pragma unroll
  for (int q = 0; q < kVal; ++q) o[q] = s[q];
}

struct VtrBlock {
  const double* src[3];  // components (one for scalars)
  int ncomp;
  int halo[3];           // source q is a (0:I+1, ...) halo array
  int I, J, K;
  int64_t n;             // values in the block
  int64_t q0, nq;        // this launch: values [q0, q0 + nq), q0 a multiple of 5
  int kind;              // 0: cell field, 1: 1-D coordinate array
  char* out;             // text of value q0
};

__global__ void __launch_bounds__(256) k_vtr_format(VtrBlock B) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B.nq) return;
  const int64_t q = B.q0 + r;
  double v;
  if (B.kind == 1) {
    v = B.src[0][q];
  } else {
    const int64_t cell = q / B.ncomp;
    const int comp = (int)(q - cell * B.ncomp);
    const int64_t IJ = (int64_t)B.I * B.J;
    const int k = (int)(cell / IJ);
    const int64_t rem = cell - (int64_t)k * IJ;
    const int j = (int)(rem / B.I), i = (int)(rem - (int64_t)j * B.I);
    const double* a = B.src[comp];
    const int64_t idx = B.halo[comp]
                            ? (i + 1) + (int64_t)(B.I + 2) * ((j + 1) + (int64_t)(B.J + 2) * (k + 1))
                            : cell;
    v = a[idx];
  }
  char* o = B.out + r * kVal + r / kPerLine;
  format_value((float)v, o);
  if (q % kPerLine == kPerLine - 1 || q == B.n - 1) o[kVal] = '\n';
}

int64_t text_bytes(int64_t n) { return n * kVal + (n + kPerLine - 1) / kPerLine; }

struct Writer {
  sp_ctx* c;
  FILE* fh = nullptr;
  char* dtext = nullptr;   // device text of one block
  char* htext = nullptr;   // pinned host copy
  int64_t cap = 0;
  std::string err;

  bool line(const std::string& s) {
    if (std::fputs(s.c_str(), fh) < 0 || std::fputc('\n', fh) < 0) { err = "write failed"; return false; }
    return true;
  }
  // one DataArray's numbers, formatted in chunks of kChunk values
  sp_status block(const VtrBlock& b0) {
    for (int64_t q0 = 0; q0 < b0.n; q0 += kChunk) {
      VtrBlock b = b0;
      b.q0 = q0;
      b.nq = b0.n - q0 < kChunk ? b0.n - q0 : kChunk;
      b.out = dtext;
      const int64_t bytes = text_bytes(b.nq);
      if (bytes > cap) return fail(c, SP_ERR_INVALID_ARG, "VTR chunk larger than its buffer");
      const int64_t nb = (b.nq + 255) / 256;
      int slot;
      timer_begin(c, "vtr_format", (double)b.nq * (8.0 + kVal) + (double)(b.nq / kPerLine), &slot);
      hipLaunchKernelGGL(k_vtr_format, dim3((unsigned)nb), dim3(256), 0, c->stream, b);
      timer_end(c, slot);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(htext, dtext, (size_t)bytes, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) return hip_check(c, e, "VTR formatting");
      if (std::fwrite(htext, 1, (size_t)bytes, fh) != (size_t)bytes)
        return fail(c, SP_ERR_INVALID_ARG, "VTR write failed");
    }
    return SP_OK;
  }
};

// mkdir -p (PrintResultVTR3D creates its result folders the same way,
// PrintResult.f90:111-129)
bool make_dirs(const std::string& path) {
  for (size_t pos = 1; pos <= path.size(); ++pos) {
    if (pos < path.size() && path[pos] != '/') continue;
    const std::string sub = path.substr(0, pos);
    if (mkdir(sub.c_str(), 0775) != 0 && errno != EEXIST) return false;
  }
  return true;
}

std::string data_array(const char* indent, const std::string& name, int ncomp) {
  return std::string(indent) + "<DataArray type=\"Float32\" Name=\"" + name +
         "\" NumberOfComponents=\"" + std::to_string(ncomp) + "\" format=\"ascii\">";
}

}  // namespace

sp_status write_vtr(sp_ctx* c, const char* dir, const char* prefix, int64_t itt,
                    const sp_vtr_fields& f, bool device) {
  const Dims& D = c->dm;
  const int I = D.I, J = D.J, K = D.K;
  const int64_t N = D.N, Nh = D.Nh;
  // <dir>Paraview/<prefix>_<itt, 8 digits>.vtr (VTR_mod.f90:116-130)
  const std::string pdir = std::string(dir) + "Paraview/";
  if (!make_dirs(pdir)) return fail(c, SP_ERR_INVALID_ARG, "cannot create " + pdir);
  char snap[32];
  std::snprintf(snap, sizeof(snap), "%08lld", (long long)itt);
  const std::string path = pdir + prefix + "_" + snap + ".vtr";

  const double* cell[9] = {f.u, f.v, f.w, f.p, f.phi, f.phi_l, f.vof, f.vof_l, f.mres};
  const double* coord[3] = {f.x, f.y, f.z};
  for (const double* p : cell)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "null VTR field");
  for (const double* p : coord)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "null VTR coordinate");
  int halo[9];
  for (int q = 0; q < 9; ++q) halo[q] = (f.halo_mask >> q) & 1u;

  // device copies of host arrays (staging slots past the ones of the
  // projection and the adjoint)
  const double* dcell[9];
  const double* dcoord[3];
  std::vector<void*> tmp;
  auto release = [&]() { for (void* p : tmp) hipFree(p); tmp.clear(); };
  if (device) {
    for (int q = 0; q < 9; ++q) dcell[q] = cell[q];
    for (int q = 0; q < 3; ++q) dcoord[q] = coord[q];
  } else {
    for (int q = 0; q < 12; ++q) {
      const double* src = q < 9 ? cell[q] : coord[q - 9];
      const int64_t n = q < 9 ? (halo[q] ? Nh : N) : (q == 9 ? I : (q == 10 ? J : K));
      void* d = nullptr;
      hipError_t e = hipMalloc(&d, sizeof(double) * (size_t)n);
      if (e != hipSuccess) { release(); return hip_check(c, e, "hipMalloc(VTR staging)"); }
      tmp.push_back(d);
      e = hipMemcpyAsync(d, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) { release(); return hip_check(c, e, "hipMemcpyAsync(VTR)"); }
      if (q < 9) dcell[q] = static_cast<const double*>(d);
      else dcoord[q - 9] = static_cast<const double*>(d);
    }
  }

  Writer w;
  w.c = c;
  w.cap = text_bytes(kChunk);
  hipError_t e = hipMalloc(&w.dtext, (size_t)w.cap);
  if (e == hipSuccess) e = hipHostMalloc(&w.htext, (size_t)w.cap, hipHostMallocDefault);
  if (e != hipSuccess) {
    if (w.dtext) hipFree(w.dtext);
    release();
    return hip_check(c, e, "VTR buffers");
  }
  w.fh = std::fopen(path.c_str(), "w");
  sp_status s = SP_OK;
  if (!w.fh) s = fail(c, SP_ERR_INVALID_ARG, "cannot open " + path);

  auto scalar = [&](const double* src, int hal, int64_t n, int kind) {
    VtrBlock b{};
    b.src[0] = src;
    b.ncomp = 1;
    b.halo[0] = hal;
    b.I = kind == 1 ? (int)n : I; b.J = J; b.K = K;
    b.n = n;
    b.kind = kind;
    return w.block(b);
  };
  const std::string ext = "1 " + std::to_string(I) + " 1 " + std::to_string(J) + " 1 " +
                          std::to_string(K);
  if (!s) {
    bool ok = w.line("<VTKFile type=\"RectilinearGrid\" version=\"0.1\" format=\"ascii\">") &&
              w.line(" <RectilinearGrid WholeExtent=\"" + ext + "\">") &&
              w.line("  <Piece Extent=\"" + ext + "\">") && w.line("   <Coordinates>");
    const char* cname[3] = {"X_COORDINATES", "Y_COORDINATES", "Z_COORDINATES"};
    const int64_t cn[3] = {I, J, K};
    for (int q = 0; ok && !s && q < 3; ++q) {
      ok = w.line(data_array("    ", cname[q], 1));
      if (ok) s = scalar(dcoord[q], 0, cn[q], 1);
      if (ok && !s) ok = w.line("    </DataArray>");
    }
    if (ok && !s) ok = w.line("   </Coordinates>") && w.line("   <PointData>");
    // VTR_write_vector_3d: interleaved vector, then its three components
    if (ok && !s) {
      ok = w.line(data_array("        ", "Velocity", 3));
      VtrBlock b{};
      b.src[0] = dcell[0]; b.src[1] = dcell[1]; b.src[2] = dcell[2];
      b.halo[0] = halo[0]; b.halo[1] = halo[1]; b.halo[2] = halo[2];
      b.ncomp = 3; b.I = I; b.J = J; b.K = K; b.n = 3 * N; b.kind = 0;
      if (ok) s = w.block(b);
      if (ok && !s) ok = w.line("        </DataArray>");
    }
    const char* comp[3] = {"X_Velocity", "Y_Velocity", "Z_Velocity"};
    for (int q = 0; ok && !s && q < 3; ++q) {
      ok = w.line(data_array("        ", comp[q], 1));
      if (ok) s = scalar(dcell[q], halo[q], N, 0);
      if (ok && !s) ok = w.line("        </DataArray>");
    }
    // VTR_write_scalar_3d in PrintResultVTR3D's order
    const char* sname[6] = {"Pressure", "FluidLvs", "LiquidLvs", "FluidVof", "LiquidVof",
                            "Mass error"};
    for (int q = 0; ok && !s && q < 6; ++q) {
      ok = w.line(data_array("         ", sname[q], 1));
      if (ok) s = scalar(dcell[3 + q], halo[3 + q], N, 0);
      if (ok && !s) ok = w.line("         </DataArray>");
    }
    if (ok && !s)
      ok = w.line("     </PointData>") && w.line("   </Piece>") && w.line("  </RectilinearGrid>") &&
           w.line("</VTKFile>");
    if (!ok && !s) s = fail(c, SP_ERR_INVALID_ARG, "write failed: " + path);
  }
  if (w.fh && std::fclose(w.fh) != 0 && !s) s = fail(c, SP_ERR_INVALID_ARG, "close failed: " + path);
  hipFree(w.dtext);
  hipHostFree(w.htext);
  release();
  return s;
}

}  // namespace sp
