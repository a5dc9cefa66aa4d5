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

// ---------------------------------------------------------------------------
// PrintResultVTK (PrintResult.f90:328-420): VTKFortran's binary XML
// RectilinearGrid.  Each DataArray's content is the base64 code of the byte
// stream [int32 4 n, float32 value 0 .. n-1]; thread g encodes stream bytes
// 3g .. 3g+2 into characters 4g .. 4g+3, reading the (at most two) values
// those bytes come from, so the text of every array is made in one pass with
// no scan.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kGroups = int64_t(1) << 22;  // base64 groups per launch (16 MiB of text)

__constant__ char kB64[64] = {
    'A', 'B', 'C', 'D', 'E', 'F', 'G', 'H', 'I', 'J', 'K', 'L', 'M', 'N', 'O', 'P',
    'Q', 'R', 'S', 'T', 'U', 'V', 'W', 'X', 'Y', 'Z', 'a', 'b', 'c', 'd', 'e', 'f',
    'g', 'h', 'i', 'j', 'k', 'l', 'm', 'n', 'o', 'p', 'q', 'r', 's', 't', 'u', 'v',
    'w', 'x', 'y', 'z', '0', '1', '2', '3', '4', '5', '6', '7', '8', '9', '+', '/'};

struct VtkBlock {
  const double* src;    // cell field, or cell centres for a coordinate array
  const double* width;  // cell widths for a coordinate array, else null
  int halo;             // src is a (0:I+1, ...) halo array
  int I, J, K;
  int64_t n;            // float32 values in the array
  int64_t g0, ng;       // this launch: groups [g0, g0 + ng)
  char* out;            // text of group g0
};

// value m of the array, rounded to float32 as real(., kind=R4P) does
__device__ inline float vtk_value(const VtkBlock& B, int64_t m) {
  if (B.width) {
    // faces: x(i) = xc(i) - dx(i)/2, x(n) = xc(n-1) + dx(n-1)/2 (0-based)
    const int64_t last = B.n - 1;
    const double x = m < last ? B.src[m] - B.width[m] / 2.0
                              : B.src[last - 1] + B.width[last - 1] / 2.0;
    return (float)x;
  }
  const int64_t IJ = (int64_t)B.I * B.J;
  const int k = (int)(m / IJ);
  const int64_t rem = m - (int64_t)k * IJ;
  const int j = (int)(rem / B.I), i = (int)(rem - (int64_t)j * B.I);
  const int64_t idx =
      B.halo ? (i + 1) + (int64_t)(B.I + 2) * ((j + 1) + (int64_t)(B.J + 2) * (k + 1)) : m;
  return (float)B.src[idx];
}

// 32-bit word w of the stream (little endian): the byte count, then the values
__device__ inline uint32_t vtk_word(const VtkBlock& B, int64_t w) {
  if (w == 0) return (uint32_t)(4 * B.n);
  if (w > B.n) return 0u;
  return __float_as_uint(vtk_value(B, w - 1));
}

__global__ void __launch_bounds__(256) k_vtk_base64(VtkBlock B) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B.ng) return;
  const int64_t L = 4 * (B.n + 1);  // stream bytes
  const int64_t b0 = 3 * (B.g0 + r);
  const int64_t w0 = b0 >> 2;
  const int s = (int)(b0 & 3);
  // bytes s .. s+2 of words w0, w0+1 (the second is needed only when s >= 2)
  uint64_t pair = vtk_word(B, w0);
  if (s >= 2) pair |= (uint64_t)vtk_word(B, w0 + 1) << 32;
  const uint32_t n1 = b0 + 1 < L, n2 = b0 + 2 < L;  // bytes 2 and 3 of the group exist
  const uint32_t c0 = (uint32_t)(pair >> (8 * s)) & 255u;
  const uint32_t c1 = n1 ? (uint32_t)(pair >> (8 * s + 8)) & 255u : 0u;
  const uint32_t c2 = n2 ? (uint32_t)(pair >> (8 * s + 16)) & 255u : 0u;
  const uint32_t v = c0 << 16 | c1 << 8 | c2;
  const uint32_t a0 = (uint8_t)kB64[v >> 18];
  const uint32_t a1 = (uint8_t)kB64[(v >> 12) & 63];
  const uint32_t a2 = n1 ? (uint8_t)kB64[(v >> 6) & 63] : (uint32_t)'=';
  const uint32_t a3 = n2 ? (uint8_t)kB64[v & 63] : (uint32_t)'=';
  reinterpret_cast<uint32_t*>(B.out)[r] = a0 | a1 << 8 | a2 << 16 | a3 << 24;
}

// PENF's str for an integer: '+' before a non-negative value
std::string penf_str(int64_t n) { return (n >= 0 ? "+" : "") + std::to_string(n); }

}  // namespace

sp_status write_vtk(sp_ctx* c, const char* dir, int64_t itt, const sp_vtk_fields& f,
                    bool device) {
  const Dims& D = c->dm;
  const int I = D.I, J = D.J, K = D.K;
  const int64_t N = D.N, Nh = D.Nh;
  if (4 * (N + 1) > (int64_t)INT32_MAX)
    return fail(c, SP_ERR_INVALID_ARG, "VTK binary arrays carry a 32-bit byte count: grid too large");
  const double* cell[8] = {f.p, f.u, f.v, f.w, f.phi_l, f.phi, f.vof_l, f.vof};
  const double* coord[6] = {f.x, f.y, f.z, f.dx, f.dy, f.dz};
  for (const double* p : cell)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "null VTK field");
  for (const double* p : coord)
    if (!p) return fail(c, SP_ERR_INVALID_ARG, "null VTK coordinate or width");
  int halo[8];
  for (int q = 0; q < 8; ++q) halo[q] = (f.halo_mask >> q) & 1u;

  // 'PCell_Binary_3D'//curd//'.vtr', curd = itt in i8.8
  char curd[16];
  if (itt >= 0 && itt <= 99999999) std::snprintf(curd, sizeof(curd), "%08lld", (long long)itt);
  else std::snprintf(curd, sizeof(curd), "********");
  const std::string sdir(dir);
  if (!sdir.empty() && !make_dirs(sdir)) return fail(c, SP_ERR_INVALID_ARG, "cannot create " + sdir);
  const std::string path = sdir + "PCell_Binary_3D" + curd + ".vtr";

  // device copies of host arrays
  const double* dcell[8];
  const double* dcoord[6];
  std::vector<void*> tmp;
  auto release = [&]() { for (void* p : tmp) hipFree(p); tmp.clear(); };
  if (device) {
    for (int q = 0; q < 8; ++q) dcell[q] = cell[q];
    for (int q = 0; q < 6; ++q) dcoord[q] = coord[q];
  } else {
    for (int q = 0; q < 14; ++q) {
      const double* src = q < 8 ? cell[q] : coord[q - 8];
      const int axis = (q - 8) % 3;
      const int64_t n = q < 8 ? (halo[q] ? Nh : N) : (axis == 0 ? I : (axis == 1 ? J : K));
      void* d = nullptr;
      hipError_t e = hipMalloc(&d, sizeof(double) * (size_t)n);
      if (e != hipSuccess) { release(); return hip_check(c, e, "hipMalloc(VTK staging)"); }
      tmp.push_back(d);
      e = hipMemcpyAsync(d, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) { release(); return hip_check(c, e, "hipMemcpyAsync(VTK)"); }
      if (q < 8) dcell[q] = static_cast<const double*>(d);
      else dcoord[q - 8] = static_cast<const double*>(d);
    }
  }

  char* dtext = nullptr;
  char* htext = nullptr;
  const size_t cap = (size_t)(4 * kGroups);
  hipError_t e = hipMalloc(&dtext, cap);
  if (e == hipSuccess) e = hipHostMalloc(&htext, cap, hipHostMallocDefault);
  if (e != hipSuccess) {
    if (dtext) hipFree(dtext);
    release();
    return hip_check(c, e, "VTK buffers");
  }
  FILE* fh = std::fopen(path.c_str(), "wb");
  sp_status s = SP_OK;
  if (!fh) s = fail(c, SP_ERR_INVALID_ARG, "cannot open " + path);
  bool ok = true;
  auto put = [&](const std::string& t) {
    if (ok && std::fwrite(t.data(), 1, t.size(), fh) != t.size()) ok = false;
  };
  // one DataArray tag: start tag, content line, end tag (FoXy's layout)
  auto data_array = [&](const std::string& name, const VtkBlock& b0) {
    put("        <DataArray type=\"Float32\" NumberOfComponents=\"1\" Name=\"" + name +
        "\" format=\"binary\">\n          ");
    const int64_t G = (4 * (b0.n + 1) + 2) / 3;
    for (int64_t g0 = 0; ok && !s && g0 < G; g0 += kGroups) {
      VtkBlock b = b0;
      b.g0 = g0;
      b.ng = G - g0 < kGroups ? G - g0 : kGroups;
      b.out = dtext;
      int slot;
      // algorithmic bytes: 3/4 of a source double (two for a coordinate) read
      // and 4 characters written per group
      timer_begin(c, "vtk_base64", (double)b.ng * ((b.width ? 12.0 : 6.0) + 4.0), &slot);
      hipLaunchKernelGGL(k_vtk_base64, dim3((unsigned)((b.ng + 255) / 256)), dim3(256), 0,
                         c->stream, b);
      timer_end(c, slot);
      hipError_t e2 = hipGetLastError();
      if (e2 == hipSuccess)
        e2 = hipMemcpyAsync(htext, dtext, (size_t)(4 * b.ng), hipMemcpyDeviceToHost, c->stream);
      if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
      if (e2 != hipSuccess) { s = hip_check(c, e2, "VTK base64"); break; }
      if (std::fwrite(htext, 1, (size_t)(4 * b.ng), fh) != (size_t)(4 * b.ng)) ok = false;
    }
    put("\n        </DataArray>\n");
  };
  if (!s) {
    const std::string ext = penf_str(1) + " " + penf_str(I + 1) + " " + penf_str(1) + " " +
                            penf_str(J + 1) + " " + penf_str(1) + " " + penf_str(K + 1);
    put("<?xml version=\"1.0\"?>\n"
        "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
    put("  <RectilinearGrid WholeExtent=\"" + ext + "\">\n");
    put("    <Piece Extent=\"" + ext + "\">\n");
    put("      <Coordinates>\n");
    const char* cname[3] = {"X", "Y", "Z"};
    const int cn[3] = {I, J, K};
    for (int q = 0; q < 3 && ok && !s; ++q) {
      VtkBlock b{};
      b.src = dcoord[q];
      b.width = dcoord[3 + q];
      b.n = cn[q] + 1;
      data_array(cname[q], b);
    }
    put("      </Coordinates>\n");
    put("      <CellData>\n");
    const char* name[8] = {"p", "u", "v", "w", "LiquidLvs", "FluidLvs", "LiquidVof", "FluidVof"};
    for (int q = 0; q < 8 && ok && !s; ++q) {
      VtkBlock b{};
      b.src = dcell[q];
      b.halo = halo[q];
      b.I = I; b.J = J; b.K = K;
      b.n = N;
      data_array(name[q], b);
    }
    put("      </CellData>\n");
    put("    </Piece>\n");
    put("  </RectilinearGrid>\n");
    put("</VTKFile>\n");
    if (!ok && !s) s = fail(c, SP_ERR_INVALID_ARG, "write failed: " + path);
  }
  if (fh && std::fclose(fh) != 0 && !s) s = fail(c, SP_ERR_INVALID_ARG, "close failed: " + path);
  hipFree(dtext);
  hipHostFree(htext);
  release();
  return s;
}

}  // namespace sp
