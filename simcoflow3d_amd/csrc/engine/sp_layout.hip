// Interior copies between a host's halo arrays (0:Imax+1,0:Jmax+1,0:Kmax+1)
// and the cell layout (1:Imax,1:Jmax,1:Kmax) the geometry, liquid-field
// and predictor kernels index.  The reference's drivers allocate the level
// set, normals and volume fractions of every Cell with the halo
// (Test/FlowSphere2Phases/FlowSphere2Phases.f90:316-322, 431-440), its
// AllocateVar without it (Main.f90:264-270, 383-387); the entry points take
// either and convert at the boundary.

#include <hip/hip_runtime.h>

#include "sp_internal.h"

namespace sp {
namespace {

__global__ void __launch_bounds__(256) k_halo_to_cell(const double* __restrict__ h,
                                                      double* __restrict__ c, int I, int J,
                                                      int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int i = (int)(n % I);
  const int64_t r = n / I;
  const int j = (int)(r % J), k = (int)(r / J);
  c[n] = h[(i + 1) + (int64_t)(I + 2) * ((j + 1) + (int64_t)(J + 2) * (k + 1))];
}

__global__ void __launch_bounds__(256) k_cell_to_halo(const double* __restrict__ c,
                                                      double* __restrict__ h, int I, int J,
                                                      int64_t N) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int i = (int)(n % I);
  const int64_t r = n / I;
  const int j = (int)(r % J), k = (int)(r / J);
  h[(i + 1) + (int64_t)(I + 2) * ((j + 1) + (int64_t)(J + 2) * (k + 1))] = c[n];
}

}  // namespace

sp_status halo_to_cell(sp_ctx* c, const double* halo, double* cell) {
  const Dims& D = c->dm;
  hipLaunchKernelGGL(k_halo_to_cell, dim3((unsigned)((D.N + 255) / 256)), dim3(256), 0, c->stream,
                     halo, cell, D.I, D.J, D.N);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "halo_to_cell");
}

sp_status cell_to_halo(sp_ctx* c, const double* cell, double* halo) {
  const Dims& D = c->dm;
  hipLaunchKernelGGL(k_cell_to_halo, dim3((unsigned)((D.N + 255) / 256)), dim3(256), 0, c->stream,
                     cell, halo, D.I, D.J, D.N);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SP_OK : hip_check(c, e, "cell_to_halo");
}

}  // namespace sp
