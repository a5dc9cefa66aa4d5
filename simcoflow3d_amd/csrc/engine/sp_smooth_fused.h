// Temporally blocked red-black Gauss-Seidel for one multigrid level.
//
// A workgroup of FEX x FEY threads owns a TX x TY tile of (i,j) columns plus a
// halo of H = 2*NU columns on every side and marches over a chunk of k
// planes.  The 2*NU half-sweeps are pipelined one plane apart: at step t the
// plane t is loaded, half-sweep s updates plane t-s+1 (pre) or t-s (post),
// so every array is streamed from HBM once per kernel instead of once per
// half-sweep.  Values along k live in per-thread register windows; the
// in-plane neighbours of a plane come from a ring of LDS planes written one
// step earlier (one barrier per step).  The halo shrinks by one column per
// half-sweep, so the tile's results equal those of the plain sweep sequence:
//   pre  (MODE 0): x = 0; red, black, red, black ...; r = b - A x restricted
//                  (sum over 2x2x2 children) into the coarse right-hand side;
//                  x is written to xout.
//   post (MODE 1): x = xin + P xc (active cells); black, red, black, red ...;
//                  x is written to xout.
// The neighbour sums use the order of the plain kernels (W,S,B,E,N,T).
//
// In both modes the cell (i,j) is updated at step t exactly when i + j + t is
// even (half-sweep s touches plane t - s + 1 with colour s - 1, or plane
// t - s with colour s), so one predicate per step replaces the per-half-sweep
// colour tests, and all lanes run the same instructions.  The march is
// unrolled by the window length W; with the chunk start t0 even and R = W,
// every register-window slot and LDS ring slot is a compile-time constant.
#pragma once

#include <hip/hip_runtime.h>

#include "sp_device.h"

namespace sp {

constexpr int kFEX = 40;  // extended tile width (threads along i)
constexpr int kFEY = 24;  // extended tile height (threads along j)

template <typename T>
struct FusedArgs {
  int nx, ny, nz;
  const T *cx, *cy, *cz, *e;
  const T* b;
  const T* xin;  // post: pre-smoothed x
  const T* xc;   // post: coarse correction
  T* xout;
  T* bc;         // pre: coarse right-hand side
  int cnx, cny;
  int kchunk;    // even
  T omega;       // over-relaxation: x += omega (x_gs - x)
  // finest post-smoother only: the CG dot products (r,z), (z,q), (dw,z) of
  // the z it writes, as per-workgroup partials (r and the diagonal weight:
  // the FP32 right-hand side and diagonal the leg already holds)
  int dots;
  const double* dq;
  RedSite dsite;
};

// Coarse levels too small in (i, j) to fill the chip with extended 40 x 24
// tiles march with 16 x 16 (an 8 x 8 interior at NU = 2): four waves per
// workgroup, several workgroups per CU.
constexpr int kFEXs = 16, kFEYs = 16;

template <int NU, int EX = kFEX, int EY = kFEY>
struct FusedShape {
  static constexpr int H = 2 * NU;
  static constexpr int TX = EX - 2 * H;
  static constexpr int TY = EY - 2 * H;
  static constexpr int W = H + 2;  // register window (planes)
  static constexpr int R = W;      // LDS ring (planes)
  static_assert(TX > 0 && TY > 0 && TX % 2 == 0 && TY % 2 == 0,
                "the restriction pairs cells inside the tile interior");
};

// Barrier that orders LDS only: global loads issued for later planes stay in
// flight across it (the kernel's global stores are never read back inside
// the kernel).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Element at a 32-bit byte offset (every level array is below 4 GiB), so the
// loads use a scalar base plus one 32-bit vector offset.
template <typename T>
__device__ inline T at_off(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <typename T>
__device__ inline void put_off(T* base, uint32_t byte_off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// 1/d of a smoother: FP32 uses the hardware reciprocal (1 ulp), which is
// ample for a preconditioner; FP64 divides exactly.
__device__ inline float smoother_rcp(float d) { return __builtin_amdgcn_rcpf(d); }
__device__ inline double smoother_rcp(double d) { return 1.0 / d; }

// Raw loads of one plane.  The loads always target a valid address (the
// plane index is clamped, neighbours fall back to the cell itself) so their
// destination registers carry straight into the next pipeline step; validity
// is applied when the values are consumed.
template <typename T>
struct PlaneLoad {
  T b, cx, cy, cz, cxm, cym, e, xin, xc;
  bool vt;
  double dq;  // dot-product input of the step's output plane t - H
};

// Per-thread constants of the march.
struct FusedCol {
  int i, j, tx, ty;
  bool in_ij, tile, vxm, vym;
  bool act_even;         // updated at even steps (i + j even)
  uint32_t col;          // byte offset of the (i, j) column in plane 0
  uint32_t plane;        // bytes per plane
  uint32_t dxm, dym;     // byte distance to the -x / -y neighbour (0 at the edge)
  uint32_t ccol;         // post: byte offset of the coarse column
  uint32_t cplane;       // post: bytes per coarse plane
};

template <typename T, int MODE, bool HAS_E, bool DOTS, int H>
__device__ inline void load_plane(PlaneLoad<T>& L, const FusedArgs<T>& A, const FusedCol& C, int t) {
  const int tc = t < 0 ? 0 : (t >= A.nz ? A.nz - 1 : t);
  const uint32_t n = C.col + C.plane * (uint32_t)tc;
  L.vt = C.in_ij && t >= 0 && t < A.nz;
  L.b = at_off(A.b, n);
  L.cx = at_off(A.cx, n);
  L.cy = at_off(A.cy, n);
  L.cz = at_off(A.cz, n);
  L.cxm = at_off(A.cx, n - C.dxm);
  L.cym = at_off(A.cy, n - C.dym);
  L.e = HAS_E ? at_off(A.e, n) : T(0);
  if (MODE == 1) {
    L.xin = at_off(A.xin, n);
    L.xc = at_off(A.xc, C.ccol + C.cplane * (uint32_t)(tc >> 1));
  } else {
    L.xin = T(0);
    L.xc = T(0);
  }
  // with the plane loads, so the in-order load counter never makes a step
  // wait for planes fetched after them
  if (DOTS && A.dots && C.tile && C.in_ij) {
    const int q = t - H, qc = q < 0 ? 0 : (q >= A.nz ? A.nz - 1 : q);
    const uint32_t e = C.col / (uint32_t)sizeof(T) + (C.plane / (uint32_t)sizeof(T)) * (uint32_t)qc;
    L.dq = at_off(A.dq, e * 8u);
  }
}

// Register windows along k.  Logical plane o (0 = newest) of step u lives in
// physical slot (o + W - u % W) % W, so advancing one plane renames slots
// instead of moving registers (the march is unrolled by W).
template <typename T, int NU, int MODE>
struct FusedState {
  static constexpr int H = 2 * NU, W = H + 2;
  T xw[W], bw[W], dw[W], rdw[W], czw[W], cxw[W], cxmw[W], cyw[W], cymw[W];
  bool dpos[W];  // d > 0 of each plane, tested once when the plane enters
  T racc;
  double dacc[3];  // CG dot products (finest post-smoother)
};

// One pipeline step: plane t enters (from L), half-sweeps on the planes
// behind it, output/residual of plane t - H, restriction after the barrier.
// U = position of this step in the W-step unrolled march (t - t0 = U mod W).
template <typename T, int NU, int MODE, bool HAS_E, bool DOTS, int U, int EX, int EY>
__device__ inline void fused_step(FusedState<T, NU, MODE>& F, const PlaneLoad<T>& L,
                                  const FusedArgs<T>& A, const FusedCol& C,
                                  T (*xs)[EY + 2][EX + 2], T (*rs)[EY][EX], int t, int k0, int k1) {
  using Sh = FusedShape<NU, EX, EY>;
  constexpr int H = Sh::H, W = Sh::W, R = Sh::R;
  static_assert(R == W && (W % 2) == 0, "ring slots are compile-time only for R == W, W even");
  auto S = [](int o) constexpr { return (o + W - U) % W; };        // register slot
  auto Q = [](int o) constexpr { return ((U - o) % R + R) % R; };  // LDS ring slot
  // FP32 keeps a window of 1/d (one reciprocal per cell and plane); FP64 has
  // no registers to spare for it and divides in every half-sweep
  constexpr bool kRcp = sizeof(T) == 4;
  const int tx = C.tx, ty = C.ty;
  // LDS planes carry a one-cell border of zeros: cell (tx, ty) is at
  // [ty + 1][tx + 1] and every neighbour read is unconditional
  const int lx = tx + 1, ly = ty + 1;

  const int q = t - H;
  const bool outq = q >= k0 && q < k1;
  const bool wr = outq && C.tile && C.in_ij;

  const T lb = L.vt ? L.b : T(0), lcx = L.vt ? L.cx : T(0), lcy = L.vt ? L.cy : T(0);
  const T lcz = L.vt ? L.cz : T(0), le = L.vt ? L.e : T(0);
  const T lcxm = (L.vt && C.vxm) ? L.cxm : T(0), lcym = (L.vt && C.vym) ? L.cym : T(0);
  T d0 = lcxm;
  d0 = d0 + lcym;
  d0 = d0 + F.czw[S(1)];
  d0 = d0 + lcx;
  d0 = d0 + lcy;
  d0 = d0 + lcz;
  // the diagonal-only term; without one the sum stays as it is (adding +0
  // to a sum of non-negative couplings changes nothing)
  if (HAS_E) d0 = d0 + le;
  T x0 = T(0);
  if (MODE == 1) x0 = d0 > T(0) ? L.xin + L.xc : T(0);
  F.bw[S(0)] = lb; F.cxw[S(0)] = lcx; F.cxmw[S(0)] = lcxm; F.cyw[S(0)] = lcy;
  F.cymw[S(0)] = lcym; F.czw[S(0)] = lcz; F.dw[S(0)] = d0; F.xw[S(0)] = x0;
  F.dpos[S(0)] = d0 > T(0);
  if (kRcp) F.rdw[S(0)] = d0 > T(0) ? smoother_rcp(d0) : T(0);
  xs[Q(0)][ly][lx] = x0;

  // cells of this thread's colour move this step (t0 even, W even)
  const bool act = (U & 1) ? !C.act_even : C.act_even;
#pragma unroll
  for (int s = 1; s <= H; ++s) {
    const int o = MODE == 0 ? s - 1 : s;
    // planes outside [0, nz) have d = 0 and stay 0, so no range test
    T v;
    if (MODE == 0 && s == 1) {
      v = kRcp ? F.bw[S(0)] * F.rdw[S(0)] : F.bw[S(0)] / F.dw[S(0)];
    } else {
      const T xW = xs[Q(o)][ly][lx - 1], xS = xs[Q(o)][ly - 1][lx];
      const T xE = xs[Q(o)][ly][lx + 1], xN = xs[Q(o)][ly + 1][lx];
      T acc = T(0);
      acc += F.cxmw[S(o)] * xW;
      acc += F.cymw[S(o)] * xS;
      acc += F.czw[S(o + 1)] * F.xw[S(o + 1)];
      acc += F.cxw[S(o)] * xE;
      acc += F.cyw[S(o)] * xN;
      acc += F.czw[S(o)] * F.xw[S(o - 1)];
      v = kRcp ? (F.bw[S(o)] + acc) * F.rdw[S(o)] : (F.bw[S(o)] + acc) / F.dw[S(o)];
    }
    v = F.xw[S(o)] + A.omega * (v - F.xw[S(o)]);
    // keep v in a register on every lane: a select whose operand is only
    // needed on some lanes would otherwise become an exec-masked branch
    __asm__ volatile("" : "+v"(v));
    const T xn = (act && F.dpos[S(o)]) ? v : F.xw[S(o)];
    F.xw[S(o)] = xn;
    xs[Q(o)][ly][lx] = xn;
  }

  if (MODE == 0) {
    T r = T(0);
    if (wr && F.dpos[S(H)]) {
      const T xW = xs[Q(H)][ly][lx - 1], xS = xs[Q(H)][ly - 1][lx];
      const T xE = xs[Q(H)][ly][lx + 1], xN = xs[Q(H)][ly + 1][lx];
      T acc = T(0);
      acc += F.cxmw[S(H)] * xW;
      acc += F.cymw[S(H)] * xS;
      acc += F.czw[S(H + 1)] * F.xw[S(H + 1)];
      acc += F.cxw[S(H)] * xE;
      acc += F.cyw[S(H)] * xN;
      acc += F.czw[S(H)] * F.xw[S(H - 1)];
      r = F.bw[S(H)] - (F.dw[S(H)] * F.xw[S(H)] - acc);
    }
    if (wr) put_off(A.xout, C.col + C.plane * (uint32_t)q, F.xw[S(H)]);
    // q = t0 + U - H (mod W) with t0 and H even: parity of q = parity of U
    if (q >= 0) rs[U & 1][ty][tx] = r;
  } else if (wr) {
    put_off(A.xout, C.col + C.plane * (uint32_t)q, F.xw[S(H)]);
    if (DOTS && A.dots) {
      const double z = (double)F.xw[S(H)];
      // (r, z) with the FP32 r the leg smooths with (the b window): an
      // rz rounded at 1e-7 only perturbs the step length, and the x and r
      // updates stay consistent (both use the same alpha)
      F.dacc[0] += (double)F.bw[S(H)] * z;
      F.dacc[1] += z * L.dq;
      F.dacc[2] += (double)F.dw[S(H)] * z;
    }
  }
  lds_barrier();
  if (MODE == 0 && wr && ((tx - H) & 1) == 0 && ((ty - H) & 1) == 0) {
    constexpr int sl = U & 1;
    T ps = rs[sl][ty][tx];
    ps += rs[sl][ty][tx + 1];
    ps += rs[sl][ty + 1][tx];
    ps += rs[sl][ty + 1][tx + 1];
    constexpr bool first = (U & 1) == 0;  // q even
    F.racc = first ? ps : F.racc + ps;
    if (!first || q == A.nz - 1) {
      // C.ccol: byte offset of the coarse column (i/2, j/2)
      put_off(A.bc, C.ccol + C.cplane * (uint32_t)(q >> 1), F.racc);
    }
  }
}

// Planes of loads in flight: FP32 three, FP64 one (register budget; six
// planes in FP32 measured slower).  The window length W is a multiple of it,
// so the buffer of every step of the unrolled march is a compile-time choice.
template <typename T, int NU>
constexpr int fused_pipe() {
  return sizeof(T) == 4 ? (FusedShape<NU>::W % 3 == 0 ? 3 : 2) : 1;
}

// Steps t + U, t + U + 1, ..., t + W - 1 of the march, each followed by the
// load of a later plane into the buffer it used.  CHECK: stop at t1 (the
// last, partial turn only).  Full turns carry no exit test, so the march loop
// has one back-edge whose in-flight loads are the same on every pass: an exit
// path per step would merge into the loop head with fewer loads in flight,
// and the compiler's wait counts there would drain the load pipeline on
// every turn.
template <typename T, int NU, int MODE, bool HAS_E, bool DOTS, bool CHECK, int U, int EX, int EY>
__device__ inline void march_steps(FusedState<T, NU, MODE>& F, PlaneLoad<T> (&LB)[fused_pipe<T, NU>()],
                                   const FusedArgs<T>& A, const FusedCol& C,
                                   T (*xs)[EY + 2][EX + 2], T (*rs)[EY][EX], int t, int t1,
                                   int k0, int k1) {
  if constexpr (U < FusedShape<NU>::W) {
    if (CHECK && t + U >= t1) return;
    constexpr int P = fused_pipe<T, NU>();
    static_assert(FusedShape<NU>::W % P == 0, "load buffers must tile the window");
    PlaneLoad<T>& L = LB[U % P];
    fused_step<T, NU, MODE, HAS_E, DOTS, U, EX, EY>(F, L, A, C, xs, rs, t + U, k0, k1);
    load_plane<T, MODE, HAS_E, DOTS, FusedShape<NU>::H>(L, A, C, t + U + P);
    march_steps<T, NU, MODE, HAS_E, DOTS, CHECK, U + 1, EX, EY>(F, LB, A, C, xs, rs, t, t1, k0, k1);
  }
}

template <typename T, int NU, int MODE, bool FINE, bool HAS_E, int EX = kFEX, int EY = kFEY>
__global__ void __launch_bounds__(EX * EY) k_smooth_fused(FusedArgs<T> A, const Scalars* S) {
  if (S->done) return;
  using Sh = FusedShape<NU, EX, EY>;
  constexpr int H = Sh::H, TX = Sh::TX, TY = Sh::TY, W = Sh::W, R = Sh::R;
  __shared__ T xs[R][EY + 2][EX + 2];
  __shared__ T rs[2][EY][EX];
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  FusedCol C;
  C.tx = threadIdx.x;
  C.ty = threadIdx.y;
  C.i = (int)bx * TX - H + C.tx;
  C.j = (int)by * TY - H + C.ty;
  C.in_ij = C.i >= 0 && C.i < A.nx && C.j >= 0 && C.j < A.ny;
  C.tile = C.tx >= H && C.tx < H + TX && C.ty >= H && C.ty < H + TY;
  C.vxm = C.i > 0;
  C.vym = C.j > 0;
  C.act_even = ((C.i + C.j) & 1) == 0;
  C.col = C.in_ij ? (uint32_t)(C.i + A.nx * C.j) * (uint32_t)sizeof(T) : 0u;
  C.plane = (uint32_t)A.nx * (uint32_t)A.ny * (uint32_t)sizeof(T);
  C.dxm = (C.in_ij && C.vxm) ? (uint32_t)sizeof(T) : 0u;
  C.dym = (C.in_ij && C.vym) ? (uint32_t)(A.nx * sizeof(T)) : 0u;
  C.ccol = C.in_ij ? (uint32_t)((C.i >> 1) + A.cnx * (C.j >> 1)) * (uint32_t)sizeof(T) : 0u;
  C.cplane = (uint32_t)A.cnx * (uint32_t)A.cny * (uint32_t)sizeof(T);
  const int k0 = (int)bz * A.kchunk;
  const int k1 = min(k0 + A.kchunk, A.nz);

  FusedState<T, NU, MODE> F;
#pragma unroll
  for (int o = 0; o < W; ++o) {
    F.xw[o] = F.bw[o] = F.dw[o] = F.rdw[o] = F.czw[o] = F.cxw[o] = F.cxmw[o] = F.cyw[o] =
        F.cymw[o] = T(0);
    F.dpos[o] = false;
  }
  F.racc = T(0);
  F.dacc[0] = F.dacc[1] = F.dacc[2] = 0.0;
  for (int q = C.ty * EX + C.tx; q < R * (EY + 2) * (EX + 2); q += EX * EY) (&xs[0][0][0])[q] = T(0);
  for (int q = C.ty * EX + C.tx; q < 2 * EY * EX; q += EX * EY) (&rs[0][0][0])[q] = T(0);
  // cz of the plane below the first loaded one (bottom face of that plane):
  // logical plane 1 of the first step, physical slot 1
  {
    const int pb = k0 - H - 1;
    if (C.in_ij && pb >= 0 && pb < A.nz) F.czw[1] = at_off(A.cz, C.col + C.plane * (uint32_t)pb);
  }
  __syncthreads();

  // buffer u % P holds the planes of steps t0 + u (mod P)
  const int t0 = k0 - H, t1 = k1 + H;  // steps [t0, t1); t0 even
  constexpr int P = fused_pipe<T, NU>();
  constexpr bool DOTS = FINE && MODE == 1 && sizeof(T) == 4;  // FP64: separate dots kernel
  PlaneLoad<T> LB[P];
#pragma unroll
  for (int u = 0; u < P; ++u) load_plane<T, MODE, HAS_E, DOTS, H>(LB[u], A, C, t0 + u);
  int t = t0;
  for (; t + W <= t1; t += W)
    march_steps<T, NU, MODE, HAS_E, DOTS, false, 0, EX, EY>(F, LB, A, C, xs, rs, t, t1, k0, k1);
  if (t < t1)
    march_steps<T, NU, MODE, HAS_E, DOTS, true, 0, EX, EY>(F, LB, A, C, xs, rs, t, t1, k0, k1);
  if constexpr (DOTS) {
    if (A.dots) block_partials<EX * EY, 3, 0u>(A.dsite, F.dacc);
  }
}

}  // namespace sp
