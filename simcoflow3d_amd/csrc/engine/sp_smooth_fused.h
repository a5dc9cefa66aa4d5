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
  int kchunk;
};

template <int NU>
struct FusedShape {
  static constexpr int H = 2 * NU;
  static constexpr int TX = kFEX - 2 * H;
  static constexpr int TY = kFEY - 2 * H;
  static constexpr int W = H + 2;  // register window (planes)
  static constexpr int R = H + 2;  // LDS ring (planes)
};

// Barrier that orders LDS only: global loads issued for later planes stay in
// flight across it (the kernel's global stores are never read back inside
// the kernel).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Raw loads of one plane.  The loads always target a valid address (the
// plane index is clamped, neighbours fall back to the cell itself) so their
// destination registers carry straight into the next pipeline step; validity
// is applied when the values are consumed.
template <typename T>
struct PlaneLoad {
  T b, cx, cy, cz, cxm, cym, e, xin, xc;
  bool vt, vxm, vym;
};

template <typename T, int MODE, bool HAS_E>
__device__ inline void load_plane(PlaneLoad<T>& L, const FusedArgs<T>& A, bool in_ij, int i, int j,
                                  int64_t col, int64_t plane, int t) {
  const int tc = t < 0 ? 0 : (t >= A.nz ? A.nz - 1 : t);
  const int64_t n = col + plane * tc;
  L.vt = in_ij && t >= 0 && t < A.nz;
  L.vxm = i > 0;
  L.vym = j > 0;
  L.b = A.b[n];
  L.cx = A.cx[n];
  L.cy = A.cy[n];
  L.cz = A.cz[n];
  L.cxm = A.cx[L.vxm && in_ij ? n - 1 : n];
  L.cym = A.cy[L.vym && in_ij ? n - A.nx : n];
  L.e = HAS_E ? A.e[n] : T(0);
  if (MODE == 1) {
    L.xin = A.xin[n];
    const int ic = in_ij ? (i >> 1) : 0, jc = in_ij ? (j >> 1) : 0;
    L.xc = A.xc[ic + (int64_t)A.cnx * (jc + (int64_t)A.cny * (tc >> 1))];
  } else {
    L.xin = T(0);
    L.xc = T(0);
  }
}

// Register windows along k.  Logical plane o (0 = newest) of step u lives in
// physical slot (o + W - u % W) % W, so advancing one plane renames slots
// instead of moving registers (the march is unrolled by W).
template <typename T, int NU, int MODE>
struct FusedState {
  static constexpr int H = 2 * NU, W = H + 2;
  T xw[W], bw[W], dw[W], rdw[W], czw[W], cxw[W], cxmw[W], cyw[W], cymw[W];
  T racc;
};

// One pipeline step: plane t enters (from L), half-sweeps on the planes
// behind it, output/residual of plane t - H, restriction after the barrier.
// U = position of this step in the W-step unrolled march.
template <typename T, int NU, int MODE, int U>
__device__ inline void fused_step(FusedState<T, NU, MODE>& F, const PlaneLoad<T>& L,
                                  const FusedArgs<T>& A, T (*xs)[kFEY][kFEX], T (*rs)[kFEY][kFEX],
                                  int t, int k0, int k1, int i, int j, int tx, int ty, bool in_ij,
                                  bool tile, int64_t col, int64_t plane) {
  using Sh = FusedShape<NU>;
  constexpr int H = Sh::H, W = Sh::W, R = Sh::R;
  auto S = [](int o) constexpr { return (o + W - U) % W; };
  // FP32 keeps a window of 1/d (one division per cell and plane); FP64 has no
  // registers to spare for it and divides in every half-sweep
  constexpr bool kRcp = sizeof(T) == 4;
  const T lb = L.vt ? L.b : T(0), lcx = L.vt ? L.cx : T(0), lcy = L.vt ? L.cy : T(0);
  const T lcz = L.vt ? L.cz : T(0), le = L.vt ? L.e : T(0);
  const T lcxm = (L.vt && L.vxm) ? L.cxm : T(0), lcym = (L.vt && L.vym) ? L.cym : T(0);
  T d0 = lcxm;
  d0 = d0 + lcym;
  d0 = d0 + F.czw[S(1)];
  d0 = d0 + lcx;
  d0 = d0 + lcy;
  d0 = d0 + lcz;
  d0 = d0 + le;
  T x0 = T(0);
  if (MODE == 1 && d0 > T(0)) x0 = L.xin + L.xc;
  F.bw[S(0)] = lb; F.cxw[S(0)] = lcx; F.cxmw[S(0)] = lcxm; F.cyw[S(0)] = lcy;
  F.cymw[S(0)] = lcym; F.czw[S(0)] = lcz; F.dw[S(0)] = d0; F.xw[S(0)] = x0;
  if (kRcp) F.rdw[S(0)] = d0 > T(0) ? T(1) / d0 : T(0);
  xs[((t % R) + R) % R][ty][tx] = x0;

#pragma unroll
  for (int s = 1; s <= H; ++s) {
    const int o = MODE == 0 ? s - 1 : s;
    const int p = t - o;
    const int color = MODE == 0 ? ((s - 1) & 1) : (s & 1);
    if (p < 0 || p >= A.nz || ((i + j + p) & 1) != color || !(F.dw[S(o)] > T(0))) continue;
    T v;
    if (MODE == 0 && s == 1) {
      v = kRcp ? F.bw[S(0)] * F.rdw[S(0)] : F.bw[S(0)] / F.dw[S(0)];
    } else {
      const int slot = ((p % R) + R) % R;
      const T xW = tx > 0 ? xs[slot][ty][tx - 1] : T(0);
      const T xS = ty > 0 ? xs[slot][ty - 1][tx] : T(0);
      const T xE = tx < kFEX - 1 ? xs[slot][ty][tx + 1] : T(0);
      const T xN = ty < kFEY - 1 ? xs[slot][ty + 1][tx] : T(0);
      T acc = T(0);
      acc += F.cxmw[S(o)] * xW;
      acc += F.cymw[S(o)] * xS;
      acc += F.czw[S(o + 1)] * F.xw[S(o + 1)];
      acc += F.cxw[S(o)] * xE;
      acc += F.cyw[S(o)] * xN;
      acc += F.czw[S(o)] * F.xw[S(o - 1)];
      v = kRcp ? (F.bw[S(o)] + acc) * F.rdw[S(o)] : (F.bw[S(o)] + acc) / F.dw[S(o)];
    }
    F.xw[S(o)] = v;
    xs[((p % R) + R) % R][ty][tx] = v;
  }

  const int q = t - H;
  const bool outq = q >= k0 && q < k1;
  if (MODE == 0) {
    T r = T(0);
    if (outq && tile && in_ij) {
      if (F.dw[S(H)] > T(0)) {
        const int slot = ((q % R) + R) % R;
        const T xW = xs[slot][ty][tx - 1], xS = xs[slot][ty - 1][tx];
        const T xE = xs[slot][ty][tx + 1], xN = xs[slot][ty + 1][tx];
        T acc = T(0);
        acc += F.cxmw[S(H)] * xW;
        acc += F.cymw[S(H)] * xS;
        acc += F.czw[S(H + 1)] * F.xw[S(H + 1)];
        acc += F.cxw[S(H)] * xE;
        acc += F.cyw[S(H)] * xN;
        acc += F.czw[S(H)] * F.xw[S(H - 1)];
        r = F.bw[S(H)] - (F.dw[S(H)] * F.xw[S(H)] - acc);
      }
      A.xout[col + plane * q] = F.xw[S(H)];
    }
    if (q >= 0) rs[q & 1][ty][tx] = r;
  } else if (outq && tile && in_ij) {
    A.xout[col + plane * q] = F.xw[S(H)];
  }
  lds_barrier();
  if (MODE == 0 && outq && tile && ((tx - H) & 1) == 0 && ((ty - H) & 1) == 0 && in_ij) {
    const int sl = q & 1;
    T ps = rs[sl][ty][tx];
    ps += rs[sl][ty][tx + 1];
    ps += rs[sl][ty + 1][tx];
    ps += rs[sl][ty + 1][tx + 1];
    const bool first = (q & 1) == 0;
    F.racc = first ? ps : F.racc + ps;
    if (!first || q == A.nz - 1)
      A.bc[(i >> 1) + (int64_t)A.cnx * ((j >> 1) + (int64_t)A.cny * (q >> 1))] = F.racc;
  }
}

// Steps t + U, t + U + 1, ..., t + W - 1 of the march (stopping at t1), each
// followed by the load of the plane two steps ahead into the buffer it used.
template <typename T, int NU, int MODE, bool HAS_E, int U>
__device__ inline void march_steps(FusedState<T, NU, MODE>& F, PlaneLoad<T>& LA, PlaneLoad<T>& LB,
                                   const FusedArgs<T>& A, T (*xs)[kFEY][kFEX],
                                   T (*rs)[kFEY][kFEX], int t, int t1, int k0, int k1, int i,
                                   int j, int tx, int ty, bool in_ij, bool tile, int64_t col,
                                   int64_t plane) {
  if constexpr (U < FusedShape<NU>::W) {
    if (t + U >= t1) return;
    // FP32: two planes of loads in flight; FP64: one (register budget)
    constexpr int kPipe = sizeof(T) == 4 ? 2 : 1;
    PlaneLoad<T>& L = (kPipe == 2 && (U & 1)) ? LB : LA;
    fused_step<T, NU, MODE, U>(F, L, A, xs, rs, t + U, k0, k1, i, j, tx, ty, in_ij, tile, col,
                               plane);
    load_plane<T, MODE, HAS_E>(L, A, in_ij, i, j, col, plane, t + U + kPipe);
    march_steps<T, NU, MODE, HAS_E, U + 1>(F, LA, LB, A, xs, rs, t, t1, k0, k1, i, j, tx, ty,
                                           in_ij, tile, col, plane);
  }
}

template <typename T, int NU, int MODE, bool FINE, bool HAS_E>
__global__ void __launch_bounds__(kFEX * kFEY) k_smooth_fused(FusedArgs<T> A, const Scalars* S) {
  if (S->done) return;
  using Sh = FusedShape<NU>;
  constexpr int H = Sh::H, TX = Sh::TX, TY = Sh::TY, W = Sh::W, R = Sh::R;
  __shared__ T xs[R][kFEY][kFEX];
  __shared__ T rs[2][kFEY][kFEX];
  uint32_t bx, by, bz;
  xcd_block3(bx, by, bz);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = (int)bx * TX - H + tx, j = (int)by * TY - H + ty;
  const bool in_ij = i >= 0 && i < A.nx && j >= 0 && j < A.ny;
  const bool tile = tx >= H && tx < H + TX && ty >= H && ty < H + TY;
  const int k0 = (int)bz * A.kchunk;
  const int k1 = min(k0 + A.kchunk, A.nz);
  const int64_t plane = (int64_t)A.nx * A.ny;
  const int64_t col = in_ij ? (int64_t)i + (int64_t)A.nx * j : 0;

  FusedState<T, NU, MODE> F;
#pragma unroll
  for (int o = 0; o < W; ++o) {
    F.xw[o] = F.bw[o] = F.dw[o] = F.rdw[o] = F.czw[o] = F.cxw[o] = F.cxmw[o] = F.cyw[o] =
        F.cymw[o] = T(0);
  }
  F.racc = T(0);
  for (int q = ty * kFEX + tx; q < R * kFEY * kFEX; q += kFEX * kFEY) (&xs[0][0][0])[q] = T(0);
  for (int q = ty * kFEX + tx; q < 2 * kFEY * kFEX; q += kFEX * kFEY) (&rs[0][0][0])[q] = T(0);
  // cz of the plane below the first loaded one (bottom face of that plane):
  // logical plane 1 of the first step, physical slot 1
  {
    const int pb = k0 - H - 1;
    if (in_ij && pb >= 0 && pb < A.nz) F.czw[1] = A.cz[col + plane * pb];
  }
  __syncthreads();

  // FP32: two planes of loads in flight, buffer LA holds even steps, LB odd
  // steps (W is even, so the parity of a step's position in the unrolled
  // march is the parity of the step).  FP64: LA only.
  const int t0 = k0 - H, t1 = k1 + H;  // steps [t0, t1)
  PlaneLoad<T> LA, LB;
  load_plane<T, MODE, HAS_E>(LA, A, in_ij, i, j, col, plane, t0);
  if (sizeof(T) == 4) load_plane<T, MODE, HAS_E>(LB, A, in_ij, i, j, col, plane, t0 + 1);
  for (int t = t0; t < t1; t += W) {
    march_steps<T, NU, MODE, HAS_E, 0>(F, LA, LB, A, xs, rs, t, t1, k0, k1, i, j, tx, ty, in_ij,
                                       tile, col, plane);
  }
}

}  // namespace sp
