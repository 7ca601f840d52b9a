// Device-side tile sources shared by the GEMM kernel families: a kernel runs
// `for (t = first(); t < total_tiles; t += stride())` and fetches the problem
// owning tile t.  GemmBatch launches one workgroup per tile (the loop runs once,
// tiles remapped so that each XCD gets a contiguous range); GemmGroup launches a
// persistent grid over a device-resident problem array (remapped the same way
// within every round of gridDim.x tiles).
#pragma once
#include "gemm.h"

namespace mk {

constexpr int NXCD = 8;

// Workgroups are dispatched round-robin over the XCDs.  Give XCD x a
// contiguous range of tile ids so that neighbouring tiles (which share A row
// panels / B column panels) are resident in the same L2.
__device__ __forceinline__ int xcd_remap(int t, int total) {
  const int q = total / NXCD, r = total % NXCD;
  const int x = t % NXCD, local = t / NXCD;
  return x * q + min(x, r) + local;
}

__device__ __forceinline__ int src_first(const GemmBatch& b) { return xcd_remap(blockIdx.x, b.total_tiles); }
__device__ __forceinline__ int src_stride(const GemmBatch& b) { return b.total_tiles; }
__device__ __forceinline__ const GemmProb& src_prob(const GemmBatch& b, int t) {
  int pi = 0;
  while (pi + 1 < b.nprob && t >= b.p[pi + 1].tile_start) ++pi;
  return b.p[pi];
}

// Persistent grid: round r covers tiles [r G, r G + G) (G = gridDim.x); the
// remap gives each XCD a contiguous range of every round.
__device__ __forceinline__ int src_first(const GemmGroup&) { return xcd_remap(blockIdx.x, gridDim.x); }
__device__ __forceinline__ int src_stride(const GemmGroup&) { return gridDim.x; }
__device__ __forceinline__ const GemmProb& src_prob(const GemmGroup& g, int t) {
  int lo = 0, hi = g.nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g.probs[mid].tile_start <= t) lo = mid;
    else hi = mid - 1;
  }
  return g.probs[lo];
}

// Resolve a descriptor pointer: absolute for GemmBatch, (slot, offset) for GemmGroup.
template <class T>
__device__ __forceinline__ T* resolve(const GemmBatch&, T* p) { return p; }
// Every operand lives in device (global) memory.  A pointer rebuilt from an
// integer loses that fact and would be accessed with flat instructions, whose
// completion counts on both the memory and the LDS counters; the round trip
// through the global address space lets the compiler emit global loads and
// stores again.
template <class T>
__device__ __forceinline__ T* as_global(T* p) {
  return (T*)((__attribute__((address_space(1))) T*)p);
}
// The same pointer typed in the global address space, for the access sites.
template <class T>
using gptr_t = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr_t<T> gptr(T* p) {
  return (gptr_t<T>)p;
}
template <class T>
__device__ __forceinline__ T* resolve(const GemmGroup& g, T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const int slot = (int)(v >> 62);
  T* r = slot == 0 ? p : reinterpret_cast<T*>(const_cast<char*>(g.base[slot]) + (v & REL_MASK));
  return as_global(r);
}

// Epilogue C = alpha * acc + beta * C of an FM x FN grid of 16 x 16 accumulator
// blocks, item r of lane (li, lq) at C[row0 + 16 i + li][col0 + 16 j + cq(lq, r)]
// with cq = lq + 4 r (QMAJOR = false: the fp64 MFMA layout) or 4 lq + r
// (QMAJOR = true: the fp32-accumulating MFMA layouts).  With beta != 0 the
// 4 JB values of C of JB consecutive blocks of a row block are all fetched
// before any of them is stored, so their reads overlap (one round trip per
// batch instead of one per element; JB = 2 keeps the batch within the
// registers the main loop leaves free); out-of-range items read a clamped
// in-range address (branch-free) and are not stored.  beta == 0 never reads C.
template <bool QMAJOR, int FM, int FN, class V, int JB = 2>
__device__ __forceinline__ void epilogue_c(const V (&acc)[FM][FN], double* C, int64_t ldc, int M, int N, int row0,
                                           int col0, int li, int lq, double alpha, double beta) {
  static_assert(FN % JB == 0, "batch must divide the block columns");
  gptr_t<double> Cg = gptr(C);
  auto col_of = [&](int j, int r) { return col0 + j * 16 + (QMAJOR ? 4 * lq + r : lq + 4 * r); };
  const bool rd = beta != 0.0;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = row0 + i * 16 + li;
#pragma unroll
    for (int j0 = 0; j0 < FN; j0 += JB) {
      double cin[JB * 4];
      if (rd) {
        const int64_t rc = min(row, M - 1);
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) cin[j * 4 + r] = Cg[rc + (int64_t)min(col_of(j0 + j, r), N - 1) * ldc];
      }
      if (row >= M) continue;
#pragma unroll
      for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = col_of(j0 + j, r);
          if (col >= N) continue;
          double v = alpha * (double)acc[i][j0 + j][r];
          if (rd) v += beta * cin[j * 4 + r];
          Cg[(int64_t)row + (int64_t)col * ldc] = v;
        }
    }
  }
  // Drain the stores here.  Their data registers are reused inside the next
  // tile's K loop, and without this wait the compiler's wait-count analysis
  // carries the pending stores around the tile loop and inserts a full
  // vmcnt(0) wait (on the K step's own global loads) into every K step.
  __builtin_amdgcn_s_waitcnt(0);
}

// Position of tile u of a problem: tile rows are grouped GROUP_M at a time so
// that consecutive tiles reuse the same A row panels.
constexpr int GROUP_M = 8;
__device__ __forceinline__ void tile_coords(const GemmProb& P, int u, int& tm, int& tn) {
  const int per_group = GROUP_M * P.tiles_n;
  const int g = u / per_group;
  const int first_m = g * GROUP_M;
  const int gsize = min(GROUP_M, P.tiles_m - first_m);
  tm = first_m + (u % per_group) % gsize;
  tn = (u % per_group) / gsize;
}

}  // namespace mk
