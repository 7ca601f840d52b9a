// FP64 MFMA GEMM for CDNA4 (gfx950): v_mfma_f64_16x16x4_f64, LDS-staged
// BK = 16 (or 32) deep tiles, double-buffered through registers, 256 threads =
// 4 waves of 64 lanes in a 2x2 arrangement.  One workgroup computes one BM x BN output tile
// of one problem, summing every segment of that problem into the same
// accumulators before a single fused epilogue (alpha/beta and optional CB term).
#include "gemm_dev.h"
#include "prof.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace mk {
namespace {

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef double dx2 __attribute__((ext_vector_type(2)));

constexpr int NT = 256;
constexpr int BK32 = 32;  // default K depth of a stage
constexpr int PAD = 16;     // LDS row padding (doubles): alternates 128-B bank halves

// Stage a tile whose contiguous memory direction is the LDS fast direction.
// element (r, k) lives at p[(r0 + r) + (k0 + k) * ld]; LDS tile S[k][r].
template <int R, int LDR, int BK, class Ptr>
__device__ __forceinline__ void load_fastdir(Ptr __restrict__ p, int64_t ld, int r0,
                                             int k0, int Rlim, int Klim, bool vec,
                                             dx2 (&reg)[R * BK / 2 / NT]) {
  constexpr int CH = R * BK / 2 / NT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * NT;
    const int k = id / (R / 2);
    const int r = (id % (R / 2)) * 2;
    const int gr = r0 + r, gk = k0 + k;
    dx2 v = {0.0, 0.0};
    if (gk < Klim) {
      const auto q = p + (int64_t)gr + (int64_t)gk * ld;
      if (vec && gr + 1 < Rlim) {
        v = *(const __attribute__((address_space(1))) dx2*)q;
      } else {
        if (gr < Rlim) v.x = q[0];
        if (gr + 1 < Rlim) v.y = q[1];
      }
    }
    reg[c] = v;
  }
}

template <int R, int LDR, int BK>
__device__ __forceinline__ void store_fastdir(double (*S)[LDR], const dx2 (&reg)[R * BK / 2 / NT]) {
  constexpr int CH = R * BK / 2 / NT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * NT;
    const int k = id / (R / 2);
    const int r = (id % (R / 2)) * 2;
    *reinterpret_cast<dx2*>(&S[k][r]) = reg[c];
  }
}

// Stage a tile whose contiguous memory direction is k.
// element (r, k) lives at p[(k0 + k) + (r0 + r) * ld]; LDS tile S[k][r ^ kswz(k)].
// The 16 lanes that share a column write rows k = 0, 2, ..., 30; with a row
// length that is a multiple of 32 doubles they would all hit one bank pair, so
// the column index is XORed with bits 1..4 of k (a permutation inside each
// aligned 32-column group, undone by the fragment reads).
__device__ __forceinline__ int kswz(int k) { return k & 30; }

template <int R, int LDR, int BK, class Ptr>
__device__ __forceinline__ void load_kdir(Ptr __restrict__ p, int64_t ld, int r0, int k0,
                                          int Rlim, int Klim, bool vec,
                                          dx2 (&reg)[R * BK / 2 / NT]) {
  constexpr int CH = R * BK / 2 / NT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * NT;
    const int r = id / (BK / 2);
    const int k = (id % (BK / 2)) * 2;
    const int gr = r0 + r, gk = k0 + k;
    dx2 v = {0.0, 0.0};
    if (gr < Rlim) {
      const auto q = p + (int64_t)gk + (int64_t)gr * ld;
      if (vec && gk + 1 < Klim) {
        v = *(const __attribute__((address_space(1))) dx2*)q;
      } else {
        if (gk < Klim) v.x = q[0];
        if (gk + 1 < Klim) v.y = q[1];
      }
    }
    reg[c] = v;
  }
}

template <int R, int LDR, int BK>
__device__ __forceinline__ void store_kdir(double (*S)[LDR], const dx2 (&reg)[R * BK / 2 / NT]) {
  constexpr int CH = R * BK / 2 / NT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * NT;
    const int r = id / (BK / 2);
    const int k = (id % (BK / 2)) * 2;
    S[k][r ^ kswz(k)] = reg[c].x;
    S[k + 1][r ^ kswz(k)] = reg[c].y;
  }
}

template <int BM, int BN, bool TA, bool TB, class Src, int BK = BK32>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(BK == 16 ? 2 : 1)))
void gemm_f64_kernel(Src src) {
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int CHA = BM * BK / 2 / NT;
  constexpr int CHB = BN * BK / 2 / NT;
  static_assert(CHA >= 1 && CHB >= 1, "tile too small for 256 threads");

  __shared__ double As[2][BK][LDA_S];
  __shared__ double Bs[2][BK][LDB_S];

  for (int t = src_first(src); t < src.total_tiles; t += src_stride(src)) {
  const GemmProb& P = src_prob(src, t);
  int tm, tn;
  tile_coords(P, t - P.tile_start, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  const int li = lane & 15;   // row (A) / column (B, D) inside a 16x16 block
  const int lk = lane >> 4;   // k slot (A, B) / row quad (D)

  dx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dx4{0.0, 0.0, 0.0, 0.0};

  int nsteps = 0;
  for (int s = 0; s < P.nseg; ++s) nsteps += (P.seg[s].K + BK - 1) / BK;

  dx2 ra[CHA], rb[CHB];
  // descriptor of the segment being loaded, read from the problem array once
  // per segment rather than once per K step
  typedef const __attribute__((address_space(1))) double* gptr;   // operands are in device memory
  struct Cur {
    gptr A;
    gptr B;
    int64_t lda, ldb;
    int K;
    bool vec;
  } cur{};
  int cur_s = -1;
  const int M = P.M, N = P.N;
  auto load = [&](int s, int k0) {
    if (s != cur_s) {
      const GemmSeg& S = P.seg[s];
      cur.A = (gptr)static_cast<const double*>(resolve(src, S.A));
      cur.B = (gptr)static_cast<const double*>(resolve(src, S.B));
      cur.lda = S.lda;
      cur.ldb = S.ldb;
      cur.K = S.K;
      cur.vec = S.vec != 0;
      cur_s = s;
    }
    if (!TA) load_fastdir<BM, LDA_S, BK>(cur.A, cur.lda, m0, k0, M, cur.K, cur.vec, ra);
    else     load_kdir<BM, LDA_S, BK>(cur.A, cur.lda, m0, k0, M, cur.K, cur.vec, ra);
    if (!TB) load_kdir<BN, LDB_S, BK>(cur.B, cur.ldb, n0, k0, N, cur.K, cur.vec, rb);
    else     load_fastdir<BN, LDB_S, BK>(cur.B, cur.ldb, n0, k0, N, cur.K, cur.vec, rb);
  };
  auto store = [&](int buf) {
    if (!TA) store_fastdir<BM, LDA_S, BK>(As[buf], ra);
    else     store_kdir<BM, LDA_S, BK>(As[buf], ra);
    if (!TB) store_kdir<BN, LDB_S, BK>(Bs[buf], rb);
    else     store_fastdir<BN, LDB_S, BK>(Bs[buf], rb);
  };

  int s = 0, k0 = 0;
  if (nsteps > 0) {
    load(0, 0);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  int kcur = nsteps > 0 ? P.seg[0].K : 0;
  for (int step = 0; step < nsteps; ++step) {
    int ns = s, nk0 = k0 + BK;
    if (nk0 >= kcur) {
      ns = s + 1;
      nk0 = 0;
      kcur = ns < P.nseg ? P.seg[ns].K : 0;
    }
    const bool has_next = step + 1 < nsteps;
    if (has_next) load(ns, nk0);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      double a[FM], b[FN];
      const int kr = kk * 4 + lk;
      const int sa = TA ? kswz(kr) : 0;    // K-contiguous operands were stored swizzled
      const int sb = TB ? 0 : kswz(kr);
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = As[buf][kr][(wm0 + i * 16 + li) ^ sa];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = Bs[buf][kr][(wn0 + j * 16 + li) ^ sb];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    if (has_next) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
    s = ns;
    k0 = nk0;
  }

  // epilogue.  The v_mfma_f64_16x16x4f64 D layout puts item r of lane (li, lk)
  // at block row lk + 4 r, column li; with the B fragment as first operand the
  // block is C^T, so the lane holds C[row li][col lk + 4 r]: the 16 lanes of a
  // quarter-wave store 16 consecutive rows (128 B) of one column.
  epilogue_c<false, FM, FN, dx4, 1>(acc, resolve(src, P.C), P.ldc, P.M, P.N, m0 + wm0, n0 + wn0, li, lk, P.alpha, P.beta);
  }  // tiles
}

// ---------------------------------------------------------------------------
// Same tile, waves, LDS layout and epilogue as gemm_f64_kernel (BK = 16), with
// a branch-free load path: each operand is read through a buffer resource based
// at the tile's first row, the per-lane byte offsets of a segment are computed
// once per segment, and a chunk outside the operand (row past M / N, k past K)
// gets an offset past num_records, for which the hardware returns zeros.  A K
// step then issues its eight 16-byte loads with two VALU ops each and no
// exec-mask branches.  Needs (checked on the host, buf_ok): 16-byte aligned
// operands, even leading dimensions, even K, an even row count for operands
// stored along their rows, and every in-tile offset below 2 GiB.
// ---------------------------------------------------------------------------
constexpr unsigned OOB = 0x80000000u;
typedef unsigned int ux4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff, 0x00020000);
}

template <int BM, int BN, bool TA, bool TB, class Src, bool ROT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2)))
void gemm_f64_buf_kernel(Src src) {
  constexpr int BK = 16;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int CHA = BM * BK / 2 / NT;
  constexpr int CHB = BN * BK / 2 / NT;
  static_assert(CHA >= 1 && CHB >= 1, "tile too small for 256 threads");
  // A is K-contiguous when transposed, B when not
  constexpr bool AK = TA, BKD = !TB;

  __shared__ double As[2][BK][LDA_S];
  __shared__ double Bs[2][BK][LDB_S];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  const int li = lane & 15;
  const int lk = lane >> 4;

  for (int t = src_first(src); t < src.total_tiles; t += src_stride(src)) {
  const GemmProb& P = src_prob(src, t);
  int tm, tn;
  tile_coords(P, t - P.tile_start, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int M = P.M, N = P.N;

  dx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dx4{0.0, 0.0, 0.0, 0.0};

  int nsteps = 0;
  for (int s = 0; s < P.nseg; ++s) nsteps += (P.seg[s].K + BK - 1) / BK;

  // Chunk c of a thread (the mapping of load_fastdir / load_kdir):
  //   along rows (R rows): rows r0, r0 + 1 with r0 = 2 (t % (R/2)), depth
  //     t / (R/2) + c NT / (R/2): one row pair for all c, depth steps uniform;
  //   along k: row t / 8 + 32 c, depths 2 (t % 8), + 1: one depth pair for
  //     all c, row steps uniform.
  // So a segment keeps one byte offset per operand (chunk 0 at k0 = 0, OOB
  // if its rows lie outside the operand) and uniform strides; the per-chunk
  // offsets and the bounds tests are formed at each load.  (Per-chunk offset
  // arrays held 16 more registers through the K loop, which at this kernel's
  // 256-register budget made the allocator alias load destinations and wait
  // on the global loads inside the step.)
  constexpr int RA2 = BM / 2, RB2 = BN / 2;
  const int tid = threadIdx.x;
  __amdgpu_buffer_rsrc_t rA, rB;
  unsigned oA = 0, oB = 0;         // chunk 0, k0 = 0
  unsigned cA = 0, cB = 0;         // bytes from chunk c to chunk c + 1
  unsigned stepA = 0, stepB = 0;   // bytes per unit of k0
  int K = 0;
  auto set_seg = [&](int s) {
    const GemmSeg& S = P.seg[s];
    const char* pa = static_cast<const char*>(resolve(src, S.A));
    const char* pb = static_cast<const char*>(resolve(src, S.B));
    rA = raw_rsrc(pa + (AK ? (int64_t)m0 * S.lda : (int64_t)m0) * 8);
    rB = raw_rsrc(pb + (BKD ? (int64_t)n0 * S.ldb : (int64_t)n0) * 8);
    stepA = AK ? 8u : (unsigned)(S.lda * 8);
    stepB = BKD ? 8u : (unsigned)(S.ldb * 8);
    if (AK) {
      cA = (unsigned)((NT / 8) * S.lda * 8);
      oA = (unsigned)((2 * (tid % 8) + (int64_t)(tid / 8) * S.lda) * 8);
    } else {
      cA = (unsigned)((NT / RA2) * S.lda * 8);
      oA = m0 + 2 * (tid % RA2) < M ? (unsigned)((2 * (tid % RA2) + (int64_t)(tid / RA2) * S.lda) * 8) : OOB;
    }
    if (BKD) {
      cB = (unsigned)((NT / 8) * S.ldb * 8);
      oB = (unsigned)((2 * (tid % 8) + (int64_t)(tid / 8) * S.ldb) * 8);
    } else {
      cB = (unsigned)((NT / RB2) * S.ldb * 8);
      oB = n0 + 2 * (tid % RB2) < N ? (unsigned)((2 * (tid % RB2) + (int64_t)(tid / RB2) * S.ldb) * 8) : OOB;
    }
    K = S.K;
  };

  dx2 ra[CHA], rb[CHB];
  auto load = [&](int k0) {
#pragma unroll
    for (int c = 0; c < CHA; ++c) {
      bool ok;
      if (AK) ok = (m0 + tid / 8 + c * (NT / 8) < M) & (k0 + 2 * (tid % 8) < K);
      else    ok = (oA != OOB) & (k0 + tid / RA2 + c * (NT / RA2) < K);
      const unsigned o = ok ? oA + (unsigned)c * cA + (unsigned)k0 * stepA : OOB;
      ra[c] = __builtin_bit_cast(dx2, __builtin_amdgcn_raw_buffer_load_b128(rA, (int)o, 0, 0));
    }
#pragma unroll
    for (int c = 0; c < CHB; ++c) {
      bool ok;
      if (BKD) ok = (n0 + tid / 8 + c * (NT / 8) < N) & (k0 + 2 * (tid % 8) < K);
      else     ok = (oB != OOB) & (k0 + tid / RB2 + c * (NT / RB2) < K);
      const unsigned o = ok ? oB + (unsigned)c * cB + (unsigned)k0 * stepB : OOB;
      rb[c] = __builtin_bit_cast(dx2, __builtin_amdgcn_raw_buffer_load_b128(rB, (int)o, 0, 0));
    }
  };
  auto store = [&](int buf) {
    if (!TA) store_fastdir<BM, LDA_S, BK>(As[buf], ra);
    else     store_kdir<BM, LDA_S, BK>(As[buf], ra);
    if (!TB) store_kdir<BN, LDB_S, BK>(Bs[buf], rb);
    else     store_fastdir<BN, LDB_S, BK>(Bs[buf], rb);
  };

  auto read_frag = [&](int bi, int kk, double (&a)[FM], double (&b)[FN]) {
    const int kr = kk * 4 + lk;
    const int sa = TA ? kswz(kr) : 0;   // K-contiguous operands were stored swizzled
    const int sb = TB ? 0 : kswz(kr);
#pragma unroll
    for (int i = 0; i < FM; ++i) a[i] = As[bi][kr][(wm0 + i * 16 + li) ^ sa];
#pragma unroll
    for (int j = 0; j < FN; ++j) b[j] = Bs[bi][kr][(wn0 + j * 16 + li) ^ sb];
  };
  double fa[2][FM], fb[2][FN];

  int s = 0, k0 = 0;
  if (nsteps > 0) {
    set_seg(0);
    load(0);
    store(0);
  }
  __syncthreads();
  if (ROT && nsteps > 0) read_frag(0, 0, fa[0], fb[0]);
  int buf = 0;
  for (int step = 0; step < nsteps; ++step) {
    k0 += BK;
    while (k0 >= K && s + 1 < P.nseg) {
      ++s;
      k0 = 0;
      set_seg(s);
    }
    const bool has_next = step + 1 < nsteps;
    if (has_next) load(k0);
    if constexpr (!ROT) {
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        double a[FM], b[FN];
        read_frag(buf, kk, a, b);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
      }
      if (has_next) store(buf ^ 1);
      __syncthreads();
    } else {
      // fragments of k slice kk + 1 are read while slice kk multiplies; the
      // step barrier sits before the last slice's MFMAs, which then cover the
      // latency of the next step's first fragment reads
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        if (kk + 1 < BK / 4) {
          read_frag(buf, kk + 1, fa[(kk + 1) & 1], fb[(kk + 1) & 1]);
        } else {
          // (storing the next tile earlier in the step measured slower: its
          // global loads, issued at the step start, are often still in flight)
          if (has_next) store(buf ^ 1);
          __builtin_amdgcn_sched_barrier(0);
          __syncthreads();
          __builtin_amdgcn_sched_barrier(0);
          if (has_next) read_frag(buf ^ 1, 0, fa[0], fb[0]);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[kk & 1][j], fa[kk & 1][i], acc[i][j], 0, 0, 0);
      }
    }
    buf ^= 1;
  }

  // epilogue: as in gemm_f64_kernel
  epilogue_c<false, FM, FN, dx4, 1>(acc, resolve(src, P.C), P.ldc, M, N, m0 + wm0, n0 + wn0, li, lk, P.alpha, P.beta);
  }  // tiles
}

// Host check of the branch-free kernel's conditions for one problem.
template <int BM, int BN>
bool buf_ok_prob(const ProbSpec& ps, GemmShape sh) {
  for (int s = 0; s < ps.nseg; ++s) {
    const SegSpec& g = ps.seg[s];
    if (g.K <= 0) continue;
    if ((uintptr_t)g.A % 16 || (uintptr_t)g.B % 16 || g.lda % 2 || g.ldb % 2 || g.K % 2) return false;
    if (!sh.ta && ps.M % 2) return false;
    if (sh.tb && ps.N % 2) return false;
    // largest in-tile byte offset (plus the 16-byte chunk) below 2^31
    const int64_t offA = sh.ta ? (g.K + (int64_t)(BM - 1) * g.lda) * 8 : (BM + (int64_t)(g.K - 1) * g.lda) * 8;
    const int64_t offB = !sh.tb ? (g.K + (int64_t)(BN - 1) * g.ldb) * 8 : (BN + (int64_t)(g.K - 1) * g.ldb) * 8;
    if (offA + 16 >= (1ll << 31) || offB + 16 >= (1ll << 31)) return false;
  }
  return true;
}

// Load path of the FP64 kernels, read once from MK_F64_VARIANT ("rot", "buf"
// or "branch"; default "rot"): lets benchmarks and tests pin one of them.
enum F64Variant { F64_BRANCH = 0, F64_BUF = 1, F64_ROT = 2 };
int f64_variant() {
  static const int v = [] {
    const char* e = getenv("MK_F64_VARIANT");
    if (e && !strcmp(e, "branch")) return (int)F64_BRANCH;
    if (e && !strcmp(e, "buf")) return (int)F64_BUF;
    return (int)F64_ROT;
  }();
  return v;
}

template <int BM, int BN, class Src, bool ROT>
void launch4_buf(GemmShape sh, dim3 grid, dim3 block, hipStream_t stream, const Src& b) {
  if (!sh.ta && !sh.tb) hipLaunchKernelGGL((gemm_f64_buf_kernel<BM, BN, false, false, Src, ROT>), grid, block, 0, stream, b);
  else if (sh.ta && !sh.tb) hipLaunchKernelGGL((gemm_f64_buf_kernel<BM, BN, true, false, Src, ROT>), grid, block, 0, stream, b);
  else if (!sh.ta && sh.tb) hipLaunchKernelGGL((gemm_f64_buf_kernel<BM, BN, false, true, Src, ROT>), grid, block, 0, stream, b);
  else hipLaunchKernelGGL((gemm_f64_buf_kernel<BM, BN, true, true, Src, ROT>), grid, block, 0, stream, b);
}

// K depth of a stage, read once from MK_F64_BK (16 or 32; default 16).  BK = 16
// halves the LDS footprint of a 128 x 128 workgroup (74 KB instead of 147 KB),
// so two workgroups (two waves per SIMD) share a CU: 24% faster at
// M = N = 4096, K = 8192 than BK = 32 with one wave per SIMD.
int f64_bk() {
  static const int bk = [] {
    const char* e = getenv("MK_F64_BK");
    return (e && atoi(e) == 32) ? 32 : 16;
  }();
  return bk;
}

template <int BM, int BN, class Src, int BK>
void launch4(GemmShape sh, dim3 grid, dim3 block, hipStream_t stream, const Src& b) {
  if (!sh.ta && !sh.tb) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false, false, Src, BK>), grid, block, 0, stream, b);
  else if (sh.ta && !sh.tb) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true, false, Src, BK>), grid, block, 0, stream, b);
  else if (!sh.ta && sh.tb) hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false, true, Src, BK>), grid, block, 0, stream, b);
  else hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true, true, Src, BK>), grid, block, 0, stream, b);
}

template <int BM, int BN>
mk_status launch_cfg(const ProbSpec* probs, int nprob, GemmShape sh, hipStream_t stream) {
  GemmBatch b{};
  int total = 0;
  for (int i = 0; i < nprob; ++i) {
    const ProbSpec& ps = probs[i];
    GemmProb& P = b.p[i];
    P.M = (int32_t)ps.M;
    P.N = (int32_t)ps.N;
    P.nseg = 0;
    for (int s = 0; s < ps.nseg; ++s) {
      if (ps.seg[s].K <= 0) continue;
      GemmSeg& g = P.seg[P.nseg++];
      g.A = ps.seg[s].A;
      g.B = ps.seg[s].B;
      g.Alo = g.Blo = nullptr;
      g.lda = ps.seg[s].lda;
      g.ldb = ps.seg[s].ldb;
      g.K = (int32_t)ps.seg[s].K;
      const bool al = ((uintptr_t)g.A % 16 == 0) && ((uintptr_t)g.B % 16 == 0) &&
                      (g.lda % 2 == 0) && (g.ldb % 2 == 0);
      g.vec = al ? 1 : 0;
    }
    P.tiles_m = (int32_t)cdiv(ps.M, BM);
    P.tiles_n = (int32_t)cdiv(ps.N, BN);
    P.tile_start = total;
    total += P.tiles_m * P.tiles_n;
    P.C = ps.C;
    P.ldc = ps.ldc;
    P.alpha = ps.alpha;
    P.beta = ps.beta;
  }
  b.nprob = nprob;
  b.total_tiles = total;
  if (total == 0) return MK_OK;
  const int var = f64_variant();
  bool buf = var != F64_BRANCH && f64_bk() == 16;
  for (int i = 0; i < nprob && buf; ++i) buf = buf_ok_prob<BM, BN>(probs[i], sh);
  dim3 grid(total), block(NT);
  ProfScope prof("gemm_f64", stream);
  // the rotated schedule holds two fragment sets; with a K-contiguous A the
  // extra registers would spill, so those launches keep the plain schedule
  if (buf && var == F64_ROT && !sh.ta) launch4_buf<BM, BN, GemmBatch, true>(sh, grid, block, stream, b);
  else if (buf) launch4_buf<BM, BN, GemmBatch, false>(sh, grid, block, stream, b);
  else if (f64_bk() == 16) launch4<BM, BN, GemmBatch, 16>(sh, grid, block, stream, b);
  else launch4<BM, BN, GemmBatch, BK32>(sh, grid, block, stream, b);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

template <int BM, int BN>
mk_status launch_group(const GemmGroup& g, GemmShape sh, int grid, hipStream_t stream) {
  if (g.total_tiles == 0) return MK_OK;
  dim3 gr(grid), block(NT);
  ProfScope prof("gemm_f64", stream);
  const int var = f64_variant();
  const bool buf = g.f64_buf && var != F64_BRANCH && f64_bk() == 16;
  if (buf && var == F64_ROT && !sh.ta) launch4_buf<BM, BN, GemmGroup, true>(sh, gr, block, stream, g);
  else if (buf) launch4_buf<BM, BN, GemmGroup, false>(sh, gr, block, stream, g);
  else if (f64_bk() == 16) launch4<BM, BN, GemmGroup, 16>(sh, gr, block, stream, g);
  else launch4<BM, BN, GemmGroup, BK32>(sh, gr, block, stream, g);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

template <int BM, int BN>
int occupancy() {
  static const int occ = [] {
    int n = 0;
    const hipError_t e =
        f64_bk() == 16
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_f64_kernel<BM, BN, false, false, GemmGroup, 16>, NT, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_f64_kernel<BM, BN, false, false, GemmGroup, BK32>, NT,
                                                           0);
    if (e != hipSuccess) n = 1;
    return n > 0 ? n : 1;
  }();
  return occ;
}

}  // namespace

int f64_group_grid(int kind, int total_tiles, int cus) {
  const int occ = kind == TK_128 ? occupancy<128, 128>() : occupancy<64, 64>();
  return (int)std::min<int64_t>(total_tiles, (int64_t)occ * cus);
}

bool f64_buf_ok(const ProbSpec* probs, int n, GemmShape sh, int kind) {
  for (int i = 0; i < n; ++i)
    if (!(kind == TK_128 ? buf_ok_prob<128, 128>(probs[i], sh) : buf_ok_prob<64, 64>(probs[i], sh))) return false;
  return true;
}

mk_status gemm_f64_group(const GemmGroup& g, int kind, GemmShape sh, int grid, hipStream_t st) {
  if (kind == TK_128) return launch_group<128, 128>(g, sh, grid, st);
  MK_REQUIRE(kind == TK_64, MK_BAD_ARG, "gemm_f64_group: tile kind %d", kind);
  return launch_group<64, 64>(g, sh, grid, st);
}

mk_status gemm_f64(const ProbSpec* probs, int nprob, GemmShape shape, hipStream_t stream) {
  MK_REQUIRE(nprob >= 0 && nprob <= MAX_PROB, MK_BAD_ARG, "gemm_f64: %d problems (max %d)", nprob, MAX_PROB);
  int64_t tiles128 = 0;
  for (int i = 0; i < nprob; ++i) {
    MK_REQUIRE(probs[i].nseg <= MAX_SEG, MK_BAD_ARG, "gemm_f64: too many segments");
    MK_REQUIRE(probs[i].M >= 0 && probs[i].N >= 0 && probs[i].M < (1ll << 31) && probs[i].N < (1ll << 31),
               MK_BAD_SHAPE, "gemm_f64: bad M/N");
    for (int s = 0; s < probs[i].nseg; ++s)
      MK_REQUIRE(probs[i].seg[s].K >= 0 && probs[i].seg[s].K < (1ll << 31), MK_BAD_SHAPE, "gemm_f64: bad K");
    tiles128 += cdiv(probs[i].M, 128) * cdiv(probs[i].N, 128);
  }
  // 128x128 tiles once they fill the 256 CUs, else 64x64 for more parallelism
  if (tiles128 >= 256) return launch_cfg<128, 128>(probs, nprob, shape, stream);
  return launch_cfg<64, 64>(probs, nprob, shape, stream);
}

}  // namespace mk
