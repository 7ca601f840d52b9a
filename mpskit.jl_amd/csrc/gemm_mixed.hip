// Mixed-precision MFMA GEMMs for CDNA4 (gfx950) with float64 output.
//
//  BF16X3: each float64 operand is pre-split into bf16 hi + lo; the kernel
//          issues hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 and
//          accumulates in fp32 (relative error ~2^-16 per product).
//  F32   : operands pre-rounded to fp32, v_mfma_f32_16x16x4_f32.
//
// Operands are K-contiguous: A(i,k) = A[i*lda + k], B(k,j) = B[j*ldb + k], with
// 16-byte aligned rows whose length lda/ldb is padded (with zeros) to a
// multiple of 8 elements.  One LDS row per tile row holds 64 bytes of K
// (32 bf16 or 16 fp32); lane (row r = lane%16, group g = lane/16) reads the 16
// bytes at offset 16*g, which is the bf16 MFMA k-slot layout and, for fp32,
// the k-permutation "group g owns k = 4g..4g+3 across the 4 sub-steps" shared
// by A and B.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "gemm_dev.h"
#include "prof.h"

namespace mk {
namespace {

typedef float fx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bfx8 __attribute__((ext_vector_type(8)));
typedef unsigned int ux4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int ROWB = 64;          // bytes of K per LDS row per step
constexpr int ROWP = ROWB + 16;   // padded LDS row (80 B: conflict-free 16-B reads)

enum Mode { M_BF16X3 = 0, M_F32 = 1 };

// Kernel family for large problems, read once from MK_GEMM_VARIANT ("big2",
// "big", "pipe", "tile"; default "big2"): lets benchmarks and tests pin one family.
enum Variant { V_PIPE = 0, V_BIG = 1, V_TILE = 2, V_BIG2 = 3 };
int variant() {
  static const int v = [] {
    const char* e = getenv("MK_GEMM_VARIANT");
    if (e && !strcmp(e, "pipe")) return (int)V_PIPE;
    if (e && !strcmp(e, "tile")) return (int)V_TILE;
    if (e && !strcmp(e, "big")) return (int)V_BIG;
    return (int)V_BIG2;
  }();
  return v;
}

// Epilogue of an FM x FN grid of 16x16 fp32 accumulator blocks.  The MFMAs run
// with the B fragment as first operand, so block (i, j) is C^T and item r of
// lane (li, lg) is C[row0 + 16 i + li][col0 + 16 j + 4 lg + r]: the 16 lanes of
// a quarter-wave store 16 consecutive rows (128 B) of one column.
template <int FM, int FN>
__device__ __forceinline__ void store_tile(const fx4 (&acc)[FM][FN], double* __restrict__ C, int64_t ldc, int M,
                                           int N, int row0, int col0, int li, int lg, double alpha, double beta) {
  epilogue_c<true>(acc, C, ldc, M, N, row0, col0, li, lg, alpha, beta);
}

// Load a (R rows x 64 B) slab of a K-contiguous operand into registers.
// elem_bytes = 2 (bf16) or 4 (fp32); Klim = padded row length in elements.
template <int R>
__device__ __forceinline__ void load_slab(const char* __restrict__ base, int64_t ld_bytes, int r0,
                                          int64_t kbyte0, int Rlim, int64_t Kbytes,
                                          ux4 (&reg)[R * ROWB / 16 / NT]) {
  constexpr int CH = R * ROWB / 16 / NT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * NT;
    const int r = id / (ROWB / 16);
    const int q = id % (ROWB / 16);
    const int gr = r0 + r;
    const int64_t kb = kbyte0 + q * 16;
    ux4 v = {0u, 0u, 0u, 0u};
    if (gr < Rlim && kb < Kbytes) v = *gptr(reinterpret_cast<const ux4*>(base + (int64_t)gr * ld_bytes + kb));
    reg[c] = v;
  }
}

template <int R>
__device__ __forceinline__ void store_slab(char* S, const ux4 (&reg)[R * ROWB / 16 / NT]) {
  constexpr int CH = R * ROWB / 16 / NT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * NT;
    const int r = id / (ROWB / 16);
    const int q = id % (ROWB / 16);
    *reinterpret_cast<ux4*>(S + r * ROWP + q * 16) = reg[c];
  }
}

template <int BM, int BN, int MODE, class Src>
__global__ __launch_bounds__(NT) void gemm_mixed_kernel(Src src) {
  constexpr int NPART = MODE == M_BF16X3 ? 2 : 1;   // hi/lo arrays per operand
  constexpr int EB = MODE == M_BF16X3 ? 2 : 4;      // element bytes
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int CHA = BM * ROWB / 16 / NT;
  constexpr int CHB = BN * ROWB / 16 / NT;

  __shared__ __attribute__((aligned(16))) char As[2][NPART][BM * ROWP];
  __shared__ __attribute__((aligned(16))) char Bs[2][NPART][BN * ROWP];

  for (int t = src_first(src); t < src.total_tiles; t += src_stride(src)) {
  const GemmProb& P = src_prob(src, t);
  int tm, tn;
  tile_coords(P, t - P.tile_start, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  fx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};

  constexpr int KSTEP = ROWB / EB;  // elements of K per step
  int nsteps = 0;
  for (int s = 0; s < P.nseg; ++s) nsteps += (P.seg[s].K + KSTEP - 1) / KSTEP;

  ux4 ra[NPART][CHA], rb[NPART][CHB];
  auto load = [&](int s, int k0) {
    const GemmSeg& S = P.seg[s];
    const int64_t ldab = S.lda * EB, ldbb = S.ldb * EB;
    const int64_t kb0 = (int64_t)k0 * EB, Kb = (int64_t)S.K * EB;
    load_slab<BM>(static_cast<const char*>(resolve(src, S.A)), ldab, m0, kb0, P.M, Kb, ra[0]);
    load_slab<BN>(static_cast<const char*>(resolve(src, S.B)), ldbb, n0, kb0, P.N, Kb, rb[0]);
    if (NPART == 2) {
      load_slab<BM>(static_cast<const char*>(resolve(src, S.Alo)), ldab, m0, kb0, P.M, Kb, ra[NPART - 1]);
      load_slab<BN>(static_cast<const char*>(resolve(src, S.Blo)), ldbb, n0, kb0, P.N, Kb, rb[NPART - 1]);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NPART; ++p) {
      store_slab<BM>(As[buf][p], ra[p]);
      store_slab<BN>(Bs[buf][p], rb[p]);
    }
  };

  int s = 0, k0 = 0;
  if (nsteps > 0) {
    load(0, 0);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int step = 0; step < nsteps; ++step) {
    int ns = s, nk0 = k0 + KSTEP;
    if (nk0 >= P.seg[s].K) { ns = s + 1; nk0 = 0; }
    const bool has_next = step + 1 < nsteps;
    if (has_next) load(ns, nk0);

    ux4 fa[NPART][FM], fb[NPART][FN];
#pragma unroll
    for (int p = 0; p < NPART; ++p) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
        fa[p][i] = *reinterpret_cast<const ux4*>(As[buf][p] + (wm0 + i * 16 + li) * ROWP + lg * 16);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        fb[p][j] = *reinterpret_cast<const ux4*>(Bs[buf][p] + (wn0 + j * 16 + li) * ROWP + lg * 16);
    }
    if constexpr (MODE == M_BF16X3) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const bfx8 ah = __builtin_bit_cast(bfx8, fa[0][i]);
          const bfx8 al = __builtin_bit_cast(bfx8, fa[1][i]);
          const bfx8 bh = __builtin_bit_cast(bfx8, fb[0][j]);
          const bfx8 bl = __builtin_bit_cast(bfx8, fb[1][j]);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah, acc[i][j], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int sub = 0; sub < 4; ++sub)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            // extract the element to a scalar first: bit-casting a vector-element
            // lvalue reads element 0
            const unsigned int ua = fa[0][i][sub], ub = fb[0][j][sub];
            const float a = __uint_as_float(ua);
            const float b = __uint_as_float(ub);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, acc[i][j], 0, 0, 0);
          }
    }
    if (has_next) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
    s = ns;
    k0 = nk0;
  }

  store_tile<FM, FN>(acc, resolve(src, P.C), P.ldc, P.M, P.N, m0 + wm0, n0 + wn0, li, lg, P.alpha, P.beta);
  }  // tiles
}

// ---------------------------------------------------------------------------
// Large-tile variant: 256 x 256 output tile, 512 threads = 8 waves (2 x 4), each
// wave 128 x 64.  Halves the operand re-reads of the 128 x 128 kernel and
// doubles the MFMA work per LDS fragment.  LDS rows hold 64 B of K without
// padding; 16-B chunks are XOR-swizzled by (row >> 2) & 3 so that the 16 rows
// read by one 16-lane group hit distinct banks.  Double-buffered: 128 KB (bf16).
// ---------------------------------------------------------------------------

constexpr int BIG = 256;
constexpr int BNT = 512;

__device__ __forceinline__ int swz(int r, int q) { return r * ROWB + ((q ^ ((r >> 2) & 3)) << 4); }

template <int R>
__device__ __forceinline__ void load_slab_big(const char* __restrict__ base, int64_t ld_bytes, int r0,
                                              int64_t kbyte0, int Rlim, int64_t Kbytes,
                                              ux4 (&reg)[R * ROWB / 16 / BNT]) {
  constexpr int CH = R * ROWB / 16 / BNT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * BNT;
    const int r = id >> 2, q = id & 3;
    const int gr = r0 + r;
    const int64_t kb = kbyte0 + q * 16;
    ux4 v = {0u, 0u, 0u, 0u};
    if (gr < Rlim && kb < Kbytes) v = *gptr(reinterpret_cast<const ux4*>(base + (int64_t)gr * ld_bytes + kb));
    reg[c] = v;
  }
}

template <int R>
__device__ __forceinline__ void store_slab_big(char* S, const ux4 (&reg)[R * ROWB / 16 / BNT]) {
  constexpr int CH = R * ROWB / 16 / BNT;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int id = threadIdx.x + c * BNT;
    const int r = id >> 2, q = id & 3;
    *reinterpret_cast<ux4*>(S + swz(r, q)) = reg[c];
  }
}

template <int MODE, class Src>
__global__ __launch_bounds__(BNT) void gemm_mixed_big_kernel(Src src) {
  constexpr int NPART = MODE == M_BF16X3 ? 2 : 1;
  constexpr int EB = MODE == M_BF16X3 ? 2 : 4;
  constexpr int WM = 128, WN = 64;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int CH = BIG * ROWB / 16 / BNT;   // 16-B chunks per thread per operand part

  __shared__ __attribute__((aligned(16))) char As[2][NPART][BIG * ROWB];
  __shared__ __attribute__((aligned(16))) char Bs[2][NPART][BIG * ROWB];

  for (int t = src_first(src); t < src.total_tiles; t += src_stride(src)) {
  const GemmProb& P = src_prob(src, t);
  int tm, tn;
  tile_coords(P, t - P.tile_start, tm, tn);
  const int m0 = tm * BIG, n0 = tn * BIG;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  fx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};

  constexpr int KSTEP = ROWB / EB;
  int nsteps = 0;
  for (int s = 0; s < P.nseg; ++s) nsteps += (P.seg[s].K + KSTEP - 1) / KSTEP;

  ux4 ra[NPART][CH], rb[NPART][CH];
  auto load = [&](int s, int k0) {
    const GemmSeg& S = P.seg[s];
    const int64_t ldab = S.lda * EB, ldbb = S.ldb * EB;
    const int64_t kb0 = (int64_t)k0 * EB, Kb = (int64_t)S.K * EB;
    load_slab_big<BIG>(static_cast<const char*>(resolve(src, S.A)), ldab, m0, kb0, P.M, Kb, ra[0]);
    load_slab_big<BIG>(static_cast<const char*>(resolve(src, S.B)), ldbb, n0, kb0, P.N, Kb, rb[0]);
    if (NPART == 2) {
      load_slab_big<BIG>(static_cast<const char*>(resolve(src, S.Alo)), ldab, m0, kb0, P.M, Kb, ra[NPART - 1]);
      load_slab_big<BIG>(static_cast<const char*>(resolve(src, S.Blo)), ldbb, n0, kb0, P.N, Kb, rb[NPART - 1]);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NPART; ++p) {
      store_slab_big<BIG>(As[buf][p], ra[p]);
      store_slab_big<BIG>(Bs[buf][p], rb[p]);
    }
  };

  int s = 0, k0 = 0;
  if (nsteps > 0) {
    load(0, 0);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  const int aoff = swz(wm0 + li, lg);   // same swizzle for every 16-row block
  const int boff = swz(wn0 + li, lg);
  for (int step = 0; step < nsteps; ++step) {
    int ns = s, nk0 = k0 + KSTEP;
    if (nk0 >= P.seg[s].K) { ns = s + 1; nk0 = 0; }
    const bool has_next = step + 1 < nsteps;
    if (has_next) load(ns, nk0);

    ux4 fb[NPART][FN];
#pragma unroll
    for (int p = 0; p < NPART; ++p)
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[p][j] = *reinterpret_cast<const ux4*>(Bs[buf][p] + boff + j * 16 * ROWB);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      ux4 fa[NPART];
#pragma unroll
      for (int p = 0; p < NPART; ++p) fa[p] = *reinterpret_cast<const ux4*>(As[buf][p] + aoff + i * 16 * ROWB);
      if constexpr (MODE == M_BF16X3) {
        const bfx8 ah = __builtin_bit_cast(bfx8, fa[0]);
        const bfx8 al = __builtin_bit_cast(bfx8, fa[1]);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const bfx8 bh = __builtin_bit_cast(bfx8, fb[0][j]);
          const bfx8 bl = __builtin_bit_cast(bfx8, fb[1][j]);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah, acc[i][j], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int sub = 0; sub < 4; ++sub) {
          const unsigned int ua = fa[0][sub];
          const float a = __uint_as_float(ua);
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            const unsigned int ub = fb[0][j][sub];
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ub), a, acc[i][j], 0, 0, 0);
          }
        }
      }
    }
    if (has_next) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
    s = ns;
    k0 = nk0;
  }

  store_tile<FM, FN>(acc, resolve(src, P.C), P.ldc, P.M, P.N, m0 + wm0, n0 + wn0, li, lg, P.alpha, P.beta);
  }  // tiles
}

// Raw (stride 0) buffer resource over p with num_records = 2^31 - 1 bytes and
// the gfx9 data-format word; offsets at or past num_records load zeros.
constexpr unsigned OOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff, 0x00020000);
}

// Same tile, waves and LDS layout as gemm_mixed_big_kernel, with a lighter
// load path: the segment's operand descriptors are read once per segment (not
// once per K step), each operand is addressed through a buffer resource based
// at the tile's first row, and rows or K chunks outside the operand get a
// buffer offset past num_records (the hardware returns zeros), so a K step
// issues its loads without scalar loads or branches.  Needs 256 rows of every
// operand within 2 GiB of the row base (checked on the host).
template <int MODE, class Src>
__global__ __launch_bounds__(BNT) void gemm_mixed_big2_kernel(Src src) {
  constexpr int NPART = MODE == M_BF16X3 ? 2 : 1;
  constexpr int EB = MODE == M_BF16X3 ? 2 : 4;
  constexpr int WM = 128, WN = 64;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int CH = BIG * ROWB / 16 / BNT;   // 16-B chunks per thread per operand part
  constexpr int KSTEP = ROWB / EB;

  __shared__ __attribute__((aligned(16))) char As[2][NPART][BIG * ROWB];
  __shared__ __attribute__((aligned(16))) char Bs[2][NPART][BIG * ROWB];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;
  const int cq = threadIdx.x & 3;             // 16-B chunk of the 64-B row slab
  int crow[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) crow[c] = (threadIdx.x + c * BNT) >> 2;
  const int aoff = swz(wm0 + li, lg);
  const int boff = swz(wn0 + li, lg);

  for (int t = src_first(src); t < src.total_tiles; t += src_stride(src)) {
  const GemmProb& P = src_prob(src, t);
  int tm, tn;
  tile_coords(P, t - P.tile_start, tm, tn);
  const int m0 = tm * BIG, n0 = tn * BIG;
  const int nseg = P.nseg;

  fx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};

  // per-segment load state.  A row-layout operand is addressed by row; a
  // packed one (256-row panels of contiguous 16 KB K steps, every row present)
  // by its panel, with the K step at 16 KB per 64 bytes of K.
  __amdgpu_buffer_rsrc_t rA[NPART], rB[NPART];
  unsigned vA[CH], vB[CH];
  unsigned kmulA = 1, kmulB = 1;   // bytes of offset per byte of K
  int Kb = 0, kend = 0;
  auto set_seg = [&](int s) {
    const GemmSeg& S = P.seg[s];
    const bool pa = (S.vec & 2) != 0, pb = (S.vec & 4) != 0;
    const int64_t panel = (int64_t)S.K * EB * 256;
    const int64_t ldab = S.lda * EB, ldbb = S.ldb * EB;
    const int64_t offa = pa ? (int64_t)(m0 / 256) * panel : (int64_t)m0 * ldab;
    const int64_t offb = pb ? (int64_t)(n0 / 256) * panel : (int64_t)n0 * ldbb;
    rA[0] = raw_rsrc(static_cast<const char*>(resolve(src, S.A)) + offa);
    rB[0] = raw_rsrc(static_cast<const char*>(resolve(src, S.B)) + offb);
    if constexpr (NPART == 2) {
      rA[1] = raw_rsrc(static_cast<const char*>(resolve(src, S.Alo)) + offa);
      rB[1] = raw_rsrc(static_cast<const char*>(resolve(src, S.Blo)) + offb);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      vA[c] = pa ? (unsigned)((threadIdx.x + c * BNT) * 16)
                 : (m0 + crow[c] < P.M ? (unsigned)(crow[c] * ldab + cq * 16) : OOB);
      vB[c] = pb ? (unsigned)((threadIdx.x + c * BNT) * 16)
                 : (n0 + crow[c] < P.N ? (unsigned)(crow[c] * ldbb + cq * 16) : OOB);
    }
    kmulA = pa ? 256u : 1u;
    kmulB = pb ? 256u : 1u;
    Kb = S.K * EB;
    kend = S.K;
  };

  int nsteps = 0;
  for (int s = 0; s < nseg; ++s) nsteps += (P.seg[s].K + KSTEP - 1) / KSTEP;

  ux4 ra[NPART][CH], rb[NPART][CH];
  auto load = [&](int k0) {
    const int kb0 = k0 * EB;
    const bool kin = kb0 + cq * 16 < Kb;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const unsigned oa = kin ? vA[c] + (unsigned)kb0 * kmulA : OOB;
      const unsigned ob = kin ? vB[c] + (unsigned)kb0 * kmulB : OOB;
#pragma unroll
      for (int p = 0; p < NPART; ++p) {
        ra[p][c] = __builtin_amdgcn_raw_buffer_load_b128(rA[p], (int)oa, 0, 0);
        rb[p][c] = __builtin_amdgcn_raw_buffer_load_b128(rB[p], (int)ob, 0, 0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NPART; ++p)
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        *reinterpret_cast<ux4*>(As[buf][p] + swz(crow[c], cq)) = ra[p][c];
        *reinterpret_cast<ux4*>(Bs[buf][p] + swz(crow[c], cq)) = rb[p][c];
      }
  };

  int s = 0, k0 = 0;
  if (nsteps > 0) {
    set_seg(0);
    load(0);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int step = 0; step < nsteps; ++step) {
    k0 += KSTEP;
    while (k0 >= kend && s + 1 < nseg) {
      ++s;
      k0 = 0;
      set_seg(s);
    }
    const bool has_next = step + 1 < nsteps;
    if (has_next) load(k0);

    ux4 fb[NPART][FN];
#pragma unroll
    for (int p = 0; p < NPART; ++p)
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[p][j] = *reinterpret_cast<const ux4*>(Bs[buf][p] + boff + j * 16 * ROWB);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      ux4 fa[NPART];
#pragma unroll
      for (int p = 0; p < NPART; ++p) fa[p] = *reinterpret_cast<const ux4*>(As[buf][p] + aoff + i * 16 * ROWB);
      if constexpr (MODE == M_BF16X3) {
        const bfx8 ah = __builtin_bit_cast(bfx8, fa[0]);
        const bfx8 al = __builtin_bit_cast(bfx8, fa[1]);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const bfx8 bh = __builtin_bit_cast(bfx8, fb[0][j]);
          const bfx8 bl = __builtin_bit_cast(bfx8, fb[1][j]);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah, acc[i][j], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int sub = 0; sub < 4; ++sub) {
          const unsigned int ua = fa[0][sub];
          const float a = __uint_as_float(ua);
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            const unsigned int ub = fb[0][j][sub];
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ub), a, acc[i][j], 0, 0, 0);
          }
        }
      }
    }
    if (has_next) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  store_tile<FM, FN>(acc, resolve(src, P.C), P.ldc, P.M, P.N, m0 + wm0, n0 + wn0, li, lg, P.alpha, P.beta);
  }  // tiles
}

// ---------------------------------------------------------------------------
// Pipelined variant for large problems: 256 x 128 output tile, 256 threads =
// 4 waves (2 x 2) of 128 x 64, three LDS stages.  Operand slabs go global ->
// LDS through buffer_load ... lds (no staging registers), so the loads of step
// s + 2 are in flight while step s computes; a single-stage register prefetch
// covers less than one L2-miss latency at this MFMA rate.  Lane l of a wave's
// 16-byte LDS load writes LDS bytes [16 l, 16 l + 16) after the wave base, so
// the swizzle is applied on the global side: LDS chunk qp of row r receives
// logical chunk qp ^ ((r >> 2) & 3).  Out-of-range rows and K chunks get a
// buffer offset past num_records, which the hardware turns into zeros.
// Three stages of hi + lo bf16 slabs: 3 x (32 + 16) KB = 144 KB of LDS.
// ---------------------------------------------------------------------------

constexpr int PBM = 256, PBN = 128, PNT = 256, NSTAGE = 3;

typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt on vmcnt only (gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14])
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Workgroup barrier without the release fence of __syncthreads(): that fence
// would wait for every outstanding LDS-DMA load (vmcnt(0)) and serialise the
// pipeline.  Each wave waits for its own loads of the stage it is about to read
// (wait_vm) before the barrier, and every LDS read of the previous stage has
// been consumed by an MFMA before it, so the barrier alone orders the stages.
// The empty asm statements keep the compiler from moving LDS accesses across.
__device__ __forceinline__ void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// One R x 64 B slab of a K-contiguous operand into LDS stage memory S.
template <int R>
__device__ __forceinline__ void dma_slab(__amdgpu_buffer_rsrc_t rs, char* S, int ld_bytes, int r0, int Rlim,
                                         int kb0, int Kbytes, int wave) {
  constexpr int CH = R * (ROWB / 16) / PNT;
  const int t = threadIdx.x;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int idx = c * PNT + t;
    const int r = idx >> 2, qp = idx & 3;
    const int q = qp ^ ((r >> 2) & 3);
    const int gr = r0 + r, kb = kb0 + q * 16;
    const unsigned vo = (gr < Rlim && kb < Kbytes) ? (unsigned)(gr * ld_bytes + kb) : OOB;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void*)(S + (c * PNT + wave * 64) * 16), 16, vo, 0, 0, 0);
  }
}

template <int MODE>
__global__ __launch_bounds__(PNT) void gemm_mixed_pipe_kernel(GemmBatch batch) {
  constexpr int NPART = MODE == M_BF16X3 ? 2 : 1;
  constexpr int EB = MODE == M_BF16X3 ? 2 : 4;
  constexpr int WM = 128, WN = 64;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int KSTEP = ROWB / EB;
  constexpr int ABYTES = PBM * ROWB, BBYTES = PBN * ROWB;
  constexpr int SBYTES = NPART * (ABYTES + BBYTES);
  // loads issued per thread per stage (each counts once in vmcnt)
  constexpr int LPS = NPART * (PBM + PBN) * (ROWB / 16) / PNT;

  __shared__ __attribute__((aligned(16))) char smem[NSTAGE * SBYTES];

  const int t = src_first(batch);
  const GemmProb& P = src_prob(batch, t);
  int tm, tn;
  tile_coords(P, t - P.tile_start, tm, tn);
  const int m0 = tm * PBM, n0 = tn * PBN;

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm0 = (wave & 1) * WM;
  const int wn0 = (wave >> 1) * WN;
  const int li = lane & 15;
  const int lg = lane >> 4;

  fx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};

  int nsteps = 0;
  for (int s = 0; s < P.nseg; ++s) nsteps += (P.seg[s].K + KSTEP - 1) / KSTEP;

  // load cursor (segment, k) of the next stage to fetch
  int ls = 0, lk = 0;
  auto fetch = [&](int stage) {
    const GemmSeg& S = P.seg[ls];
    const int ldab = (int)S.lda * EB, ldbb = (int)S.ldb * EB;
    const int kb0 = lk * EB, Kb = S.K * EB;
    char* base = smem + stage * SBYTES;
    dma_slab<PBM>(raw_rsrc(S.A), base, ldab, m0, P.M, kb0, Kb, wave);
    dma_slab<PBN>(raw_rsrc(S.B), base + NPART * ABYTES, ldbb, n0, P.N, kb0, Kb, wave);
    if constexpr (NPART == 2) {
      dma_slab<PBM>(raw_rsrc(S.Alo), base + ABYTES, ldab, m0, P.M, kb0, Kb, wave);
      dma_slab<PBN>(raw_rsrc(S.Blo), base + 2 * ABYTES + BBYTES, ldbb, n0, P.N, kb0, Kb, wave);
    }
    lk += KSTEP;
    if (lk >= S.K) { ++ls; lk = 0; }
  };

  if (nsteps > 0) fetch(0);
  if (nsteps > 1) fetch(1);
  const int aoff = swz(wm0 + li, lg);   // same swizzle for every 16-row block
  const int boff = swz(wn0 + li, lg);
  int stage = 0;
  for (int step = 0; step < nsteps; ++step) {
    // this wave's loads of `step` are done once at most one newer stage is pending
    if (step + 1 < nsteps) wait_vm<LPS>();
    else wait_vm<0>();
    stage_barrier();   // every wave's loads of `step` landed; stage (step + 2) % 3 is free
    if (step + 2 < nsteps) fetch(stage == 0 ? 2 : stage - 1);

    const char* As = smem + stage * SBYTES;
    const char* Bs = As + NPART * ABYTES;
    ux4 fb[NPART][FN];
#pragma unroll
    for (int p = 0; p < NPART; ++p)
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[p][j] = *reinterpret_cast<const ux4*>(Bs + p * BBYTES + boff + j * 16 * ROWB);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      ux4 fa[NPART];
#pragma unroll
      for (int p = 0; p < NPART; ++p) fa[p] = *reinterpret_cast<const ux4*>(As + p * ABYTES + aoff + i * 16 * ROWB);
      if constexpr (MODE == M_BF16X3) {
        const bfx8 ah = __builtin_bit_cast(bfx8, fa[0]);
        const bfx8 al = __builtin_bit_cast(bfx8, fa[1]);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const bfx8 bh = __builtin_bit_cast(bfx8, fb[0][j]);
          const bfx8 bl = __builtin_bit_cast(bfx8, fb[1][j]);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah, acc[i][j], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int sub = 0; sub < 4; ++sub) {
          const unsigned int ua = fa[0][sub];
          const float a = __uint_as_float(ua);
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            const unsigned int ub = fb[0][j][sub];
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ub), a, acc[i][j], 0, 0, 0);
          }
        }
      }
    }
    stage = stage == NSTAGE - 1 ? 0 : stage + 1;
  }

  store_tile<FM, FN>(acc, P.C, P.ldc, P.M, P.N, m0 + wm0, n0 + wn0, li, lg, P.alpha, P.beta);
}

template <int BM, int BN, int MODE, int VAR>
mk_status launch_mixed(const ProbSpec* probs, int nprob, hipStream_t stream) {
  constexpr int EB = MODE == M_BF16X3 ? 2 : 4;
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
      const SegSpec& q = ps.seg[s];
      MK_REQUIRE(q.lda % 8 == 0 && q.ldb % 8 == 0 && q.K <= q.lda && q.K <= q.ldb, MK_BAD_ARG,
                 "mixed gemm: K-contiguous operands need row length padded to 8 (lda=%lld ldb=%lld K=%lld)",
                 (long long)q.lda, (long long)q.ldb, (long long)q.K);
      const void* ptrs[4] = {q.A, q.B, q.Alo, q.Blo};
      for (int z = 0; z < (MODE == M_BF16X3 ? 4 : 2); ++z)
        MK_REQUIRE(ptrs[z] != nullptr && (uintptr_t)ptrs[z] % 16 == 0, MK_BAD_ARG,
                   "mixed gemm: operands must be non-null and 16-byte aligned");
      GemmSeg& gs = P.seg[P.nseg++];
      gs.A = q.A; gs.B = q.B; gs.Alo = q.Alo; gs.Blo = q.Blo;
      gs.lda = q.lda; gs.ldb = q.ldb;
      // K rounded up to the padded row: the padding is zero
      gs.K = (int32_t)((q.K + 7) / 8 * 8);
      // bit 0: vector loads; bits 1, 2: A, B in the packed panel layout
      gs.vec = 1 | (q.packA ? 2 : 0) | (q.packB ? 4 : 0);
      if (q.packA || q.packB)
        MK_REQUIRE(VAR == V_BIG2 && ps.M % 256 == 0 && ps.N % 256 == 0 && q.K % 32 == 0, MK_BAD_ARG,
                   "mixed gemm: packed operands need the 256 x 256 kernel, M, N multiples of 256 and K of 32");
    }
    P.tiles_m = (int32_t)cdiv(ps.M, BM);
    P.tiles_n = (int32_t)cdiv(ps.N, BN);
    P.tile_start = total;
    total += P.tiles_m * P.tiles_n;
    P.C = ps.C; P.ldc = ps.ldc; P.alpha = ps.alpha; P.beta = ps.beta;
  }
  (void)EB;
  b.nprob = nprob;
  b.total_tiles = total;
  if (total == 0) return MK_OK;
  ProfScope prof(MODE == M_BF16X3 ? "gemm_bf16x3" : "gemm_f32", stream);
  if constexpr (VAR == V_PIPE)
    hipLaunchKernelGGL((gemm_mixed_pipe_kernel<MODE>), dim3(total), dim3(PNT), 0, stream, b);
  else if constexpr (VAR == V_BIG)
    hipLaunchKernelGGL((gemm_mixed_big_kernel<MODE, GemmBatch>), dim3(total), dim3(BNT), 0, stream, b);
  else if constexpr (VAR == V_BIG2)
    hipLaunchKernelGGL((gemm_mixed_big2_kernel<MODE, GemmBatch>), dim3(total), dim3(BNT), 0, stream, b);
  else
    hipLaunchKernelGGL((gemm_mixed_kernel<BM, BN, MODE, GemmBatch>), dim3(total), dim3(NT), 0, stream, b);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

// 256 rows of every operand lie within the 32-bit buffer offsets of gemm_mixed_big2_kernel
template <int MODE>
bool rows256_fit(const ProbSpec* probs, int nprob) {
  constexpr int64_t EB = MODE == M_BF16X3 ? 2 : 4;
  for (int i = 0; i < nprob; ++i)
    for (int s = 0; s < probs[i].nseg; ++s)
      if (256 * std::max(probs[i].seg[s].lda, probs[i].seg[s].ldb) * EB >= (1ll << 31)) return false;
  return true;
}

template <int MODE>
mk_status dispatch(const ProbSpec* probs, int nprob, hipStream_t stream) {
  MK_REQUIRE(nprob >= 0 && nprob <= MAX_PROB, MK_BAD_ARG, "mixed gemm: too many problems");
  int64_t tiles128 = 0;
  for (int i = 0; i < nprob; ++i) {
    MK_REQUIRE(probs[i].nseg <= MAX_SEG, MK_BAD_ARG, "mixed gemm: too many segments");
    tiles128 += cdiv(probs[i].M, 128) * cdiv(probs[i].N, 128);
  }
  constexpr int64_t EB = MODE == M_BF16X3 ? 2 : 4;
  int64_t tiles256 = 0, tilesp = 0;
  bool off32 = true;   // buffer offsets of the pipelined kernel are 32-bit
  for (int i = 0; i < nprob; ++i) {
    tiles256 += cdiv(probs[i].M, 256) * cdiv(probs[i].N, 256);
    tilesp += cdiv(probs[i].M, PBM) * cdiv(probs[i].N, PBN);
    for (int s = 0; s < probs[i].nseg; ++s)
      off32 = off32 && probs[i].M * probs[i].seg[s].lda * EB < (1ll << 31) &&
              probs[i].N * probs[i].seg[s].ldb * EB < (1ll << 31);
  }
  const int var = variant();
  bool packed = false;
  for (int i = 0; i < nprob; ++i)
    for (int s = 0; s < probs[i].nseg; ++s) packed = packed || probs[i].seg[s].packA || probs[i].seg[s].packB;
  if (packed) {
    MK_REQUIRE(var == V_BIG2 && tiles256 >= 256 && rows256_fit<MODE>(probs, nprob), MK_BAD_ARG,
               "mixed gemm: packed operands are read by the 256 x 256 kernel only");
    return launch_mixed<256, 256, MODE, V_BIG2>(probs, nprob, stream);
  }
  // large tiles once they alone fill the 256 CUs; 128x128 while they fill them; else 64x64
  if (var == V_PIPE && tilesp >= 256 && off32) return launch_mixed<PBM, PBN, MODE, V_PIPE>(probs, nprob, stream);
  if (var == V_BIG2 && tiles256 >= 256 && rows256_fit<MODE>(probs, nprob))
    return launch_mixed<256, 256, MODE, V_BIG2>(probs, nprob, stream);
  if ((var == V_BIG || var == V_BIG2) && tiles256 >= 256)
    return launch_mixed<256, 256, MODE, V_BIG>(probs, nprob, stream);
  if (tiles128 >= 256) return launch_mixed<128, 128, MODE, V_TILE>(probs, nprob, stream);
  return launch_mixed<64, 64, MODE, V_TILE>(probs, nprob, stream);
}

template <int MODE>
mk_status launch_group(const GemmGroup& g, int kind, int grid, hipStream_t stream) {
  if (g.total_tiles == 0) return MK_OK;
  ProfScope prof(MODE == M_BF16X3 ? "gemm_bf16x3" : "gemm_f32", stream);
  if (kind == TK_256 && variant() == V_BIG2 && g.rows256_fit)
    hipLaunchKernelGGL((gemm_mixed_big2_kernel<MODE, GemmGroup>), dim3(grid), dim3(BNT), 0, stream, g);
  else if (kind == TK_256)
    hipLaunchKernelGGL((gemm_mixed_big_kernel<MODE, GemmGroup>), dim3(grid), dim3(BNT), 0, stream, g);
  else if (kind == TK_128)
    hipLaunchKernelGGL((gemm_mixed_kernel<128, 128, MODE, GemmGroup>), dim3(grid), dim3(NT), 0, stream, g);
  else
    hipLaunchKernelGGL((gemm_mixed_kernel<64, 64, MODE, GemmGroup>), dim3(grid), dim3(NT), 0, stream, g);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

template <class K>
int occ_of(K kern, int threads) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, threads, 0) != hipSuccess) n = 1;
  return n > 0 ? n : 1;
}

}  // namespace

int mixed_group_grid(int prec, int kind, int total_tiles, int cus) {
  // occupancy is a property of the kernel's resources; the BF16X3 instances are the larger
  static const int occ[3] = {occ_of(gemm_mixed_kernel<64, 64, M_BF16X3, GemmGroup>, NT),
                             occ_of(gemm_mixed_kernel<128, 128, M_BF16X3, GemmGroup>, NT),
                             occ_of(gemm_mixed_big_kernel<M_BF16X3, GemmGroup>, BNT)};
  (void)prec;
  return (int)std::min<int64_t>(total_tiles, (int64_t)occ[kind] * cus);
}

mk_status gemm_mixed_group(const GemmGroup& g, int prec, int kind, int grid, hipStream_t st) {
  MK_REQUIRE(kind >= TK_64 && kind <= TK_256, MK_BAD_ARG, "gemm_mixed_group: tile kind %d", kind);
  if (prec == MK_BF16X3) return launch_group<M_BF16X3>(g, kind, grid, st);
  return launch_group<M_F32>(g, kind, grid, st);
}

bool mixed_packed_ok(int64_t M, int64_t N, const int64_t* Ks, int nk) {
  if (variant() != V_BIG2 || M % 256 || N % 256 || (M / 256) * (N / 256) < 256) return false;
  for (int i = 0; i < nk; ++i)
    // a panel (256 rows x K, fp32 the wider) must stay within 32-bit buffer offsets
    if (Ks[i] % 32 || Ks[i] * 4 * 256 >= (1ll << 31)) return false;
  return true;
}

mk_status gemm_bf16x3(const ProbSpec* probs, int nprob, hipStream_t stream) {
  return dispatch<M_BF16X3>(probs, nprob, stream);
}
mk_status gemm_f32(const ProbSpec* probs, int nprob, hipStream_t stream) {
  return dispatch<M_F32>(probs, nprob, stream);
}

}  // namespace mk
