// Multi-segment, multi-problem column-major GEMM descriptors.
//
// Every contraction of the derivative / transfer path is a sum of plain
// column-major matrix products with free leg fusions (reshapes), so one GEMM
// family serves all of them:
//
//   C_p (M_p x N_p) = alpha_p * sum_s op(A_ps) op(B_ps) + beta_p * C_p
//
// A launch holds up to MAX_PROB independent problems (e.g. the d output
// blocks of a strided term) and each problem sums up to MAX_SEG segments
// (e.g. x*IC + BE*x of the two-site Jordan apply) into one accumulator, so the
// output tile is written exactly once.
#pragma once
#include "common.h"

namespace mk {

constexpr int MAX_SEG = 4;
constexpr int MAX_PROB = 4;

struct GemmSeg {
  const void* A;     // fp64 operand, or the hi part (bf16) / fp32 copy
  const void* B;
  const void* Alo;   // bf16x3 lo parts (unused otherwise)
  const void* Blo;
  int64_t lda, ldb;
  int32_t K;
  int32_t vec;       // 1 if 16-byte vector loads are legal for A and B
};

struct GemmProb {
  GemmSeg seg[MAX_SEG];
  int32_t nseg;
  int32_t M, N;
  int32_t tiles_m, tiles_n, tile_start;
  double* C;
  int64_t ldc;
  double alpha, beta;
};

// Up to MAX_PROB problems passed by value as the kernel argument.
struct GemmBatch {
  GemmProb p[MAX_PROB];
  int32_t nprob;
  int32_t total_tiles;
};

// Any number of problems in a device-resident array (problems sorted by
// decreasing per-tile work, tile_start ascending), worked through by a
// persistent grid: workgroup b takes tiles b, b + gridDim.x, ...
// Pointer fields of the array may be relative: bits 62-63 name a base slot
// (1..3) of `base` and the low bits are a byte offset into it, so one array
// serves every call whose inputs and outputs live at other addresses.
constexpr int NBASE = 4;
struct GemmGroup {
  const GemmProb* probs;
  int32_t nprob;
  int32_t total_tiles;
  const char* base[NBASE];
  int32_t rows256_fit;   // 256 rows of every mixed operand span < 2 GiB (32-bit buffer offsets)
  int32_t f64_buf;       // every FP64 problem meets the branch-free kernel's conditions (f64_buf_ok)
};

constexpr uint64_t REL_MASK = (1ull << 62) - 1;
inline const void* rel_ptr(int slot, size_t offset) {
  return reinterpret_cast<const void*>((static_cast<uint64_t>(slot) << 62) | offset);
}

// Host-side builders -------------------------------------------------------

struct SegSpec {
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  int64_t K;
  const void* Alo = nullptr;
  const void* Blo = nullptr;
  // mixed precision only: the operand is stored in the packed panel layout
  // (packed_offset below; lda / ldb then equal K)
  bool packA = false, packB = false;
};

// Packed operand layout of the 256 x 256 mixed-precision kernel.  An operand
// with MN rows (its non-K index, a multiple of 256) and K (a multiple of 32)
// K-contiguous elements per row is stored as MN / 256 panels of K / KS slabs
// (KS = 32 bf16 or 16 fp32 elements: 64 bytes); slab s of panel p holds the
// 64 bytes of K step s of rows 256 p .. 256 p + 255, row after row.  A K step
// of a tile is then one contiguous 16 KB read, and every 128-byte line it
// touches is used in that step.  Element offset of (row r, k):
__host__ __device__ inline int64_t packed_offset(int64_t r, int64_t k, int64_t K, int ks) {
  return ((r >> 8) * (K / ks) + k / ks) * (256 * (int64_t)ks) + (r & 255) * ks + k % ks;
}

struct ProbSpec {
  int64_t M = 0, N = 0;
  SegSpec seg[MAX_SEG];
  int nseg = 0;
  double* C = nullptr;
  int64_t ldc = 0;
  double alpha = 1.0, beta = 0.0;
  void add(const void* A, int64_t lda, const void* B, int64_t ldb, int64_t K,
           const void* Alo = nullptr, const void* Blo = nullptr) {
    seg[nseg].A = A; seg[nseg].lda = lda; seg[nseg].B = B; seg[nseg].ldb = ldb;
    seg[nseg].K = K; seg[nseg].Alo = Alo; seg[nseg].Blo = Blo;
    ++nseg;
  }
};

// Operand storage transposition of a GEMM launch (uniform over its segments):
//   TA = false: A(i,k) = A[i + k*lda]   TA = true: A(i,k) = A[k + i*lda]
//   TB = false: B(k,j) = B[k + j*ldb]   TB = true: B(k,j) = B[j + k*ldb]
struct GemmShape {
  bool ta = false, tb = false;
};

// FP64 MFMA path (gemm_f64.hip).  Enqueues on `stream`.
mk_status gemm_f64(const ProbSpec* probs, int nprob, GemmShape shape, hipStream_t stream);

// Mixed paths (gemm_mixed.hip).  Operands are pre-converted and K-contiguous:
//   A(i,k) = A[i*lda + k], B(k,j) = B[j*ldb + k]  (bf16 hi/lo or fp32),
// accumulation in fp32, output C is float64 column-major (alpha/beta in fp64).
mk_status gemm_bf16x3(const ProbSpec* probs, int nprob, hipStream_t stream);
mk_status gemm_f32(const ProbSpec* probs, int nprob, hipStream_t stream);
// Whether one M x N problem whose segments have the given K values runs on the
// 256 x 256 kernel with packed operands (the only kernel that reads them).
bool mixed_packed_ok(int64_t M, int64_t N, const int64_t* Ks, int nk);

// Grouped (persistent) launches over a device descriptor array.  Tile kinds:
//   FP64:  TK_64 (64x64), TK_128 (128x128)
//   mixed: TK_64, TK_128, TK_256 (256x256, 8 waves)
enum TileKind { TK_64 = 0, TK_128 = 1, TK_256 = 2 };
inline int tile_m(int kind) { return kind == TK_64 ? 64 : kind == TK_128 ? 128 : 256; }
inline int tile_n(int kind) { return tile_m(kind); }
bool tile_valid(int prec, int kind);
// Workgroups of one persistent launch: resident workgroups per CU x CUs, capped by the tiles.
int group_grid(int prec, int kind, GemmShape sh, int total_tiles);
mk_status gemm_f64_group(const GemmGroup& g, int kind, GemmShape sh, int grid, hipStream_t st);
// Whether the FP64 problems may run on the branch-free (buffer-load) kernel
// with tile kind `kind`: 16-byte aligned operands (relative pointers: their
// offsets), even leading dimensions and K, even row counts for operands stored
// along their rows, in-tile offsets below 2 GiB.
bool f64_buf_ok(const ProbSpec* probs, int n, GemmShape sh, int kind);
mk_status gemm_mixed_group(const GemmGroup& g, int prec, int kind, int grid, hipStream_t st);

}  // namespace mk
