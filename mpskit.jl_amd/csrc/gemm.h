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

struct GemmBatch {
  GemmProb p[MAX_PROB];
  int32_t nprob;
  int32_t total_tiles;
};

// Host-side builders -------------------------------------------------------

struct SegSpec {
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  int64_t K;
  const void* Alo = nullptr;
  const void* Blo = nullptr;
};

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

}  // namespace mk
