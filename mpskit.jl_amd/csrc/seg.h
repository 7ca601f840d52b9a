// Segmented-vector kernels and grouped memory-bound passes of batched applies.
#pragma once
#include "common.h"
#include "gemm.h"

namespace mk {

// Segments start on a chunk boundary and are zero-padded to the next one.
constexpr int64_t SEG_CHUNK = 4096;

// Device-side segment tables (all arrays on the device).
struct Segs {
  int nseg = 0;
  int nchunks = 0;
  const int* chunk_seg = nullptr;  // segment of every chunk
  const int* first = nullptr;      // first chunk of every segment
  const int* count = nullptr;      // chunks of every segment
};

// One conversion of a float64 column-major matrix into a K-contiguous mixed
// operand (see k_split_bf16 / k_to_f32); pointers may be relative.
struct SplitItem {
  const double* src;
  int64_t rows, cols, ld, ldo;
  int32_t trans, nblocks;
  void* hi;
  void* lo;
};

// One two-site physical term y = alpha P(x) + beta y; x, y may be relative, P absolute.
struct PhysItem {
  const double* P;
  const double* x;
  double* y;
  int64_t Dl, Dr;
  int32_t d1, d2, nblocks, pad;
  double alpha, beta;
};

// One output block of the two-site physical term on charge-block storage:
// y = alpha sum_s coef[s] src[s] + beta y over `len` contiguous elements (the
// source blocks hold the same bond sectors as y, so the sum is elementwise);
// y and src may be relative.
constexpr int CPHYS_MAX = 16;
struct CPhysItem {
  double* y;
  const double* src[CPHYS_MAX];
  double coef[CPHYS_MAX];
  int64_t len;
  int32_t nsrc, nblocks;
};

// out[s * ostride + j] = <V_j, u> restricted to segment s, j < k.
// `partial` holds k * nchunks doubles.
mk_status seg_dots(const Segs& sg, const double* V, int64_t ldv, int k, const double* u, double* out, int ostride,
                   double* partial, hipStream_t st);
// u -= sum_j V_j c[s * cstride + j] on every segment s.
mk_status seg_multiaxpy(const Segs& sg, const double* V, int64_t ldv, int k, const double* c, int cstride, double* u,
                        hipStream_t st);
// x = sum_j V_j c[s * cstride + j] on every segment s.
mk_status seg_gemv(const Segs& sg, const double* V, int64_t ldv, int k, const double* c, int cstride, double* x,
                   hipStream_t st);
// y = a[s] x + b[s] y (b[s] == 0 ignores y).
mk_status seg_axpby(const Segs& sg, const double* a, const double* x, const double* b, double* y, hipStream_t st);
// y = x / sqrt(d[s * dstride]) (zero where d <= 0).
mk_status seg_normalize(const Segs& sg, const double* x, const double* d, int dstride, double* y, hipStream_t st);

mk_status split_group(const SplitItem* items, int n, int max_blocks, bool bf16, const void* const base[NBASE],
                      hipStream_t st);
int split_blocks(int64_t rows, int64_t cols, int trans, int64_t ldo);
mk_status phys2_group(const PhysItem* items, int n, int max_blocks, int d12, const void* const base[NBASE],
                      hipStream_t st);
int phys2_blocks(int64_t Dl, int64_t Dr);
mk_status cphys_group(const CPhysItem* items, int n, int max_blocks, double alpha, double beta,
                      const void* const base[NBASE], hipStream_t st);
int cphys_blocks(int64_t len);

// One nearest-neighbour two-site prologue (d1 = d2 = 2): y = alpha CB(x) + beta y
// and the K-contiguous conversions of x as left factor (xa, row length ka) and
// as right factor (xb, row length kb); x, y, xa, xb may be relative, P absolute.
struct NNItem {
  const double* P;
  const double* x;
  double* y;
  void* xa_hi;
  void* xa_lo;
  void* xb_hi;
  void* xb_lo;
  int64_t Dl, Dr, ka, kb;
  int32_t tiles_a, tiles;
  double alpha, beta;
};
// Items along blockIdx.y, one (32 x 64) (a, c) tile per block; max_tiles = max tiles.
mk_status nn_group(const NNItem* items, int n, int max_tiles, bool bf16, const void* const base[NBASE],
                   hipStream_t st);
void nn_tiles(int64_t Dl, int64_t Dr, int32_t& tiles_a, int32_t& tiles);

}  // namespace mk
