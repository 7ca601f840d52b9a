// Kernels over "segmented" vectors: several independent vectors (one per
// instance of a batch) concatenated, each starting on a SEG_CHUNK boundary and
// zero-padded to the next one, so that every chunk of SEG_CHUNK elements
// belongs to exactly one segment.  Reductions produce one value per
// (segment, column); coefficients are read per segment.  Also the grouped
// precision conversion and two-site physical-term kernels of batched applies,
// which take device descriptor arrays whose pointers may be relative to the
// base slots passed by value (so the launches can be captured once and
// replayed).
#include <hip/hip_bf16.h>

#include "gemm.h"
#include "seg.h"
#include "convert_dev.h"

namespace mk {
namespace {

constexpr int TPB = 256;
constexpr int ITEMS = SEG_CHUNK / TPB;  // elements per thread per chunk

struct BaseArgs {
  const char* b[NBASE];
};

// Operands live in device memory: the round trip through the global address
// space keeps the compiler from accessing a pointer rebuilt from an integer
// with flat instructions.
template <class T>
__device__ __forceinline__ T* rslv(const BaseArgs& ba, T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const int slot = (int)(v >> 62);
  T* r = slot == 0 ? p : reinterpret_cast<T*>(const_cast<char*>(ba.b[slot]) + (v & REL_MASK));
  return (T*)((__attribute__((address_space(1))) T*)r);
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  __syncthreads();
  return r;
}

// partial[j * nchunks + c] = <V_j, u> over chunk c
__global__ void seg_dot_kernel(const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ u,
                               double* __restrict__ partial) {
  __shared__ double sh[TPB / 64];
  const int64_t base = (int64_t)blockIdx.x * SEG_CHUNK;
  double ur[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) ur[it] = u[base + it * TPB + threadIdx.x];
  for (int j = 0; j < k; ++j) {
    const double* Vj = V + (int64_t)j * ldv + base;
    double s = 0.0;
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) s += Vj[it * TPB + threadIdx.x] * ur[it];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = s;
  }
}

// out[s * ostride + j] = sum over the chunks of segment s (grid: nseg x k)
__global__ void seg_reduce_kernel(const double* __restrict__ partial, int nchunks, const int* __restrict__ first,
                                  const int* __restrict__ count, double* __restrict__ out, int ostride) {
  __shared__ double sh[TPB / 64];
  const int s = blockIdx.x, j = blockIdx.y;
  const double* p = partial + (int64_t)j * nchunks + first[s];
  double v = 0.0;
  for (int c = threadIdx.x; c < count[s]; c += TPB) v += p[c];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) out[(int64_t)s * ostride + j] = v;
}

// u -= sum_j V_j c[seg * cstride + j]
__global__ void seg_multiaxpy_kernel(const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ c,
                                     int cstride, const int* __restrict__ chunk_seg, double* __restrict__ u) {
  const int s = chunk_seg[blockIdx.x];
  const double* cs = c + (int64_t)s * cstride;
  const int64_t base = (int64_t)blockIdx.x * SEG_CHUNK;
#pragma unroll 4
  for (int it = 0; it < ITEMS; ++it) {
    const int64_t i = base + it * TPB + threadIdx.x;
    double v = u[i];
    for (int j = 0; j < k; ++j) v -= cs[j] * V[i + (int64_t)j * ldv];
    u[i] = v;
  }
}

// x = sum_j V_j c[seg * cstride + j]
__global__ void seg_gemv_kernel(const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ c,
                                int cstride, const int* __restrict__ chunk_seg, double* __restrict__ x) {
  const int s = chunk_seg[blockIdx.x];
  const double* cs = c + (int64_t)s * cstride;
  const int64_t base = (int64_t)blockIdx.x * SEG_CHUNK;
#pragma unroll 4
  for (int it = 0; it < ITEMS; ++it) {
    const int64_t i = base + it * TPB + threadIdx.x;
    double v = 0.0;
    for (int j = 0; j < k; ++j) v += cs[j] * V[i + (int64_t)j * ldv];
    x[i] = v;
  }
}

// y = a[seg] x + b[seg] y
__global__ void seg_axpby_kernel(const double* __restrict__ a, const double* x, const double* __restrict__ b,
                                 double* y, const int* __restrict__ chunk_seg) {
  const int s = chunk_seg[blockIdx.x];
  const double as = a[s], bs = b[s];
  const int64_t base = (int64_t)blockIdx.x * SEG_CHUNK;
#pragma unroll 4
  for (int it = 0; it < ITEMS; ++it) {
    const int64_t i = base + it * TPB + threadIdx.x;
    const double yv = bs == 0.0 ? 0.0 : bs * y[i];
    y[i] = as * x[i] + yv;
  }
}

// y = x / sqrt(d[seg * dstride]) (y = 0 where d <= 0)
__global__ void seg_normalize_kernel(const double* x, const double* __restrict__ d, int dstride, double* y,
                                     const int* __restrict__ chunk_seg) {
  const int s = chunk_seg[blockIdx.x];
  const double dv = d[(int64_t)s * dstride];
  const double f = dv > 0.0 ? 1.0 / sqrt(dv) : 0.0;
  const int64_t base = (int64_t)blockIdx.x * SEG_CHUNK;
#pragma unroll 4
  for (int it = 0; it < ITEMS; ++it) {
    const int64_t i = base + it * TPB + threadIdx.x;
    y[i] = f * x[i];
  }
}

// ---- grouped conversions ------------------------------------------------------

template <bool BF>
__device__ __forceinline__ void put(const SplitItem& d, void* hi, void* lo, int64_t idx, double v) {
  if (BF) {
    __hip_bfloat16 h, l;
    split2(v, h, l);
    static_cast<__hip_bfloat16*>(hi)[idx] = h;
    static_cast<__hip_bfloat16*>(lo)[idx] = l;
  } else {
    static_cast<float*>(hi)[idx] = (float)v;
  }
}

// One item per blockIdx.y.  trans = 1: out[c * ldo + r] = M[r + c * ld] (zero
// for rows <= r < ldo), 1024 elements per block; trans = 0: out[r * ldo + c] =
// M[r + c * ld] (zero for cols <= c < ldo) through 64 x 64 LDS tiles.
template <bool BF>
__global__ void split_group_kernel(const SplitItem* __restrict__ items, BaseArgs ba) {
  const SplitItem d = items[blockIdx.y];
  if ((int)blockIdx.x >= d.nblocks) return;
  const double* M = rslv(ba, d.src);
  void* hi = rslv(ba, d.hi);
  void* lo = BF ? rslv(ba, d.lo) : nullptr;
  if (d.trans) {
    const int64_t n = d.ldo * d.cols;
    const int64_t b0 = (int64_t)blockIdx.x * 1024;
    for (int t = threadIdx.x; t < 1024; t += TPB) {
      const int64_t idx = b0 + t;
      if (idx >= n) break;
      const int64_t r = idx % d.ldo, c = idx / d.ldo;
      put<BF>(d, hi, lo, idx, r < d.rows ? M[r + c * d.ld] : 0.0);
    }
  } else {
    __shared__ double tile[64][65];
    const int64_t tiles_r = (d.rows + 63) / 64;
    const int64_t r0 = (blockIdx.x % tiles_r) * 64, c0 = (blockIdx.x / tiles_r) * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int cc = ty; cc < 64; cc += 4) {
      const int64_t r = r0 + tx, c = c0 + cc;
      tile[cc][tx] = (r < d.rows && c < d.cols) ? M[r + c * d.ld] : 0.0;
    }
    __syncthreads();
    for (int rr = ty; rr < 64; rr += 4) {
      const int64_t r = r0 + rr, c = c0 + tx;
      if (r < d.rows && c < d.ldo) put<BF>(d, hi, lo, r * d.ldo + c, tile[tx][rr]);
    }
  }
}

// Two-site physical term of batched applies, one item per blockIdx.y:
// y[a,b,c,e] = alpha sum_{h,j} P[b,e,h,j] x[a,h,c,j] + beta y  (one thread per (a, c)).
template <int D12>
__global__ void phys2_group_kernel(const PhysItem* __restrict__ items, BaseArgs ba) {
  const PhysItem d = items[blockIdx.y];
  const int64_t n = d.Dl * d.Dr;
  const int64_t idx = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (idx >= n) return;
  typedef __attribute__((address_space(1))) double gdouble;
  const gdouble* x = (const gdouble*)rslv(ba, d.x);
  gdouble* y = (gdouble*)rslv(ba, d.y);
  const gdouble* P = (const gdouble*)d.P;
  const int d1 = d.d1, d2 = d.d2;
  const int64_t a = idx % d.Dl, c = idx / d.Dl;
  const int64_t base = a + d.Dl * d1 * c;
  const int64_t sx = d.Dl * d1 * d.Dr;
  double xv[D12];
#pragma unroll
  for (int t = 0; t < D12; ++t) xv[t] = x[base + d.Dl * (t % d1) + sx * (t / d1)];
#pragma unroll
  for (int o = 0; o < D12; ++o) {
    const int b = o % d1, e = o / d1;
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < D12; ++t) s += P[b + d1 * (e + d2 * t)] * xv[t];
    gdouble* yp = y + base + d.Dl * b + sx * e;
    double v = d.alpha * s;
    if (d.beta != 0.0) v += d.beta * *yp;
    *yp = v;
  }
}

// Physical term on charge-block storage, one output block per blockIdx.y,
// 4 elements per thread.
__global__ void cphys_group_kernel(const CPhysItem* __restrict__ items, double alpha, double beta, BaseArgs ba) {
  const CPhysItem& d = items[blockIdx.y];
  const int64_t i0 = ((int64_t)blockIdx.x * TPB + threadIdx.x) * 4;
  if (i0 >= d.len) return;
  typedef __attribute__((address_space(1))) double gdouble;
  gdouble* y = (gdouble*)rslv(ba, d.y);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int nsrc = d.nsrc;
  for (int s = 0; s < nsrc; ++s) {
    const gdouble* x = (const gdouble*)rslv(ba, d.src[s]);
    const double c = d.coef[s];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (i0 + t < d.len) acc[t] += c * x[i0 + t];
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (i0 + t >= d.len) break;
    double v = alpha * acc[t];
    if (beta != 0.0) v += beta * y[i0 + t];
    y[i0 + t] = v;
  }
}

template <bool BF>
__global__ __launch_bounds__(256) void nn_group_kernel(const NNItem* __restrict__ items, BaseArgs ba) {
  __shared__ ProTile t;
  const NNItem d = items[blockIdx.y];
  if ((int)blockIdx.x >= d.tiles) return;
  const int64_t a0 = (int64_t)(blockIdx.x % d.tiles_a) * PRO_A, c0 = (int64_t)(blockIdx.x / d.tiles_a) * PRO_C;
  nn_prologue_tile<BF>(d.P, rslv(ba, d.x), d.Dl, d.Dr, d.alpha, d.beta, rslv(ba, d.y), rslv(ba, d.xa_hi),
                       BF ? rslv(ba, d.xa_lo) : nullptr, d.ka, rslv(ba, d.xb_hi), BF ? rslv(ba, d.xb_lo) : nullptr,
                       d.kb, a0, c0, t);
}

BaseArgs to_args(const void* const base[NBASE]) {
  BaseArgs ba{};
  for (int s = 0; s < NBASE; ++s) ba.b[s] = static_cast<const char*>(base ? base[s] : nullptr);
  return ba;
}

}  // namespace

mk_status seg_dots(const Segs& sg, const double* V, int64_t ldv, int k, const double* u, double* out, int ostride,
                   double* partial, hipStream_t st) {
  if (k <= 0 || sg.nchunks == 0) return MK_OK;
  hipLaunchKernelGGL(seg_dot_kernel, dim3(sg.nchunks), dim3(TPB), 0, st, V, ldv, k, u, partial);
  MK_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(seg_reduce_kernel, dim3(sg.nseg, k), dim3(TPB), 0, st, partial, sg.nchunks, sg.first, sg.count,
                     out, ostride);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status seg_multiaxpy(const Segs& sg, const double* V, int64_t ldv, int k, const double* c, int cstride, double* u,
                        hipStream_t st) {
  if (k <= 0 || sg.nchunks == 0) return MK_OK;
  hipLaunchKernelGGL(seg_multiaxpy_kernel, dim3(sg.nchunks), dim3(TPB), 0, st, V, ldv, k, c, cstride, sg.chunk_seg, u);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status seg_gemv(const Segs& sg, const double* V, int64_t ldv, int k, const double* c, int cstride, double* x,
                   hipStream_t st) {
  if (sg.nchunks == 0) return MK_OK;
  hipLaunchKernelGGL(seg_gemv_kernel, dim3(sg.nchunks), dim3(TPB), 0, st, V, ldv, k, c, cstride, sg.chunk_seg, x);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status seg_axpby(const Segs& sg, const double* a, const double* x, const double* b, double* y, hipStream_t st) {
  if (sg.nchunks == 0) return MK_OK;
  hipLaunchKernelGGL(seg_axpby_kernel, dim3(sg.nchunks), dim3(TPB), 0, st, a, x, b, y, sg.chunk_seg);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status seg_normalize(const Segs& sg, const double* x, const double* d, int dstride, double* y, hipStream_t st) {
  if (sg.nchunks == 0) return MK_OK;
  hipLaunchKernelGGL(seg_normalize_kernel, dim3(sg.nchunks), dim3(TPB), 0, st, x, d, dstride, y, sg.chunk_seg);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status split_group(const SplitItem* items, int n, int max_blocks, bool bf16, const void* const base[NBASE],
                      hipStream_t st) {
  if (n == 0 || max_blocks == 0) return MK_OK;
  MK_REQUIRE(n < 65536, MK_BAD_ARG, "split_group: too many items");
  const BaseArgs ba = to_args(base);
  if (bf16) hipLaunchKernelGGL(split_group_kernel<true>, dim3(max_blocks, n), dim3(TPB), 0, st, items, ba);
  else hipLaunchKernelGGL(split_group_kernel<false>, dim3(max_blocks, n), dim3(TPB), 0, st, items, ba);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

int split_blocks(int64_t rows, int64_t cols, int trans, int64_t ldo) {
  if (trans) return (int)cdiv(ldo * cols, (int64_t)1024);
  return (int)(cdiv(rows, (int64_t)64) * cdiv(ldo, (int64_t)64));
}

mk_status phys2_group(const PhysItem* items, int n, int max_blocks, int d12, const void* const base[NBASE],
                      hipStream_t st) {
  if (n == 0 || max_blocks == 0) return MK_OK;
  MK_REQUIRE(n < 65536, MK_BAD_ARG, "phys2_group: too many items");
  const BaseArgs ba = to_args(base);
  const dim3 g(max_blocks, n);
  switch (d12) {
    case 4: hipLaunchKernelGGL(phys2_group_kernel<4>, g, dim3(TPB), 0, st, items, ba); break;
    case 9: hipLaunchKernelGGL(phys2_group_kernel<9>, g, dim3(TPB), 0, st, items, ba); break;
    case 16: hipLaunchKernelGGL(phys2_group_kernel<16>, g, dim3(TPB), 0, st, items, ba); break;
    default: MK_REQUIRE(false, MK_NOT_SUPPORTED, "phys2_group: d1*d2 = %d not in {4, 9, 16}", d12);
  }
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

int phys2_blocks(int64_t Dl, int64_t Dr) { return (int)cdiv(Dl * Dr, (int64_t)TPB); }

mk_status cphys_group(const CPhysItem* items, int n, int max_blocks, double alpha, double beta,
                      const void* const base[NBASE], hipStream_t st) {
  if (n == 0 || max_blocks == 0) return MK_OK;
  MK_REQUIRE(n < 65536, MK_BAD_ARG, "cphys_group: too many items");
  hipLaunchKernelGGL(cphys_group_kernel, dim3(max_blocks, n), dim3(TPB), 0, st, items, alpha, beta, to_args(base));
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

int cphys_blocks(int64_t len) { return (int)cdiv(len, (int64_t)TPB * 4); }

void nn_tiles(int64_t Dl, int64_t Dr, int32_t& tiles_a, int32_t& tiles) {
  tiles_a = (int32_t)cdiv(Dl, (int64_t)PRO_A);
  tiles = tiles_a * (int32_t)cdiv(Dr, (int64_t)PRO_C);
}

mk_status nn_group(const NNItem* items, int n, int max_tiles, bool bf16, const void* const base[NBASE],
                   hipStream_t st) {
  if (n == 0 || max_tiles == 0) return MK_OK;
  MK_REQUIRE(n < 65536, MK_BAD_ARG, "nn_group: too many items");
  const BaseArgs ba = to_args(base);
  if (bf16) hipLaunchKernelGGL(nn_group_kernel<true>, dim3(max_tiles, n), dim3(256), 0, st, items, ba);
  else hipLaunchKernelGGL(nn_group_kernel<false>, dim3(max_tiles, n), dim3(256), 0, st, items, ba);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

}  // namespace mk
