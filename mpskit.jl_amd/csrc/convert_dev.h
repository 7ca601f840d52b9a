// Device helpers shared by the conversion kernels of single applies and of
// grouped (batched) applies.
#pragma once
#include <hip/hip_bf16.h>

#include <cstdint>

#include "gemm.h"

namespace mk {

// bf16 hi/lo split of a float64 value: hi = bf16(v), lo = bf16(v - hi).
__device__ __forceinline__ void split2(double v, __hip_bfloat16& hi, __hip_bfloat16& lo) {
  const float f = (float)v;
  hi = __float2bfloat16(f);
  const double r = v - (double)__bfloat162float(hi);
  lo = __float2bfloat16((float)r);
}

// One pass over x for the mixed-precision nearest-neighbour two-site apply
// (d1 = d2 = 2, x an M x N column-major matrix, M = 2 Dl, N = 2 Dr, row
// r' = a + Dl h, column c' = c + Dr j):
//   y = alpha * CB(x) + beta * y                       (the physical term)
//   xb[c' * kb + r'] = conv(x[r', c'])                 (x as right factor, K = rows)
//   xa[r' * ka + c'] = conv(x[r', c'])                 (x as left factor, K = cols)
// with conv the bf16 hi/lo split (SPLIT) or the fp32 rounding.  A 256-thread
// block owns the (a, c) tile [a0, a0 + 32) x [c0, c0 + 64) and all four (h, j)
// sub-tiles of it, so x is read once: the thread holding x[a, :, c, :] produces
// the CB outputs of that pair, and the converted values pass through the LDS
// array t (33 KB) for the transposed xa stores.  Row padding of xa / xb is not
// written (K = N and M are multiples of 8 here, so no K chunk the GEMMs read
// reaches into it).
constexpr int PRO_A = 32, PRO_C = 64;
typedef unsigned int ProTile[4][PRO_C][PRO_A + 1];

// PACK: the two views are written in the packed panel layout of the 256 x 256
// mixed kernel (packed_offset, with ka = N and kb = M their K lengths) instead
// of K-contiguous rows of length ka / kb.
template <bool SPLIT, bool PACK = false>
__device__ __forceinline__ void nn_prologue_tile(const double* __restrict__ P, const double* __restrict__ x,
                                                 int64_t Dl, int64_t Dr, double alpha, double beta,
                                                 double* __restrict__ y, void* __restrict__ xa_hi,
                                                 void* __restrict__ xa_lo, int64_t ka, void* __restrict__ xb_hi,
                                                 void* __restrict__ xb_lo, int64_t kb, int64_t a0, int64_t c0,
                                                 ProTile& t) {
  // all arrays are in device memory: access them through global-typed pointers
  typedef __attribute__((address_space(1))) double gdouble;
  const gdouble* __restrict__ gx = (const gdouble*)x;
  gdouble* __restrict__ gy = (gdouble*)y;
  typedef __attribute__((address_space(1))) float gf;
  typedef __attribute__((address_space(1))) unsigned short gus;
  typedef __attribute__((address_space(1))) unsigned int gui;
  const int64_t M = 2 * Dl;
  double p[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) p[i] = ((const gdouble*)P)[i];
  {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // a offset, c phase (8)
    const int64_t a = a0 + tx;
#pragma unroll
    for (int k = 0; k < PRO_C / 8; ++k) {
      const int cc = ty + 8 * k;
      const int64_t c = c0 + cc;
      if (a < Dl && c < Dr) {
        double xv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[q] = gx[a + Dl * (q & 1) + M * (c + Dr * (q >> 1))];
#pragma unroll
        for (int o = 0; o < 4; ++o) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) s += p[(o & 1) + 2 * ((o >> 1) + 2 * q)] * xv[q];
          gdouble* yp = gy + a + Dl * (o & 1) + M * (c + Dr * (o >> 1));
          double v = alpha * s;
          if (beta != 0.0) v += beta * *yp;
          *yp = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t rr = a + Dl * (q & 1), cl = c + Dr * (q >> 1);
          const int64_t ob = PACK ? packed_offset(cl, rr, kb, SPLIT ? 32 : 16) : cl * kb + rr;
          unsigned int packed;
          if constexpr (SPLIT) {
            __hip_bfloat16 h, l;
            split2(xv[q], h, l);
            ((gus*)xb_hi)[ob] = __builtin_bit_cast(unsigned short, h);
            ((gus*)xb_lo)[ob] = __builtin_bit_cast(unsigned short, l);
            packed = (unsigned int)__builtin_bit_cast(unsigned short, h) |
                     ((unsigned int)__builtin_bit_cast(unsigned short, l) << 16);
          } else {
            const float f = (float)xv[q];
            ((gf*)xb_hi)[ob] = f;
            packed = __float_as_uint(f);
          }
          t[q][cc][tx] = packed;
        }
      }
    }
  }
  __syncthreads();
  {
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // c offset, a phase (4)
    const int64_t c = c0 + tx;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int k = 0; k < PRO_A / 4; ++k) {
        const int aa = ty + 4 * k;
        const int64_t ar = a0 + aa;
        if (ar >= Dl || c >= Dr) continue;
        const unsigned int v = t[q][tx][aa];
        const int64_t row = ar + Dl * (q & 1), kk = c + Dr * (q >> 1);
        const int64_t idx = PACK ? packed_offset(row, kk, ka, SPLIT ? 32 : 16) : row * ka + kk;
        if constexpr (SPLIT) {
          ((gus*)xa_hi)[idx] = (unsigned short)(v & 0xffffu);
          ((gus*)xa_lo)[idx] = (unsigned short)(v >> 16);
        } else {
          ((gui*)xa_hi)[idx] = v;
        }
      }
    }
  }
}

}  // namespace mk
