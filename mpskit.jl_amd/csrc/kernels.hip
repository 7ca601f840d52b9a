// Memory-bound helper kernels; see kernels.h for the contracts.
#include "kernels.h"
#include "convert_dev.h"

#include <hip/hip_bf16.h>

#include <algorithm>

namespace mk {
namespace {

constexpr int TPB = 256;

inline dim3 grid1d(int64_t n, int per_thread = 1) {
  int64_t b = cdiv(n, (int64_t)TPB * per_thread);
  if (b < 1) b = 1;
  if (b > (1 << 30)) b = (1 << 30);
  return dim3((unsigned)b);
}

#define GRID_STRIDE(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

__device__ __forceinline__ double env_t(const double* G, int64_t D0, int w, int64_t a, int g,
                                        int64_t f, int ident) {
  // environment with channel `ident` replaced by the identity matrix
  if (g == ident) return a == f ? 1.0 : 0.0;
  return G[a + D0 * (g + (int64_t)w * f)];
}

__global__ void prep_L_kernel(const double* __restrict__ GL, const double* __restrict__ W,
                              const int* __restrict__ klist, int nk, int64_t Da, int64_t Df,
                              int wl, int d, int wr, int g_ident, double* __restrict__ L) {
  const int64_t n = Da * d * Df * d * nk;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % Da; t /= Da;
    const int b = (int)(t % d); t /= d;
    const int64_t f = t % Df; t /= Df;
    const int h = (int)(t % d); t /= d;
    const int kk = (int)t;
    const int k = klist[kk];
    double s = 0.0;
    for (int g = 0; g < wl; ++g) {
      const double wv = W[g + wl * (b + d * (h + d * k))];
      if (wv != 0.0) s += env_t(GL, Da, wl, a, g, f, g_ident) * wv;
    }
    L[idx] = s;
  }
}

__global__ void prep_R_kernel(const double* __restrict__ GR, const double* __restrict__ W,
                              const int* __restrict__ klist, int nk, int64_t Di, int64_t Dc,
                              int wl, int d, int wr, int l_ident, double* __restrict__ R) {
  const int64_t n = Di * d * nk * Dc * d;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t i = t % Di; t /= Di;
    const int j = (int)(t % d); t /= d;
    const int kk = (int)(t % nk); t /= nk;
    const int64_t c = t % Dc; t /= Dc;
    const int e = (int)t;
    const int k = klist[kk];
    double s = 0.0;
    for (int l = 0; l < wr; ++l) {
      const double wv = W[k + wl * (e + d * (j + d * l))];
      if (wv != 0.0) s += wv * env_t(GR, Di, wr, i, l, c, l_ident);
    }
    R[idx] = s;
  }
}

__global__ void prep_Cp_kernel(const double* __restrict__ GR, const double* __restrict__ W,
                               int64_t Di, int64_t Dc, int wl, int d, int wr, int l_ident,
                               double* __restrict__ Cp) {
  const int64_t n = d * Di * d * Dc;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int h = (int)(t % d); t /= d;
    const int64_t i = t % Di; t /= Di;
    const int b = (int)(t % d); t /= d;
    const int64_t c = t;
    double s = 0.0;
    for (int l = 0; l < wr; ++l) {
      const double wv = W[0 + wl * (b + d * (h + d * l))];
      if (wv != 0.0) s += wv * env_t(GR, Di, wr, i, l, c, l_ident);
    }
    Cp[idx] = s;
  }
}

__global__ void prep_CA_kernel(const double* __restrict__ GR, const double* __restrict__ W1,
                               const double* __restrict__ W2, int64_t Dr, int wl, int d1, int wm,
                               int d2, int wr, int l_ident, double* __restrict__ CA) {
  const int64_t n = (int64_t)d1 * Dr * d2 * d1 * Dr * d2;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int h = (int)(t % d1); t /= d1;
    const int64_t i = t % Dr; t /= Dr;
    const int j = (int)(t % d2); t /= d2;
    const int b = (int)(t % d1); t /= d1;
    const int64_t c = t % Dr; t /= Dr;
    const int e = (int)t;
    double s = 0.0;
    for (int k = 0; k < wm; ++k) {
      const double w1 = W1[0 + wl * (b + d1 * (h + d1 * k))];
      if (w1 == 0.0) continue;
      double r = 0.0;
      for (int l = 0; l < wr; ++l) {
        const double w2 = W2[k + wm * (e + d2 * (j + d2 * l))];
        if (w2 != 0.0) r += w2 * env_t(GR, Dr, wr, i, l, c, l_ident);
      }
      s += w1 * r;
    }
    CA[idx] = s;
  }
}

__global__ void prep_AB_kernel(const double* __restrict__ GL, const double* __restrict__ W1,
                               const double* __restrict__ W2, int64_t Dl, int wl, int d1, int wm,
                               int d2, int wr, int g_ident, double* __restrict__ AB) {
  const int64_t n = Dl * d1 * Dl * d1 * d2 * d2;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % Dl; t /= Dl;
    const int b = (int)(t % d1); t /= d1;
    const int64_t f = t % Dl; t /= Dl;
    const int h = (int)(t % d1); t /= d1;
    const int e = (int)(t % d2); t /= d2;
    const int j = (int)t;
    double s = 0.0;
    for (int k = 0; k < wm; ++k) {
      const double w2 = W2[k + wm * (e + d2 * (j + d2 * (wr - 1)))];
      if (w2 == 0.0) continue;
      double lsum = 0.0;
      for (int g = 0; g < wl; ++g) {
        const double w1 = W1[g + wl * (b + d1 * (h + d1 * k))];
        if (w1 != 0.0) lsum += env_t(GL, Dl, wl, a, g, f, g_ident) * w1;
      }
      s += lsum * w2;
    }
    AB[idx] = s;
  }
}

__global__ void swap23_kernel(const double* __restrict__ in, int64_t D0, int64_t D1, int64_t D2,
                              double* __restrict__ out) {
  const int64_t n = D0 * D1 * D2;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % D0; t /= D0;
    const int64_t f = t % D2; t /= D2;
    const int64_t g = t;
    out[idx] = in[a + D0 * (g + D1 * f)];
  }
}

__global__ void env_channels_kernel(const double* __restrict__ src, int64_t Di, int w, int64_t Dc,
                                    const int* __restrict__ klist, int nk, double* __restrict__ dst) {
  const int64_t n = Di * nk * Dc;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t i = t % Di; t /= Di;
    const int k = (int)(t % nk); t /= nk;
    const int64_t c = t;
    dst[idx] = src[i + Di * (klist[k] + (int64_t)w * c)];
  }
}

__global__ void tl_mid_kernel(const double* __restrict__ t1, const double* __restrict__ O,
                              int64_t Dm, int wl, int d, int wr, int64_t Dc, double* __restrict__ t2) {
  const int64_t n = Dm * d * wr * Dc;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t m = t % Dm; t /= Dm;
    const int k = (int)(t % d); t /= d;
    const int b = (int)(t % wr); t /= wr;
    const int64_t c = t;
    double s = 0.0;
    for (int h = 0; h < d; ++h)
      for (int g = 0; g < wl; ++g) {
        const double ov = O[g + wl * (k + d * (h + d * b))];
        if (ov != 0.0) s += t1[m + Dm * (g + (int64_t)wl * (h + d * c))] * ov;
      }
    t2[idx] = s;
  }
}

__global__ void tr_mid_kernel(const double* __restrict__ t1, const double* __restrict__ O,
                              int64_t Da, int wl, int d, int wr, int64_t Do, double* __restrict__ t2) {
  const int64_t n = Da * wl * d * Do;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % Da; t /= Da;
    const int b = (int)(t % wl); t /= wl;
    const int k = (int)(t % d); t /= d;
    const int64_t o = t;
    double s = 0.0;
    for (int l = 0; l < wr; ++l)
      for (int h = 0; h < d; ++h) {
        const double ov = O[b + wl * (k + d * (h + d * l))];
        if (ov != 0.0) s += t1[a + Da * (h + d * (l + (int64_t)wr * o))] * ov;
      }
    t2[idx] = s;
  }
}

// out[a,q,p,b] = in[a,p,q,b]: the two middle legs of a column-major
// n0 x n1 x n2 x n3 tensor exchanged (reads and writes run along a).
__global__ void swap_mid_kernel(const double* __restrict__ in, int64_t n0, int64_t n1, int64_t n2, int64_t n3,
                                double* __restrict__ out) {
  const int64_t n = n0 * n1 * n2 * n3;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % n0; t /= n0;
    const int64_t q = t % n2; t /= n2;
    const int64_t p = t % n1;
    const int64_t b = t / n1;
    out[idx] = in[a + n0 * (p + n1 * (q + n2 * b))];
  }
}

// An MPO tensor widened by identities on a virtual and a physical passenger
// leg: out[(w,e),(k,q),(h,r),(v,f)] = O[w,k,h,v] delta(e,f) delta(q,r) with
// the fused indices w + wl e, k + d q (column-major fusion of the legs).
// O == nullptr stands for the identity MPO O[w,k,h,v] = delta(w,v) delta(k,h)
// (wl == wr).
__global__ void mpo_extend_kernel(const double* __restrict__ O, int wl, int d, int wr, int ne, int nq,
                                  double* __restrict__ out) {
  const int64_t WL = (int64_t)wl * ne, DQ = (int64_t)d * nq, WR = (int64_t)wr * ne;
  const int64_t n = WL * DQ * DQ * WR;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int i0 = (int)(t % WL); t /= WL;
    const int i1 = (int)(t % DQ); t /= DQ;
    const int i2 = (int)(t % DQ);
    const int i3 = (int)(t / DQ);
    const int w = i0 % wl, e = i0 / wl, k = i1 % d, q = i1 / d;
    const int h = i2 % d, r = i2 / d, v = i3 % wr, f = i3 / wr;
    double val = 0.0;
    if (e == f && q == r)
      val = O ? O[w + (int64_t)wl * (k + (int64_t)d * (h + (int64_t)d * v))] : (w == v && k == h ? 1.0 : 0.0);
    out[idx] = val;
  }
}

// The MPO step of the left transfer written straight into the K-contiguous
// mixed-precision operand of the second GEMM (one pass instead of a float64
// round trip and a conversion):
//   t2[(m,k),(b,c)] = sum_{g,h} t1[m + Dm (g + wl (h + d c))] O[g + wl (k + d (h + d b))]
// stored out[(b + wr c) * ldo + m + Dm k], zero for Dm d <= kk < ldo.
template <bool SPLIT>
__global__ void tl_mid_mixed_kernel(const double* __restrict__ t1, const double* __restrict__ O, int64_t Dm,
                                    int wl, int d, int wr, int64_t Dc, int64_t ldo, void* __restrict__ hi,
                                    void* __restrict__ lo) {
  const int64_t n = ldo * wr * Dc, K = Dm * d;
  GRID_STRIDE(idx, n) {
    const int64_t kk = idx % ldo, col = idx / ldo;
    const int b = (int)(col % wr);
    const int64_t c = col / wr;
    double s = 0.0;
    if (kk < K) {
      const int64_t m = kk % Dm;
      const int k = (int)(kk / Dm);
      for (int h = 0; h < d; ++h)
        for (int g = 0; g < wl; ++g) {
          const double ov = O[g + wl * (k + d * (h + d * b))];
          if (ov != 0.0) s += t1[m + Dm * (g + (int64_t)wl * (h + d * c))] * ov;
        }
    }
    if constexpr (SPLIT) {
      __hip_bfloat16 h, l;
      split2(s, h, l);
      static_cast<__hip_bfloat16*>(hi)[idx] = h;
      static_cast<__hip_bfloat16*>(lo)[idx] = l;
    } else {
      static_cast<float*>(hi)[idx] = (float)s;
    }
  }
}

// The MPO step of the right transfer written straight into the K-contiguous
// mixed-precision left operand of the second GEMM (64 x 64 LDS transpose tiles:
// reads run along a, writes along (k, o)):
//   t2[(a,b),(k,o)] = sum_{l,h} t1[a + Da (h + d (l + wr o))] O[b + wl (k + d (h + d l))]
// stored out[(a + Da b) * ldo + k + d o], zero for d Do <= kk < ldo.
template <bool SPLIT>
__global__ __launch_bounds__(256) void tr_mid_mixed_kernel(const double* __restrict__ t1,
                                                           const double* __restrict__ O, int64_t Da, int wl, int d,
                                                           int wr, int64_t Do, int64_t ldo, void* __restrict__ hi,
                                                           void* __restrict__ lo) {
  __shared__ double tile[64][65];
  const int64_t rows = Da * wl, K = (int64_t)d * Do;
  const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int cc = ty; cc < 64; cc += 4) {
    const int64_t r = r0 + tx, kk = c0 + cc;
    double s = 0.0;
    if (r < rows && kk < K) {
      const int64_t a = r % Da;
      const int b = (int)(r / Da);
      const int k = (int)(kk % d);
      const int64_t o = kk / d;
      for (int l = 0; l < wr; ++l)
        for (int h = 0; h < d; ++h) {
          const double ov = O[b + wl * (k + d * (h + d * l))];
          if (ov != 0.0) s += t1[a + Da * (h + d * (l + (int64_t)wr * o))] * ov;
        }
    }
    tile[cc][tx] = s;
  }
  __syncthreads();
  for (int rr = ty; rr < 64; rr += 4) {
    const int64_t r = r0 + rr, kk = c0 + tx;
    if (r < rows && kk < ldo) {
      const double v = tile[tx][rr];
      if constexpr (SPLIT) {
        __hip_bfloat16 h, l;
        split2(v, h, l);
        static_cast<__hip_bfloat16*>(hi)[r * ldo + kk] = h;
        static_cast<__hip_bfloat16*>(lo)[r * ldo + kk] = l;
      } else {
        static_cast<float*>(hi)[r * ldo + kk] = (float)v;
      }
    }
  }
}

__global__ void phys2_kernel(const double* __restrict__ P, const double* __restrict__ x, int64_t Dl,
                             int d1, int64_t Dr, int d2, double alpha, double beta,
                             double* __restrict__ y) {
  const int64_t n = Dl * d1 * Dr * d2;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % Dl; t /= Dl;
    const int b = (int)(t % d1); t /= d1;
    const int64_t c = t % Dr; t /= Dr;
    const int e = (int)t;
    double s = 0.0;
    for (int j = 0; j < d2; ++j)
      for (int h = 0; h < d1; ++h)
        s += P[b + d1 * (e + d2 * (h + d1 * j))] * x[a + Dl * (h + d1 * (c + Dr * j))];
    double v = alpha * s;
    if (beta != 0.0) v += beta * y[idx];
    y[idx] = v;
  }
}

// Same contraction with one thread per (a, c): the d1*d2 inputs x[a,:,c,:] are
// read once into registers and all d1*d2 outputs y[a,:,c,:] written from them,
// so x is read once and y written once (loads and stores coalesced along a).
template <int D12>
__global__ void phys2_reg_kernel(const double* __restrict__ P, const double* __restrict__ x, int64_t Dl,
                                 int d1, int64_t Dr, int d2, double alpha, double beta,
                                 double* __restrict__ y) {
  const int64_t n = Dl * Dr;
  const int64_t sx = Dl * d1 * Dr;   // stride of the second physical leg
  GRID_STRIDE(idx, n) {
    const int64_t a = idx % Dl, c = idx / Dl;
    const int64_t base = a + Dl * d1 * c;
    double xv[D12];
#pragma unroll
    for (int t = 0; t < D12; ++t) {
      const int h = t % d1, j = t / d1;
      xv[t] = x[base + Dl * h + sx * j];
    }
#pragma unroll
    for (int o = 0; o < D12; ++o) {
      const int b = o % d1, e = o / d1;
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < D12; ++t) s += P[b + d1 * (e + d2 * t)] * xv[t];
      double* yp = y + base + Dl * b + sx * e;
      double v = alpha * s;
      if (beta != 0.0) v += beta * *yp;
      *yp = v;
    }
  }
}

// Diagonal of the nearest-neighbour two-site Jordan operator x*IC + BE*x + CB(x):
// diag[row + M col] = IC[col, col] + BE[row, row] + CB[b, e, b, e], row = a + Dl b,
// col = c + Dr e (absent terms are null).
__global__ void diag_nn_kernel(const double* __restrict__ IC, const double* __restrict__ BE,
                               const double* __restrict__ P, int64_t Dl, int64_t Dr, int d1, int d2,
                               double* __restrict__ out) {
  const int64_t M = Dl * d1, N = Dr * d2;
  GRID_STRIDE(idx, M * N) {
    const int64_t row = idx % M, col = idx / M;
    double v = 0.0;
    if (IC) v += IC[col + N * col];
    if (BE) v += BE[row + M * row];
    if (P) {
      const int b = (int)(row / Dl), e = (int)(col / Dr);
      v += P[b + d1 * (e + d2 * (b + d1 * e))];
    }
    out[idx] = v;
  }
}

__global__ void phys1_kernel(const double* __restrict__ P, const double* __restrict__ x, int64_t Dl,
                             int d, int64_t Dr, double alpha, double beta, double* __restrict__ y) {
  const int64_t n = Dl * d * Dr;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % Dl; t /= Dl;
    const int b = (int)(t % d); t /= d;
    const int64_t c = t;
    double s = 0.0;
    for (int h = 0; h < d; ++h) s += P[b + d * h] * x[a + Dl * (h + (int64_t)d * c)];
    double v = alpha * s;
    if (beta != 0.0) v += beta * y[idx];
    y[idx] = v;
  }
}

// ---- reductions -------------------------------------------------------------

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// lam[g] = sum_{i<D2, j<D1} lvec[i + D2*j] * v[j + D1*(g + w*i)]
__global__ void reg_lambda_kernel(const double* __restrict__ v, const double* __restrict__ lvec,
                                  int64_t D1, int w, int64_t D2, double* __restrict__ lam) {
  __shared__ double sh[TPB / 64];
  const int g = blockIdx.x;
  double s = 0.0;
  const int64_t n = D1 * D2;
  for (int64_t idx = threadIdx.x; idx < n; idx += blockDim.x) {
    const int64_t j = idx % D1, i = idx / D1;
    s += lvec[i + D2 * j] * v[j + D1 * (g + (int64_t)w * i)];
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) lam[g] = s;
}

__global__ void reg_apply_kernel(double* __restrict__ v, const double* __restrict__ rvec,
                                 const double* __restrict__ lam, int64_t D1, int w, int64_t D2) {
  const int64_t n = D1 * w * D2;
  GRID_STRIDE(idx, n) {
    int64_t t = idx;
    const int64_t a = t % D1; t /= D1;
    const int g = (int)(t % w); t /= w;
    const int64_t c = t;
    v[idx] -= lam[g] * rvec[a + D1 * c];
  }
}

constexpr int MD_ITEMS = 16;  // elements of w held per thread in multidot

__global__ void multidot_kernel(const double* __restrict__ V, int64_t ldv, int k,
                                const double* __restrict__ w, int64_t n, double* __restrict__ partial) {
  __shared__ double sh[TPB / 64];
  const int64_t base = (int64_t)blockIdx.x * TPB * MD_ITEMS;
  double wr[MD_ITEMS];
#pragma unroll
  for (int it = 0; it < MD_ITEMS; ++it) {
    const int64_t i = base + it * TPB + threadIdx.x;
    wr[it] = i < n ? w[i] : 0.0;
  }
  for (int j = 0; j < k; ++j) {
    const double* Vj = V + (int64_t)j * ldv;
    double s = 0.0;
#pragma unroll
    for (int it = 0; it < MD_ITEMS; ++it) {
      const int64_t i = base + it * TPB + threadIdx.x;
      if (i < n) s += Vj[i] * wr[it];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void reduce_rows_kernel(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double sh[TPB / 64];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[(int64_t)j * nb + b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[j] = s;
}

// Fused classical Gram-Schmidt update: u <- u - V c (c = cin[0..k)) and, with
// the updated u, the partial sums of V^T u (MODE 0, k outputs) or of u^T u and
// w^T u (MODE 1, 2 outputs).  Each thread keeps its row of V in registers
// between the update and the dot products, so V is read once for both.
// k <= CGS_KMAX; partial sums go to partial[o * gridDim.x + blockIdx.x].
// KT (4, 8, 16 or 32 >= k) sizes the register arrays: the rows of V held per
// thread, and the accumulators, scale with the basis, so small bases keep
// the occupancy (and the memory parallelism) high.  MODE 2: no update, the
// partial sums of V^T u only (the first Gram-Schmidt pass).
constexpr int CGS_KMAX = 32;
template <int MODE, int KT, typename VT>
__global__ __launch_bounds__(TPB) void cgs_update_kernel(const VT* __restrict__ V, int64_t ldv, int k,
                                                         const double* __restrict__ cin, double* __restrict__ u,
                                                         const double* __restrict__ w, int64_t n,
                                                         double* __restrict__ partial) {
  __shared__ double sh[TPB / 64];
  constexpr int NO = MODE == 1 ? 2 : KT;
  double c[KT], acc[NO];
#pragma unroll
  for (int j = 0; j < KT; ++j) c[j] = (MODE != 2 && j < k) ? cin[j] : 0.0;
#pragma unroll
  for (int o = 0; o < NO; ++o) acc[o] = 0.0;
  GRID_STRIDE(r, n) {
    double q[KT];
#pragma unroll
    for (int j = 0; j < KT; ++j) q[j] = j < k ? (double)V[r + (int64_t)j * ldv] : 0.0;
    double v = u[r];
    if constexpr (MODE != 2) {
#pragma unroll
      for (int j = 0; j < KT; ++j) v -= c[j] * q[j];
      u[r] = v;
    }
    if constexpr (MODE != 1) {
#pragma unroll
      for (int j = 0; j < KT; ++j) acc[j] += q[j] * v;
    } else {
      acc[0] += v * v;
      acc[1] += w[r] * v;
    }
  }
  const int nout = MODE == 1 ? 2 : k;
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    if (o >= nout) break;
    const double t = block_sum(acc[o], sh);
    if (threadIdx.x == 0) partial[(int64_t)o * gridDim.x + blockIdx.x] = t;
  }
}

__global__ void multiaxpy_kernel(const double* __restrict__ V, int64_t ldv, int k,
                                 const double* __restrict__ c, double* __restrict__ w, int64_t n) {
  GRID_STRIDE(i, n) {
    double s = w[i];
    for (int j = 0; j < k; ++j) s -= c[j] * V[i + (int64_t)j * ldv];
    w[i] = s;
  }
}

__global__ void gemv_kernel(const double* __restrict__ V, int64_t ldv, int k,
                            const double* __restrict__ s, double* __restrict__ x, int64_t n) {
  GRID_STRIDE(i, n) {
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += s[j] * V[i + (int64_t)j * ldv];
    x[i] = acc;
  }
}

__global__ void axpby_kernel(double a, const double* x, double b, double* y,
                             int64_t n) {
  GRID_STRIDE(i, n) {
    double v = a * x[i];
    if (b != 0.0) v += b * y[i];
    y[i] = v;
  }
}

__global__ void scale_dev_kernel(double* __restrict__ y, const double* __restrict__ sdev, int inv,
                                 int64_t n) {
  const double s = inv ? 1.0 / *sdev : *sdev;
  GRID_STRIDE(i, n) y[i] *= s;
}

__global__ void nonfinite_kernel(const double* __restrict__ x, int64_t n, int* flag) {
  int bad = 0;
  GRID_STRIDE(i, n) {
    if (!isfinite(x[i])) bad = 1;
  }
  if (bad) atomicOr(flag, 1);
}

// ---- precision conversion ---------------------------------------------------

// plain conversion (no transposition): out[c*ldo + r] = M[r + c*ld], zero for rows <= r < ldo
__global__ void split_plain_kernel(const double* __restrict__ M, int64_t rows, int64_t cols, int64_t ld,
                                   int64_t ldo, __hip_bfloat16* __restrict__ hi,
                                   __hip_bfloat16* __restrict__ lo) {
  const int64_t n = ldo * cols;
  GRID_STRIDE(idx, n) {
    const int64_t r = idx % ldo, c = idx / ldo;
    __hip_bfloat16 h = __float2bfloat16(0.f), l = __float2bfloat16(0.f);
    if (r < rows) split2(M[r + c * ld], h, l);
    hi[idx] = h;
    lo[idx] = l;
  }
}

// transposed conversion: out[r*ldo + c] = M[r + c*ld] (zero for cols <= c < ldo), 64x64 LDS tiles
__global__ void split_trans_kernel(const double* __restrict__ M, int64_t rows, int64_t cols, int64_t ld,
                                   int64_t ldo, __hip_bfloat16* __restrict__ hi,
                                   __hip_bfloat16* __restrict__ lo) {
  __shared__ double tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 4 rows of threads
  for (int cc = ty; cc < 64; cc += 4) {
    const int64_t r = r0 + tx, c = c0 + cc;
    tile[cc][tx] = (r < rows && c < cols) ? M[r + c * ld] : 0.0;
  }
  __syncthreads();
  for (int rr = ty; rr < 64; rr += 4) {
    const int64_t r = r0 + rr, c = c0 + tx;
    if (r < rows && c < ldo) {
      __hip_bfloat16 h, l;
      split2(tile[tx][rr], h, l);
      hi[r * ldo + c] = h;
      lo[r * ldo + c] = l;
    }
  }
}

__global__ void f32_plain_kernel(const double* __restrict__ M, int64_t rows, int64_t cols, int64_t ld,
                                 int64_t ldo, float* __restrict__ out) {
  const int64_t n = ldo * cols;
  GRID_STRIDE(idx, n) {
    const int64_t r = idx % ldo, c = idx / ldo;
    out[idx] = r < rows ? (float)M[r + c * ld] : 0.f;
  }
}

__global__ void f32_trans_kernel(const double* __restrict__ M, int64_t rows, int64_t cols, int64_t ld,
                                 int64_t ldo, float* __restrict__ out) {
  __shared__ double tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int cc = ty; cc < 64; cc += 4) {
    const int64_t r = r0 + tx, c = c0 + cc;
    tile[cc][tx] = (r < rows && c < cols) ? M[r + c * ld] : 0.0;
  }
  __syncthreads();
  for (int rr = ty; rr < 64; rr += 4) {
    const int64_t r = r0 + rr, c = c0 + tx;
    if (r < rows && c < ldo) out[r * ldo + c] = (float)tile[tx][rr];
  }
}

template <bool SPLIT, bool PACK>
__global__ __launch_bounds__(256) void nn_prologue_kernel(const double* __restrict__ P,
                                                          const double* __restrict__ x, int64_t Dl, int64_t Dr,
                                                          double alpha, double beta, double* __restrict__ y,
                                                          void* __restrict__ xa_hi, void* __restrict__ xa_lo,
                                                          int64_t ka, void* __restrict__ xb_hi,
                                                          void* __restrict__ xb_lo, int64_t kb) {
  __shared__ ProTile t;
  nn_prologue_tile<SPLIT, PACK>(P, x, Dl, Dr, alpha, beta, y, xa_hi, xa_lo, ka, xb_hi, xb_lo, kb,
                                (int64_t)blockIdx.x * PRO_A, (int64_t)blockIdx.y * PRO_C, t);
}

// Packed panel layout from K-contiguous rows (row length kpad >= K, element
// size eb): one 16-byte chunk per thread.
__global__ void repack_kernel(const char* __restrict__ in, int64_t MN, int64_t K, int64_t kpad, int eb,
                              char* __restrict__ out) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(1))) u4 gu4;
  const int64_t steps = K * eb / 64;
  GRID_STRIDE(o, MN * K * eb / 16) {
    const int64_t slab = o / 1024, r = (o % 1024) / 4, q = o % 4;
    const int64_t row = (slab / steps) * 256 + r, kbyte = (slab % steps) * 64 + q * 16;
    ((gu4*)out)[o] = ((const gu4*)(in + row * kpad * eb + kbyte))[0];
  }
}

// ---- Jordan operator matrices from constructed terms (prepare_operator!! folds)
// Absent terms are null pointers; flags say which optional terms a target takes.

__global__ void asm_ic_kernel(const double* __restrict__ IC, const double* __restrict__ ID,
                              const double* __restrict__ II, int64_t Dr, int d2, double* __restrict__ out) {
  const int64_t N = Dr * d2;
  GRID_STRIDE(idx, N * N) {
    const int64_t row = idx % N, col = idx / N;
    const int64_t i = row % Dr, c = col % Dr;
    const int j = (int)(row / Dr), e = (int)(col / Dr);
    double v = IC ? IC[e + d2 * (c + Dr * (j + d2 * i))] : 0.0;
    if (ID && c == i) v += ID[e + d2 * j];
    if (II && e == j) v += II[c + Dr * i];
    out[idx] = v;
  }
}

__global__ void asm_be_kernel(const double* __restrict__ BE, const double* __restrict__ DE,
                              const double* __restrict__ EE, int64_t Dl, int d1, double* __restrict__ out) {
  const int64_t M = Dl * d1;
  GRID_STRIDE(idx, M * M) {
    const int64_t row = idx % M, col = idx / M;
    const int64_t a = row % Dl, f = col % Dl;
    const int b = (int)(row / Dl), h = (int)(col / Dl);
    double v = BE ? BE[idx] : 0.0;
    if (DE && a == f) v += DE[b + d1 * h];
    if (EE && b == h) v += EE[a + Dl * f];
    out[idx] = v;
  }
}

__global__ void asm_ca_kernel(const double* __restrict__ CA, const double* __restrict__ CB,
                              const double* __restrict__ IC, const double* __restrict__ ID,
                              const double* __restrict__ II, int64_t Dr, int d1, int d2, double* __restrict__ out) {
  const int64_t N = Dr * d2, R = d1 * N;
  GRID_STRIDE(idx, R * R) {
    int64_t t = idx % R;
    const int h = (int)(t % d1); t /= d1;
    const int64_t i = t % Dr;
    const int j = (int)(t / Dr);
    t = idx / R;
    const int b = (int)(t % d1); t /= d1;
    const int64_t c = t % Dr;
    const int e = (int)(t / Dr);
    double v = CA[b + d1 * (e + d2 * (c + Dr * (h + d1 * (j + d2 * i))))];
    if (CB && c == i) v += CB[b + d1 * (e + d2 * (h + d1 * j))];
    if (b == h) {
      if (IC) v += IC[e + d2 * (c + Dr * (j + d2 * i))];
      if (ID && c == i) v += ID[e + d2 * j];
      if (II && e == j) v += II[c + Dr * i];
    }
    out[idx] = v;
  }
}

__global__ void asm_ab_kernel(const double* __restrict__ AB, const double* __restrict__ CB,
                              const double* __restrict__ BE, const double* __restrict__ DE,
                              const double* __restrict__ EE, int64_t Dl, int d1, int d2, double* __restrict__ out) {
  const int64_t M = Dl * d1, MM = M * M;
  GRID_STRIDE(idx, MM * d2 * d2) {
    const int64_t blk = idx / MM, w = idx % MM;
    const int e = (int)(blk % d2), j = (int)(blk / d2);
    const int64_t row = w % M, col = w / M;
    const int64_t a = row % Dl, f = col % Dl;
    const int b = (int)(row / Dl), h = (int)(col / Dl);
    double v = AB[a + Dl * (b + d1 * (e + d2 * (f + Dl * (h + d1 * (int64_t)j))))];
    if (CB && a == f) v += CB[b + d1 * (e + d2 * (h + d1 * j))];
    if (e == j) {
      if (BE) v += BE[w];
      if (DE && a == f) v += DE[b + d1 * h];
      if (EE && b == h) v += EE[a + Dl * f];
    }
    out[idx] = v;
  }
}

__global__ void asm_cp_kernel(const double* __restrict__ C, const double* __restrict__ D,
                              const double* __restrict__ I, int64_t Dr, int d, double* __restrict__ out) {
  const int64_t N = Dr * d;
  GRID_STRIDE(idx, N * N) {
    const int64_t row = idx % N, col = idx / N;
    const int h = (int)(row % d), b = (int)(col % d);
    const int64_t i = row / d, c = col / d;
    double v = C ? C[b + d * (c + Dr * (h + d * i))] : 0.0;
    if (D && i == c) v += D[b + d * h];
    if (I && b == h) v += I[i + Dr * c];
    out[idx] = v;
  }
}

}  // namespace

#define LAUNCH_1D(kern, n, st, ...)                                          \
  do {                                                                        \
    if ((n) > 0) {                                                            \
      hipLaunchKernelGGL(kern, grid1d(n, 4), dim3(TPB), 0, st, __VA_ARGS__);  \
      MK_CHECK_HIP(hipGetLastError());                                        \
    }                                                                         \
    return MK_OK;                                                             \
  } while (0)

mk_status k_prep_L(const double* GL, const double* W, const int* klist, int nk, int64_t Da,
                   int64_t Df, int wl, int d, int wr, int g_ident, double* L, hipStream_t st) {
  LAUNCH_1D(prep_L_kernel, Da * d * Df * d * nk, st, GL, W, klist, nk, Da, Df, wl, d, wr, g_ident, L);
}
mk_status k_prep_R(const double* GR, const double* W, const int* klist, int nk, int64_t Di,
                   int64_t Dc, int wl, int d, int wr, int l_ident, double* R, hipStream_t st) {
  LAUNCH_1D(prep_R_kernel, Di * d * nk * Dc * d, st, GR, W, klist, nk, Di, Dc, wl, d, wr, l_ident, R);
}
mk_status k_prep_Cp(const double* GR, const double* W, int64_t Di, int64_t Dc, int wl, int d,
                    int wr, int l_ident, double* Cp, hipStream_t st) {
  LAUNCH_1D(prep_Cp_kernel, (int64_t)d * Di * d * Dc, st, GR, W, Di, Dc, wl, d, wr, l_ident, Cp);
}
mk_status k_prep_CA(const double* GR, const double* W1, const double* W2, int64_t Dr, int wl,
                    int d1, int wm, int d2, int wr, int l_ident, double* CA, hipStream_t st) {
  LAUNCH_1D(prep_CA_kernel, (int64_t)d1 * Dr * d2 * d1 * Dr * d2, st, GR, W1, W2, Dr, wl, d1, wm, d2,
            wr, l_ident, CA);
}
mk_status k_prep_AB(const double* GL, const double* W1, const double* W2, int64_t Dl, int wl,
                    int d1, int wm, int d2, int wr, int g_ident, double* AB, hipStream_t st) {
  LAUNCH_1D(prep_AB_kernel, Dl * d1 * Dl * d1 * d2 * d2, st, GL, W1, W2, Dl, wl, d1, wm, d2, wr,
            g_ident, AB);
}
mk_status k_swap23(const double* in, int64_t D0, int64_t D1, int64_t D2, double* out, hipStream_t st) {
  LAUNCH_1D(swap23_kernel, D0 * D1 * D2, st, in, D0, D1, D2, out);
}
mk_status k_env_channels(const double* src, int64_t Di, int w, int64_t Dc, const int* klist, int nk,
                         double* dst, hipStream_t st) {
  LAUNCH_1D(env_channels_kernel, Di * nk * Dc, st, src, Di, w, Dc, klist, nk, dst);
}
mk_status k_tl_mid(const double* t1, const double* O, int64_t Dm, int wl, int d, int wr, int64_t Dc,
                   double* t2, hipStream_t st) {
  LAUNCH_1D(tl_mid_kernel, Dm * d * wr * Dc, st, t1, O, Dm, wl, d, wr, Dc, t2);
}
mk_status k_tr_mid(const double* t1, const double* O, int64_t Da, int wl, int d, int wr, int64_t Do,
                   double* t2, hipStream_t st) {
  LAUNCH_1D(tr_mid_kernel, Da * wl * d * Do, st, t1, O, Da, wl, d, wr, Do, t2);
}
mk_status k_swap_mid(const double* in, int64_t n0, int64_t n1, int64_t n2, int64_t n3, double* out,
                     hipStream_t st) {
  LAUNCH_1D(swap_mid_kernel, n0 * n1 * n2 * n3, st, in, n0, n1, n2, n3, out);
}
mk_status k_mpo_extend(const double* O, int wl, int d, int wr, int ne, int nq, double* out, hipStream_t st) {
  LAUNCH_1D(mpo_extend_kernel, (int64_t)wl * ne * d * nq * d * nq * wr * ne, st, O, wl, d, wr, ne, nq, out);
}
mk_status k_tl_mid_mixed(const double* t1, const double* O, int64_t Dm, int wl, int d, int wr, int64_t Dc,
                         int64_t ldo, int split, void* hi, void* lo, hipStream_t st) {
  if (split) LAUNCH_1D(tl_mid_mixed_kernel<true>, ldo * wr * Dc, st, t1, O, Dm, wl, d, wr, Dc, ldo, hi, lo);
  LAUNCH_1D(tl_mid_mixed_kernel<false>, ldo * wr * Dc, st, t1, O, Dm, wl, d, wr, Dc, ldo, hi, lo);
}
mk_status k_tr_mid_mixed(const double* t1, const double* O, int64_t Da, int wl, int d, int wr, int64_t Do,
                         int64_t ldo, int split, void* hi, void* lo, hipStream_t st) {
  if (Da * wl * ldo == 0) return MK_OK;
  dim3 g((unsigned)cdiv(Da * wl, (int64_t)64), (unsigned)cdiv(ldo, (int64_t)64));
  if (split)
    hipLaunchKernelGGL(tr_mid_mixed_kernel<true>, g, dim3(256), 0, st, t1, O, Da, wl, d, wr, Do, ldo, hi, lo);
  else
    hipLaunchKernelGGL(tr_mid_mixed_kernel<false>, g, dim3(256), 0, st, t1, O, Da, wl, d, wr, Do, ldo, hi, lo);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}
mk_status k_phys2(const double* P, const double* x, int64_t Dl, int d1, int64_t Dr, int d2,
                  double alpha, double beta, double* y, hipStream_t st) {
  switch (d1 * d2) {
    case 4: LAUNCH_1D(phys2_reg_kernel<4>, Dl * Dr, st, P, x, Dl, d1, Dr, d2, alpha, beta, y);
    case 9: LAUNCH_1D(phys2_reg_kernel<9>, Dl * Dr, st, P, x, Dl, d1, Dr, d2, alpha, beta, y);
    case 16: LAUNCH_1D(phys2_reg_kernel<16>, Dl * Dr, st, P, x, Dl, d1, Dr, d2, alpha, beta, y);
    default: LAUNCH_1D(phys2_kernel, Dl * d1 * Dr * d2, st, P, x, Dl, d1, Dr, d2, alpha, beta, y);
  }
}
mk_status k_diag_nn(const double* IC, const double* BE, const double* P, int64_t Dl, int64_t Dr, int d1, int d2,
                    double* out, hipStream_t st) {
  LAUNCH_1D(diag_nn_kernel, Dl * d1 * Dr * d2, st, IC, BE, P, Dl, Dr, d1, d2, out);
}
mk_status k_phys1(const double* P, const double* x, int64_t Dl, int d, int64_t Dr, double alpha,
                  double beta, double* y, hipStream_t st) {
  LAUNCH_1D(phys1_kernel, Dl * d * Dr, st, P, x, Dl, d, Dr, alpha, beta, y);
}

mk_status k_regularize(double* v, const double* lvec, const double* rvec, int64_t D1, int w, int64_t D2,
                       double* lam, hipStream_t st) {
  if (D1 * D2 * w == 0) return MK_OK;
  hipLaunchKernelGGL(reg_lambda_kernel, dim3(w), dim3(TPB), 0, st, v, lvec, D1, w, D2, lam);
  MK_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(reg_apply_kernel, grid1d(D1 * w * D2, 4), dim3(TPB), 0, st, v, rvec, lam, D1, w, D2);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

int multidot_blocks(int64_t n) { return (int)std::max<int64_t>(1, cdiv(n, (int64_t)TPB * MD_ITEMS)); }

int cgs_blocks(int64_t n);

template <int MODE, typename VT>
void launch_cgs(int k, int nb, hipStream_t st, const VT* V, int64_t ldv, const double* cin, double* u,
                const double* w, int64_t n, double* partial) {
  if (k <= 4)
    hipLaunchKernelGGL((cgs_update_kernel<MODE, 4, VT>), dim3(nb), dim3(TPB), 0, st, V, ldv, k, cin, u, w, n,
                       partial);
  else if (k <= 8)
    hipLaunchKernelGGL((cgs_update_kernel<MODE, 8, VT>), dim3(nb), dim3(TPB), 0, st, V, ldv, k, cin, u, w, n,
                       partial);
  else if (k <= 16)
    hipLaunchKernelGGL((cgs_update_kernel<MODE, 16, VT>), dim3(nb), dim3(TPB), 0, st, V, ldv, k, cin, u, w, n,
                       partial);
  else
    hipLaunchKernelGGL((cgs_update_kernel<MODE, 32, VT>), dim3(nb), dim3(TPB), 0, st, V, ldv, k, cin, u, w, n,
                       partial);
}

mk_status k_multidot(const double* V, int64_t ldv, int k, const double* w, int64_t n, double* out,
                     double* partial, hipStream_t st) {
  if (k <= 0) return MK_OK;
  if (k <= CGS_KMAX) {
    // one read of V with register accumulators; at most multidot_blocks(n)
    // blocks, so callers' partial buffers sized for the wide kernel suffice
    const int nb = std::min(cgs_blocks(n), multidot_blocks(n));
    launch_cgs<2, double>(k, nb, st, V, ldv, nullptr, const_cast<double*>(w), nullptr, n, partial);
    MK_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(k), dim3(TPB), 0, st, partial, nb, out);
    MK_CHECK_HIP(hipGetLastError());
    return MK_OK;
  }
  const int nb = multidot_blocks(n);
  hipLaunchKernelGGL(multidot_kernel, dim3(nb), dim3(TPB), 0, st, V, ldv, k, w, n, partial);
  MK_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(k), dim3(TPB), 0, st, partial, nb, out);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}
int cgs_blocks(int64_t n) { return (int)std::min<int64_t>(1024, std::max<int64_t>(1, cdiv(n, (int64_t)TPB * 4))); }

mk_status k_cgs_update_dot(const double* V, int64_t ldv, int k, const double* cin, double* u, int64_t n,
                           double* out, double* partial, hipStream_t st) {
  MK_REQUIRE(k >= 1 && k <= CGS_KMAX, MK_BAD_ARG, "cgs_update_dot: k = %d not in [1, %d]", k, CGS_KMAX);
  const int nb = cgs_blocks(n);
  launch_cgs<0, double>(k, nb, st, V, ldv, cin, u, nullptr, n, partial);
  MK_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(k), dim3(TPB), 0, st, partial, nb, out);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status k_cgs_update_norm(const double* V, int64_t ldv, int k, const double* cin, double* u, const double* w,
                            int64_t n, double* out, double* partial, hipStream_t st) {
  MK_REQUIRE(k >= 1 && k <= CGS_KMAX, MK_BAD_ARG, "cgs_update_norm: k = %d not in [1, %d]", k, CGS_KMAX);
  const int nb = cgs_blocks(n);
  launch_cgs<1, double>(k, nb, st, V, ldv, cin, u, w, n, partial);
  MK_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(2), dim3(TPB), 0, st, partial, nb, out);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status k_multiaxpy(const double* V, int64_t ldv, int k, const double* c, double* w, int64_t n,
                      hipStream_t st) {
  if (k <= 0) return MK_OK;
  LAUNCH_1D(multiaxpy_kernel, n, st, V, ldv, k, c, w, n);
}
mk_status k_gemv(const double* V, int64_t ldv, int k, const double* s, double* x, int64_t n,
                 hipStream_t st) {
  LAUNCH_1D(gemv_kernel, n, st, V, ldv, k, s, x, n);
}

mk_status k_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st) {
  LAUNCH_1D(axpby_kernel, n, st, a, x, b, y, n);
}
mk_status k_scale_dev(double* y, const double* sdev, int inv, int64_t n, hipStream_t st) {
  LAUNCH_1D(scale_dev_kernel, n, st, y, sdev, inv, n);
}
mk_status k_nonfinite(const double* x, int64_t n, int* flag, hipStream_t st) {
  LAUNCH_1D(nonfinite_kernel, n, st, x, n, flag);
}

mk_status k_split_bf16(const double* M, int64_t rows, int64_t cols, int64_t ld, int trans, int64_t ldo,
                       void* hi, void* lo, hipStream_t st) {
  if (rows * cols == 0) return MK_OK;
  auto* h = static_cast<__hip_bfloat16*>(hi);
  auto* l = static_cast<__hip_bfloat16*>(lo);
  if (trans) {
    hipLaunchKernelGGL(split_plain_kernel, grid1d(ldo * cols, 4), dim3(TPB), 0, st, M, rows, cols, ld, ldo, h, l);
  } else {
    dim3 g((unsigned)cdiv(rows, 64), (unsigned)cdiv(ldo, 64));
    hipLaunchKernelGGL(split_trans_kernel, g, dim3(TPB), 0, st, M, rows, cols, ld, ldo, h, l);
  }
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status k_to_f32(const double* M, int64_t rows, int64_t cols, int64_t ld, int trans, int64_t ldo,
                   float* out, hipStream_t st) {
  if (rows * cols == 0) return MK_OK;
  if (trans) {
    hipLaunchKernelGGL(f32_plain_kernel, grid1d(ldo * cols, 4), dim3(TPB), 0, st, M, rows, cols, ld, ldo, out);
  } else {
    dim3 g((unsigned)cdiv(rows, 64), (unsigned)cdiv(ldo, 64));
    hipLaunchKernelGGL(f32_trans_kernel, g, dim3(TPB), 0, st, M, rows, cols, ld, ldo, out);
  }
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status k_nn_prologue(const double* P, const double* x, int64_t Dl, int64_t Dr, double alpha, double beta,
                        double* y, int split, void* xa_hi, void* xa_lo, int64_t ka, void* xb_hi, void* xb_lo,
                        int64_t kb, int packed, hipStream_t st) {
  if (Dl * Dr == 0) return MK_OK;
  dim3 g((unsigned)cdiv(Dl, (int64_t)PRO_A), (unsigned)cdiv(Dr, (int64_t)PRO_C));
  if (packed)
    MK_REQUIRE((2 * Dl) % 256 == 0 && (2 * Dr) % 256 == 0 && ka == 2 * Dr && kb == 2 * Dl, MK_BAD_ARG,
               "nn_prologue: packed views need 2 Dl, 2 Dr multiples of 256 and unpadded rows");
  if (split && packed)
    hipLaunchKernelGGL((nn_prologue_kernel<true, true>), g, dim3(256), 0, st, P, x, Dl, Dr, alpha, beta, y, xa_hi,
                       xa_lo, ka, xb_hi, xb_lo, kb);
  else if (split)
    hipLaunchKernelGGL((nn_prologue_kernel<true, false>), g, dim3(256), 0, st, P, x, Dl, Dr, alpha, beta, y, xa_hi,
                       xa_lo, ka, xb_hi, xb_lo, kb);
  else if (packed)
    hipLaunchKernelGGL((nn_prologue_kernel<false, true>), g, dim3(256), 0, st, P, x, Dl, Dr, alpha, beta, y, xa_hi,
                       nullptr, ka, xb_hi, nullptr, kb);
  else
    hipLaunchKernelGGL((nn_prologue_kernel<false, false>), g, dim3(256), 0, st, P, x, Dl, Dr, alpha, beta, y, xa_hi,
                       nullptr, ka, xb_hi, nullptr, kb);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status k_repack(const void* in, int64_t MN, int64_t K, int64_t kpad, int eb, void* out, hipStream_t st) {
  MK_REQUIRE(MN % 256 == 0 && (K * eb) % 64 == 0 && kpad >= K && (kpad * eb) % 16 == 0 &&
                 (uintptr_t)in % 16 == 0 && (uintptr_t)out % 16 == 0,
             MK_BAD_ARG, "repack: rows %lld, K %lld, kpad %lld do not fit the packed layout", (long long)MN,
             (long long)K, (long long)kpad);
  if (MN * K == 0) return MK_OK;
  LAUNCH_1D(repack_kernel, MN * K * eb / 16, st, static_cast<const char*>(in), MN, K, kpad, eb,
            static_cast<char*>(out));
}


mk_status k_asm_ic(const double* IC, const double* ID, const double* II, int64_t Dr, int d2, double* out,
                   hipStream_t st) {
  LAUNCH_1D(asm_ic_kernel, Dr * d2 * Dr * d2, st, IC, ID, II, Dr, d2, out);
}
mk_status k_asm_be(const double* BE, const double* DE, const double* EE, int64_t Dl, int d1, double* out,
                   hipStream_t st) {
  LAUNCH_1D(asm_be_kernel, Dl * d1 * Dl * d1, st, BE, DE, EE, Dl, d1, out);
}
mk_status k_asm_ca(const double* CA, const double* CB, const double* IC, const double* ID, const double* II,
                   int64_t Dr, int d1, int d2, double* out, hipStream_t st) {
  const int64_t R = (int64_t)d1 * Dr * d2;
  LAUNCH_1D(asm_ca_kernel, R * R, st, CA, CB, IC, ID, II, Dr, d1, d2, out);
}
mk_status k_asm_ab(const double* AB, const double* CB, const double* BE, const double* DE, const double* EE,
                   int64_t Dl, int d1, int d2, double* out, hipStream_t st) {
  const int64_t M = Dl * d1;
  LAUNCH_1D(asm_ab_kernel, M * M * d2 * d2, st, AB, CB, BE, DE, EE, Dl, d1, d2, out);
}
mk_status k_asm_cp(const double* C, const double* D, const double* I, int64_t Dr, int d, double* out,
                   hipStream_t st) {
  LAUNCH_1D(asm_cp_kernel, Dr * d * Dr * d, st, C, D, I, Dr, d, out);
}

}  // namespace mk
