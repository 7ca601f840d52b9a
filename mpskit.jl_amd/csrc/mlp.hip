// Learned Davidson preconditioner: a 3-layer FP32 MLP acting on 256-element
// blocks of the residual, with hand-written inference, training (backward +
// Adam) and feature kernels.
//
// Features of block b (512 floats): the residual block and the diagonally
// preconditioned residual block q = r / (theta - diag H), each taken from a
// globally normalised vector and scaled by sqrt(n / 256) so that block norms
// are O(1).  Output y (256 floats): a multiplicative refinement of the
// diagonal (Davidson) correction, P = q (1 + y) elementwise, rescaled to the
// norm of q in every block; the last layer starts at zero, so the untrained
// network is the diagonal correction.
// Layers: 512 -> 256 (tanh) -> 256 (tanh) -> 256 (linear); parameters are
// row-major (out x in) weights and biases.  Training minimises, over blocks,
// |t*_b| (1 - cos(q_b (1 + y_b), t*_b)) with t* = x* - x (x . x*) the ideal
// correction of a Davidson iterate x (x* the converged eigenvector).
#include <cmath>
#include <vector>

#include "mlp.h"

namespace mk {
namespace {

constexpr int TPB = 256;

// ---- FP32 GEMM: C (M x N, row-major ldc) = alpha op(A) op(B) + beta C --------
// op(A) is M x K: A row-major M x K (lda) or, with ta, A stored K x M.
// op(B) is K x N: B row-major K x N (ldb) or, with tb, B stored N x K.
// 64 x 64 tiles, BK = 16, 256 threads with a 4 x 4 register block each.
constexpr int GB = 64, GK = 16;

template <bool TA, bool TB>
__global__ __launch_bounds__(TPB) void sgemm_kernel(int M, int N, int K, float alpha, const float* __restrict__ A,
                                                    int lda, const float* __restrict__ B, int ldb, float beta,
                                                    float* __restrict__ C, int ldc) {
  __shared__ float As[GK][GB + 4];
  __shared__ float Bs[GK][GB + 4];
  const int m0 = blockIdx.y * GB, n0 = blockIdx.x * GB;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += GK) {
    // 64 x 16 elements of each operand, 4 per thread
    for (int e = threadIdx.x; e < GB * GK; e += TPB) {
      int r, k;
      if (TA) { k = e / GB; r = e % GB; } else { r = e / GK; k = e % GK; }
      const int gm = m0 + r, gk = k0 + k;
      float v = 0.f;
      if (gm < M && gk < K) v = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
      As[k][r] = v;
    }
    for (int e = threadIdx.x; e < GB * GK; e += TPB) {
      int c, k;
      if (TB) { c = e / GK; k = e % GK; } else { k = e / GB; c = e % GB; }
      const int gn = n0 + c, gk = k0 + k;
      float v = 0.f;
      if (gn < N && gk < K) v = TB ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
      Bs[k][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tx * 4 + j;
      if (gn >= N) continue;
      float* cp = C + (int64_t)gm * ldc + gn;
      const float v = alpha * acc[i][j];
      *cp = beta == 0.f ? v : v + beta * *cp;
    }
  }
}

// Y = tanh(Y + b) (act = 1) or Y + b (act = 0), Y is rows x cols
__global__ void bias_act_kernel(float* Y, const float* __restrict__ b, int64_t rows, int cols, int act) {
  const int64_t n = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    const float v = Y[i] + b[i % cols];
    Y[i] = act ? tanhf(v) : v;
  }
}

// dZ = dY * (1 - Y^2) in place on dY
__global__ void tanh_bwd_kernel(float* dY, const float* __restrict__ Y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    dY[i] *= 1.f - Y[i] * Y[i];
}

// db[j] = sum_r dZ[r, j]  (one block per column)
__global__ void colsum_kernel(const float* __restrict__ dZ, int64_t rows, int cols, float* __restrict__ db) {
  __shared__ float sh[TPB];
  const int j = blockIdx.x;
  float s = 0.f;
  for (int64_t r = threadIdx.x; r < rows; r += TPB) s += dZ[r * cols + j];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = TPB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = sh[0];
}

// Loss of one 256-element row (one block per row, one thread per element):
// prediction P = q (1 + Y), q = X[:, 256:512] the diagonally preconditioned
// residual; loss_r = w (1 - cos(P, T)) with w = |T|, so that rows with a large
// ideal correction weigh more and the loss is blind to each row's scale (the
// Davidson step only uses directions).  dY = dloss/dY / rows, part[r] = loss_r.
__global__ __launch_bounds__(MLP_OUT) void cos_loss_kernel(const float* __restrict__ Y, const float* __restrict__ X,
                                                           const float* __restrict__ T, int64_t rows,
                                                           float* __restrict__ dY, float* __restrict__ part) {
  __shared__ float s3[3][MLP_OUT / 64];
  const int64_t r = blockIdx.x;
  const int k = threadIdx.x;
  const float q = X[r * MLP_IN + MLP_BLOCK + k];
  const float y = Y[r * MLP_OUT + k];
  const float t = T[r * MLP_OUT + k];
  const float p = q * (1.f + y);
  float v[3] = {p * p, p * t, t * t};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float a = v[c];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if ((k & 63) == 0) s3[c][k >> 6] = a;
  }
  __syncthreads();
  float pp = 0.f, pt = 0.f, tt = 0.f;
#pragma unroll
  for (int w = 0; w < MLP_OUT / 64; ++w) {
    pp += s3[0][w];
    pt += s3[1][w];
    tt += s3[2][w];
  }
  float g = 0.f, loss = 0.f;
  if (pp > 0.f && tt > 0.f) {
    const float np = sqrtf(pp), nt = sqrtf(tt);
    loss = nt * (1.f - pt / (np * nt));
    // d loss / d p = -(1 / np) (t - (pt / pp) p)
    g = -(t - (pt / pp) * p) / np;
  }
  dY[r * MLP_OUT + k] = g * q / (float)rows;
  if (k == 0) part[r] = loss;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2, float c1, float c2,
                            float eps) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Features of nb blocks: block blk[b] of r (scale s_r) and of the diagonally
// preconditioned residual q (scale s_q), blk = null meaning block b.  X is nb x 512.
__global__ void features_kernel(const double* __restrict__ r, const double* __restrict__ q,
                                const int64_t* __restrict__ blk, int64_t nb, int64_t n, double s_r, double s_q,
                                float* __restrict__ X) {
  const int64_t total = nb * MLP_BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (int64_t)gridDim.x * TPB) {
    const int64_t b = i / MLP_BLOCK, k = i % MLP_BLOCK;
    const int64_t e = (blk ? blk[b] : b) * MLP_BLOCK + k;
    X[b * MLP_IN + k] = e < n ? (float)(r[e] * s_r) : 0.f;
    X[b * MLP_IN + MLP_BLOCK + k] = e < n ? (float)(q[e] * s_q) : 0.f;
  }
}

// q = r / (theta - diag), the denominator kept away from zero (|.| >= dmin)
__global__ void diagprec_kernel(const double* __restrict__ r, const double* __restrict__ diag, double theta,
                                double dmin, int64_t n, double* __restrict__ q) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    const double den = theta - diag[i];
    q[i] = r[i] / (fabs(den) < dmin ? (den < 0 ? -dmin : dmin) : den);
  }
}

// rows[i] = src[idx[i]] for row length len (float)
__global__ void gather_rows_kernel(const float* __restrict__ src, const int64_t* __restrict__ idx, int64_t rows,
                                   int len, float* __restrict__ dst) {
  const int64_t total = rows * len;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (int64_t)gridDim.x * TPB)
    dst[i] = src[idx[i / len] * len + i % len];
}

// t_b = |q_b| P_b / |P_b| with P_b = q_b (1 + Y_b), q already in t: the network
// turns each block of the diagonal correction but keeps its norm (the training
// loss only sees directions).  One 256-thread workgroup per block; elements at
// or past n belong to no block.
__global__ __launch_bounds__(MLP_OUT) void scatter_kernel(const float* __restrict__ Y, int64_t n,
                                                          double* __restrict__ t) {
  __shared__ double s2[2][MLP_OUT / 64];
  const int64_t i = (int64_t)blockIdx.x * MLP_BLOCK + threadIdx.x;
  const double q = i < n ? t[i] : 0.0;
  const double p = q * (1.0 + (double)Y[i]);
  double v[2] = {q * q, p * p};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    double a = v[c];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if ((threadIdx.x & 63) == 0) s2[c][threadIdx.x >> 6] = a;
  }
  __syncthreads();
  double qq = 0.0, pp = 0.0;
#pragma unroll
  for (int w = 0; w < MLP_OUT / 64; ++w) {
    qq += s2[0][w];
    pp += s2[1][w];
  }
  if (i < n) t[i] = pp > 0.0 ? p * sqrt(qq / pp) : q;
}

// Targets of sampled blocks: T[b] = (sgn x* - ov x) block * s_t
__global__ void targets_kernel(const double* __restrict__ xstar, double sgn, const double* __restrict__ x, double ov,
                               const int64_t* __restrict__ blk, int64_t nb, int64_t n, double s_t,
                               float* __restrict__ T) {
  const int64_t total = nb * MLP_BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (int64_t)gridDim.x * TPB) {
    const int64_t b = i / MLP_BLOCK, k = i % MLP_BLOCK;
    const int64_t e = blk[b] * MLP_BLOCK + k;
    T[b * MLP_OUT + k] = e < n ? (float)((sgn * xstar[e] - ov * x[e]) * s_t) : 0.f;
  }
}

dim3 grid_for(int64_t n) {
  int64_t b = (n + TPB - 1) / TPB;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return dim3((unsigned)b);
}

}  // namespace

mk_status sgemm(bool ta, bool tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B, int ldb,
                float beta, float* C, int ldc, hipStream_t st) {
  if (M == 0 || N == 0) return MK_OK;
  dim3 g((N + GB - 1) / GB, (M + GB - 1) / GB);
  if (!ta && !tb) hipLaunchKernelGGL((sgemm_kernel<false, false>), g, dim3(TPB), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (ta && !tb) hipLaunchKernelGGL((sgemm_kernel<true, false>), g, dim3(TPB), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (!ta && tb) hipLaunchKernelGGL((sgemm_kernel<false, true>), g, dim3(TPB), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else hipLaunchKernelGGL((sgemm_kernel<true, true>), g, dim3(TPB), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

// ---- Mlp ----------------------------------------------------------------------

namespace {
const int kIn[3] = {MLP_IN, MLP_HID, MLP_HID};
const int kOut[3] = {MLP_HID, MLP_HID, MLP_OUT};
}  // namespace

int64_t Mlp::param_count() {
  int64_t n = 0;
  for (int l = 0; l < 3; ++l) n += (int64_t)kIn[l] * kOut[l] + kOut[l];
  return n;
}
float* Mlp::W(int l) const {
  float* p = static_cast<float*>(params.ptr);
  for (int i = 0; i < l; ++i) p += (int64_t)kIn[i] * kOut[i] + kOut[i];
  return p;
}
float* Mlp::b(int l) const { return W(l) + (int64_t)kIn[l] * kOut[l]; }
float* Mlp::dW(int l) const { return static_cast<float*>(grads.ptr) + (W(l) - static_cast<float*>(params.ptr)); }
float* Mlp::db(int l) const { return dW(l) + (int64_t)kIn[l] * kOut[l]; }

mk_status Mlp::init(uint64_t seed, hipStream_t st) {
  const int64_t np = param_count();
  MK_TRY(params.alloc((size_t)np * 4));
  MK_TRY(grads.alloc((size_t)np * 4));
  MK_TRY(m1.alloc((size_t)np * 4));
  MK_TRY(m2.alloc((size_t)np * 4));
  // Glorot-uniform weights from a 64-bit LCG, zero biases; the last layer
  // starts at zero, so an untrained network is the diagonal correction
  std::vector<float> h((size_t)np, 0.f);
  uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto u01 = [&]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  };
  int64_t o = 0;
  for (int l = 0; l < 3; ++l) {
    const double lim = l == 2 ? 0.0 : std::sqrt(6.0 / (kIn[l] + kOut[l]));
    for (int64_t i = 0; i < (int64_t)kIn[l] * kOut[l]; ++i) h[o + i] = (float)((2.0 * u01() - 1.0) * lim);
    o += (int64_t)kIn[l] * kOut[l] + kOut[l];
  }
  MK_CHECK_HIP(hipMemcpyAsync(params.ptr, h.data(), (size_t)np * 4, hipMemcpyHostToDevice, st));
  MK_CHECK_HIP(hipMemsetAsync(m1.ptr, 0, (size_t)np * 4, st));
  MK_CHECK_HIP(hipMemsetAsync(m2.ptr, 0, (size_t)np * 4, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  step = 0;
  return MK_OK;
}

mk_status Mlp::ensure_act(int64_t rows) {
  if (rows <= act_rows) return MK_OK;
  // earlier launches (all on the context's stream) may still use the old buffers
  if (act_rows > 0) MK_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const size_t per = (size_t)rows * 4;
  MK_TRY(act1.alloc(per * MLP_HID));
  MK_TRY(act2.alloc(per * MLP_HID));
  MK_TRY(act3.alloc(per * MLP_OUT));
  MK_TRY(d1.alloc(per * MLP_HID));
  MK_TRY(d2.alloc(per * MLP_HID));
  MK_TRY(d3.alloc(per * MLP_OUT));
  MK_TRY(part.alloc(per));
  act_rows = rows;
  return MK_OK;
}

mk_status Mlp::forward(const float* X, int64_t rows, float* Y, hipStream_t st) {
  MK_TRY(ensure_act(rows));
  float* Y1 = static_cast<float*>(act1.ptr);
  float* Y2 = static_cast<float*>(act2.ptr);
  float* Y3 = Y ? Y : static_cast<float*>(act3.ptr);
  MK_TRY(sgemm(false, true, (int)rows, MLP_HID, MLP_IN, 1.f, X, MLP_IN, W(0), MLP_IN, 0.f, Y1, MLP_HID, st));
  hipLaunchKernelGGL(bias_act_kernel, grid_for(rows * MLP_HID), dim3(TPB), 0, st, Y1, b(0), rows, MLP_HID, 1);
  MK_TRY(sgemm(false, true, (int)rows, MLP_HID, MLP_HID, 1.f, Y1, MLP_HID, W(1), MLP_HID, 0.f, Y2, MLP_HID, st));
  hipLaunchKernelGGL(bias_act_kernel, grid_for(rows * MLP_HID), dim3(TPB), 0, st, Y2, b(1), rows, MLP_HID, 1);
  MK_TRY(sgemm(false, true, (int)rows, MLP_OUT, MLP_HID, 1.f, Y2, MLP_HID, W(2), MLP_HID, 0.f, Y3, MLP_OUT, st));
  hipLaunchKernelGGL(bias_act_kernel, grid_for(rows * MLP_OUT), dim3(TPB), 0, st, Y3, b(2), rows, MLP_OUT, 0);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status Mlp::backward(const float* X, const float* T, int64_t rows, float* loss, hipStream_t st) {
  MK_TRY(forward(X, rows, nullptr, st));
  float* Y1 = static_cast<float*>(act1.ptr);
  float* Y2 = static_cast<float*>(act2.ptr);
  float* Y3 = static_cast<float*>(act3.ptr);
  float* D1 = static_cast<float*>(d1.ptr);
  float* D2 = static_cast<float*>(d2.ptr);
  float* D3 = static_cast<float*>(d3.ptr);
  hipLaunchKernelGGL(cos_loss_kernel, dim3((unsigned)rows), dim3(MLP_OUT), 0, st, Y3, X, T, rows, D3,
                     static_cast<float*>(part.ptr));
  // layer 3
  MK_TRY(sgemm(true, false, MLP_OUT, MLP_HID, (int)rows, 1.f, D3, MLP_OUT, Y2, MLP_HID, 0.f, dW(2), MLP_HID, st));
  hipLaunchKernelGGL(colsum_kernel, dim3(MLP_OUT), dim3(TPB), 0, st, D3, rows, MLP_OUT, db(2));
  MK_TRY(sgemm(false, false, (int)rows, MLP_HID, MLP_OUT, 1.f, D3, MLP_OUT, W(2), MLP_HID, 0.f, D2, MLP_HID, st));
  hipLaunchKernelGGL(tanh_bwd_kernel, grid_for(rows * MLP_HID), dim3(TPB), 0, st, D2, Y2, rows * MLP_HID);
  // layer 2
  MK_TRY(sgemm(true, false, MLP_HID, MLP_HID, (int)rows, 1.f, D2, MLP_HID, Y1, MLP_HID, 0.f, dW(1), MLP_HID, st));
  hipLaunchKernelGGL(colsum_kernel, dim3(MLP_HID), dim3(TPB), 0, st, D2, rows, MLP_HID, db(1));
  MK_TRY(sgemm(false, false, (int)rows, MLP_HID, MLP_HID, 1.f, D2, MLP_HID, W(1), MLP_HID, 0.f, D1, MLP_HID, st));
  hipLaunchKernelGGL(tanh_bwd_kernel, grid_for(rows * MLP_HID), dim3(TPB), 0, st, D1, Y1, rows * MLP_HID);
  // layer 1
  MK_TRY(sgemm(true, false, MLP_HID, MLP_IN, (int)rows, 1.f, D1, MLP_HID, X, MLP_IN, 0.f, dW(0), MLP_IN, st));
  hipLaunchKernelGGL(colsum_kernel, dim3(MLP_HID), dim3(TPB), 0, st, D1, rows, MLP_HID, db(0));
  MK_CHECK_HIP(hipGetLastError());
  if (loss) {
    std::vector<float> p((size_t)rows);
    MK_CHECK_HIP(hipMemcpyAsync(p.data(), part.ptr, (size_t)rows * 4, hipMemcpyDeviceToHost, st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (float v : p) s += v;
    *loss = (float)(s / (double)rows);
  }
  return MK_OK;
}

mk_status Mlp::adam(float lr, hipStream_t st) {
  ++step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  const int64_t np = param_count();
  hipLaunchKernelGGL(adam_kernel, grid_for(np), dim3(TPB), 0, st, static_cast<float*>(params.ptr),
                     static_cast<const float*>(grads.ptr), static_cast<float*>(m1.ptr), static_cast<float*>(m2.ptr),
                     np, lr, b1, b2, c1, c2, 1e-8f);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status mlp_features(const double* r, const double* q, const int64_t* blk, int64_t nb, int64_t n, double s_r,
                       double s_q, float* X, hipStream_t st) {
  if (nb == 0) return MK_OK;
  hipLaunchKernelGGL(features_kernel, grid_for(nb * MLP_BLOCK), dim3(TPB), 0, st, r, q, blk, nb, n, s_r, s_q, X);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status mlp_diagprec(const double* r, const double* diag, double theta, double dmin, int64_t n, double* q,
                       hipStream_t st) {
  hipLaunchKernelGGL(diagprec_kernel, grid_for(n), dim3(TPB), 0, st, r, diag, theta, dmin, n, q);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status mlp_scatter(const float* Y, int64_t nb, int64_t n, double* t, hipStream_t st) {
  if (nb == 0) return MK_OK;
  MK_REQUIRE(nb * MLP_BLOCK >= n, MK_BAD_ARG, "mlp scatter: %lld blocks do not cover %lld elements", (long long)nb,
             (long long)n);
  hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)nb), dim3(MLP_OUT), 0, st, Y, n, t);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status mlp_targets(const double* xstar, double sgn, const double* x, double ov, const int64_t* blk, int64_t nb,
                      int64_t n, double s_t, float* T, hipStream_t st) {
  if (nb == 0) return MK_OK;
  hipLaunchKernelGGL(targets_kernel, grid_for(nb * MLP_BLOCK), dim3(TPB), 0, st, xstar, sgn, x, ov, blk, nb, n, s_t,
                     T);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

mk_status mlp_gather_rows(const float* src, const int64_t* idx, int64_t rows, int len, float* dst, hipStream_t st) {
  if (rows == 0) return MK_OK;
  hipLaunchKernelGGL(gather_rows_kernel, grid_for(rows * len), dim3(TPB), 0, st, src, idx, rows, len, dst);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

}  // namespace mk
