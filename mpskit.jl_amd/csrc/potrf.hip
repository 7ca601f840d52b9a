// Blocked Cholesky factorisation A = L L^T of a symmetric positive definite
// float64 matrix (column-major, lower triangle used), for the Cholesky QR of
// the gauge steps.  A 64-column step is three launches: the diagonal block
// factored in LDS by one workgroup, the panel below it solved row-parallel
// against that block (right-looking, so the updates of a step are
// independent), and the trailing update A22 -= L21 L21^T on the FP64 MFMA
// GEMM.  At n = 1024: 2.3 ms against 2.7 ms for rocSOLVER's potrf.  Both are
// latency bound: the diagonal block and the panel (~65 us each per step) are
// serial chains of dependent steps on one or a few waves.
#include <algorithm>

#include "gemm.h"
#include "kernels.h"
#include "ops.h"

namespace mk {
namespace {

constexpr int NB = 64;        // block columns per step
constexpr int PT = 256;       // threads of the diagonal kernel
constexpr int PR = 64;        // rows per workgroup of the panel solve

typedef __attribute__((address_space(1))) double gdouble;

// Unblocked right-looking Cholesky of the nb x nb diagonal block at (k0, k0)
// in LDS.  Thread t owns row t % 64 of the columns t / 64 + 4 q (q < 16), so a
// step is a handful of FMAs per thread between barriers.  A non-positive or
// non-finite pivot sets *info = k0 + j + 1 (the LAPACK convention) and stops
// the factorisation.
__global__ __launch_bounds__(PT) void potrf_diag_kernel(double* __restrict__ A, int64_t lda, int k0, int nb,
                                                        int* __restrict__ info) {
  __shared__ double S[NB][NB + 1];
  __shared__ int bad;
  const int t = threadIdx.x;
  const int r = t & (NB - 1), cg = t / NB;
  gdouble* gA = (gdouble*)A;
  if (*(volatile int*)info != 0) return;   // an earlier block failed
#pragma unroll
  for (int q = 0; q < NB / 4; ++q) {
    const int c = cg + 4 * q;
    S[r][c] = (r < nb && c < nb) ? gA[(int64_t)(k0 + r) + (int64_t)(k0 + c) * lda] : 0.0;
  }
  if (t == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (t == 0) {
      const double d = S[j][j];
      if (!(d > 0.0) || !isfinite(d)) bad = j + 1;
      else S[j][j] = sqrt(d);
    }
    __syncthreads();
    if (bad) break;
    if (cg == 0 && r > j && r < nb) S[r][j] /= S[j][j];
    __syncthreads();
    const double lrj = S[r][j];
#pragma unroll
    for (int q = 0; q < NB / 4; ++q) {
      const int c = cg + 4 * q;
      if (c > j && r >= c && r < nb) S[r][c] -= lrj * S[c][j];
    }
    __syncthreads();
  }
  if (bad) {
    if (t == 0) *info = k0 + bad;
    return;
  }
#pragma unroll
  for (int q = 0; q < NB / 4; ++q) {
    const int c = cg + 4 * q;
    if (r < nb && c < nb) gA[(int64_t)(k0 + r) + (int64_t)(k0 + c) * lda] = r >= c ? S[r][c] : 0.0;
  }
}

// Panel solve L21 = A21 L11^-T for the rows below the diagonal block: one row
// per thread, right-looking (x_j /= L_jj, then x_i -= x_j L_ij for i > j), so
// the updates of a step are independent; L11 in LDS.
__global__ __launch_bounds__(PR) void potrf_panel_kernel(double* __restrict__ A, int64_t lda, int k0, int nb,
                                                         int64_t n, const int* __restrict__ info) {
  __shared__ double L[NB][NB + 1];
  gdouble* gA = (gdouble*)A;
  if (*(volatile const int*)info != 0) return;
  for (int e = threadIdx.x; e < NB * NB; e += PR) {
    const int r = e % NB, c = e / NB;
    L[r][c] = (r < nb && c < nb) ? gA[(int64_t)(k0 + r) + (int64_t)(k0 + c) * lda] : 0.0;
  }
  __syncthreads();
  const int64_t row = (int64_t)k0 + nb + (int64_t)blockIdx.x * PR + threadIdx.x;
  if (row >= n) return;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) x[j] = j < nb ? gA[row + (int64_t)(k0 + j) * lda] : 0.0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nb) {
      x[j] /= L[j][j];
#pragma unroll
      for (int i = j + 1; i < NB; ++i) x[i] -= x[j] * L[i][j];
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j)
    if (j < nb) gA[row + (int64_t)(k0 + j) * lda] = x[j];
}

// Zero the strict upper triangle of the first n columns (the factor is L).
__global__ void zero_upper_kernel(double* __restrict__ A, int64_t lda, int64_t n) {
  gdouble* gA = (gdouble*)A;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % n, c = e / n;
    if (r < c) gA[r + c * lda] = 0.0;
  }
}

}  // namespace

mk_status potrf_lower(double* A, int64_t n, int64_t lda, int* info_dev, hipStream_t st) {
  MK_REQUIRE(n >= 0 && lda >= std::max<int64_t>(n, 1), MK_BAD_SHAPE, "potrf: n %lld, lda %lld", (long long)n,
             (long long)lda);
  MK_REQUIRE(n < (1 << 30), MK_BAD_SHAPE, "potrf: n %lld too large", (long long)n);
  if (n == 0) return MK_OK;
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, n - k0);
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(PT), 0, st, A, lda, (int)k0, nb, info_dev);
    MK_CHECK_HIP(hipGetLastError());
    const int64_t m = n - k0 - nb;
    if (m <= 0) break;
    hipLaunchKernelGGL(potrf_panel_kernel, dim3((unsigned)cdiv(m, (int64_t)PR)), dim3(PR), 0, st, A, lda, (int)k0,
                       nb, n, info_dev);
    MK_CHECK_HIP(hipGetLastError());
    // A22 -= L21 L21^T (full square; the strict upper part is zeroed at the end)
    ProbSpec p;
    p.M = m; p.N = m;
    p.C = A + (k0 + nb) + (k0 + nb) * lda;
    p.ldc = lda;
    p.alpha = -1.0;
    p.beta = 1.0;
    const double* L21 = A + (k0 + nb) + k0 * lda;
    p.add(L21, lda, L21, lda, nb);
    GemmShape sh;
    sh.tb = true;   // B(k, j) = L21[j, k]
    MK_TRY(gemm_f64(&p, 1, sh, st));
  }
  hipLaunchKernelGGL(zero_upper_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n * n, (int64_t)256), 4096)), dim3(256),
                     0, st, A, lda, n);
  MK_CHECK_HIP(hipGetLastError());
  return MK_OK;
}

}  // namespace mk

using namespace mk;

extern "C" mk_status mk_potrf(mk_ctx h, double* A, int64_t n, int64_t lda, int* info) {
  MK_REQUIRE(h != nullptr && A != nullptr && info != nullptr, MK_BAD_ARG, "potrf: null argument");
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  void* mem = nullptr;
  MK_TRY(ctx->small(256, &mem));
  int* dinfo = static_cast<int*>(mem);
  MK_CHECK_HIP(hipMemsetAsync(dinfo, 0, sizeof(int), st));
  MK_TRY(potrf_lower(A, n, lda, dinfo, st));
  MK_CHECK_HIP(hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}
