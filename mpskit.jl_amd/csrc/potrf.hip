// Blocked Cholesky factorisation A = L L^T of a symmetric positive definite
// float64 matrix (column-major, lower triangle used), for the Cholesky QR of
// the gauge steps.  A 64-column step is one small kernel and two MFMA GEMMs:
// the diagonal block is factored in registers by one workgroup, which also
// inverts its factor (L11^-1); the panel below it is then the GEMM
// L21 = A21 L11^-T, and the trailing update A22 -= L21 L21^T a second one
// (the inverted diagonal block turns the row-serial triangular solve into a
// parallel product, as GPU Cholesky codes commonly do).
#include <algorithm>

#include "gemm.h"
#include "kernels.h"
#include "ops.h"

namespace mk {
namespace {

constexpr int NB = 64;        // block columns per step
constexpr int PT = 256;       // threads of the diagonal kernel
constexpr int NQ = NB / 4;    // columns a thread owns (row t % 64, columns t / 64 + 4 q)

typedef __attribute__((address_space(1))) double gdouble;

// Factor the nb x nb diagonal block at (k0, k0) and invert the factor.  Thread
// t keeps row r = t % 64 of the columns c = t / 64 + 4 q in registers; step j
// broadcasts the current column j through LDS (double-buffered, so one
// barrier per step) and every thread updates its own entries.  The inverse is
// the right-looking forward substitution L X = 1 over the same ownership,
// broadcasting row i of X per step.  A non-positive or non-finite pivot sets
// *info = k0 + j + 1 (LAPACK's convention); every thread sees the same pivot,
// so all stop together.  Writes L11 (zeros above the diagonal) into A and
// L11^-1 into Linv (ld NB, zeros above the diagonal).
__global__ __launch_bounds__(PT) void potrf_diag_kernel(double* __restrict__ A, int64_t lda, int k0, int nb,
                                                        double* __restrict__ Linv, int* __restrict__ info) {
  __shared__ double col[2][NB];
  __shared__ double Ls[NB][NB + 1];
  const int t = threadIdx.x;
  const int r = t & (NB - 1), cg = t / NB;
  gdouble* gA = (gdouble*)A;
  if (*(volatile int*)info != 0) return;   // an earlier block failed
  double s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int c = cg + 4 * q;
    s[q] = (r < nb && c < nb) ? gA[(int64_t)(k0 + r) + (int64_t)(k0 + c) * lda] : 0.0;
  }
  int bad = 0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nb) {
      if (cg == (j & 3)) col[j & 1][r] = s[j >> 2];
      __syncthreads();
      const double d = col[j & 1][j];
      if (!(d > 0.0) || !isfinite(d)) {
        bad = j + 1;
        break;
      }
      const double ljj = sqrt(d);
      const double inv = 1.0 / ljj;
      const double xr = col[j & 1][r] * inv;   // L[r][j] for r > j
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int c = cg + 4 * q;
        if (c > j && r >= c) s[q] -= xr * (col[j & 1][c] * inv);
      }
      if (cg == (j & 3)) s[j >> 2] = r > j ? xr : (r == j ? ljj : s[j >> 2]);
    }
  }
  if (bad) {
    if (t == 0) *info = k0 + bad;
    return;
  }
  // L11 to A and to LDS (zero above the diagonal and outside nb)
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int c = cg + 4 * q;
    const double v = (r >= c && r < nb && c < nb) ? s[q] : 0.0;
    Ls[r][c] = v;
    if (r < nb && c < nb) gA[(int64_t)(k0 + r) + (int64_t)(k0 + c) * lda] = v;
  }
  __syncthreads();
  // X = L11^-1: W = 1; step i: X[i][.] = W[i][.] / L[i][i], then W[r][.] -= L[r][i] X[i][.] for r > i
#pragma unroll
  for (int q = 0; q < NQ; ++q) s[q] = (r == cg + 4 * q && r < nb) ? 1.0 : 0.0;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (i < nb) {
      if (r == i) {
        const double inv = 1.0 / Ls[i][i];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          s[q] *= inv;
          col[i & 1][cg + 4 * q] = s[q];
        }
      }
      __syncthreads();
      if (r > i) {
        const double lri = Ls[r][i];
#pragma unroll
        for (int q = 0; q < NQ; ++q) s[q] -= lri * col[i & 1][cg + 4 * q];
      }
    }
  }
  gdouble* gL = (gdouble*)Linv;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int c = cg + 4 * q;
    gL[r + c * NB] = (r >= c && r < nb && c < nb) ? s[q] : 0.0;
  }
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

mk_status potrf_lower(double* A, int64_t n, int64_t lda, int* info_dev, double* Linv, double* X,
                      hipStream_t st) {
  MK_REQUIRE(n >= 0 && lda >= std::max<int64_t>(n, 1), MK_BAD_SHAPE, "potrf: n %lld, lda %lld", (long long)n,
             (long long)lda);
  MK_REQUIRE(n < (1 << 30), MK_BAD_SHAPE, "potrf: n %lld too large", (long long)n);
  if (n == 0) return MK_OK;
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, n - k0);
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(PT), 0, st, A, lda, (int)k0, nb, Linv, info_dev);
    MK_CHECK_HIP(hipGetLastError());
    const int64_t m = n - k0 - nb;
    if (m <= 0) break;
    double* A21 = A + (k0 + nb) + k0 * lda;
    // X = A21 L11^-T (m x nb, ld m)
    ProbSpec pp;
    pp.M = m; pp.N = nb; pp.C = X; pp.ldc = m;
    pp.add(A21, lda, Linv, NB, nb);
    GemmShape tb;
    tb.tb = true;   // B(k, j) = Linv[j + k NB] = L11^-1(j, k)
    MK_TRY(gemm_f64(&pp, 1, tb, st));
    MK_CHECK_HIP(hipMemcpy2DAsync(A21, lda * sizeof(double), X, m * sizeof(double), m * sizeof(double), nb,
                                  hipMemcpyDeviceToDevice, st));
    // A22 -= X X^T (full square; the strict upper part is zeroed at the end)
    ProbSpec p;
    p.M = m; p.N = m;
    p.C = A + (k0 + nb) + (k0 + nb) * lda;
    p.ldc = lda;
    p.alpha = -1.0;
    p.beta = 1.0;
    p.add(X, m, X, m, nb);
    MK_TRY(gemm_f64(&p, 1, tb, st));   // B(k, j) = X[j + k m] = X(j, k)
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
  double *Linv = nullptr, *X = nullptr;   // inverted diagonal block, panel of one step
  MK_TRY(carve(ctx, {{(size_t)NB * NB * 8, (void**)&Linv}, {(size_t)std::max<int64_t>(n, 1) * NB * 8, (void**)&X}}));
  MK_CHECK_HIP(hipMemsetAsync(dinfo, 0, sizeof(int), st));
  MK_TRY(potrf_lower(A, n, lda, dinfo, Linv, X, st));
  MK_CHECK_HIP(hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}
