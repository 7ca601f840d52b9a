// Learned Davidson preconditioner: FP32 MLP on 256-element residual blocks.
#pragma once
#include <cstdint>

#include "ops.h"

namespace mk {

constexpr int MLP_BLOCK = 256;
constexpr int MLP_IN = 2 * MLP_BLOCK;
constexpr int MLP_HID = 256;
constexpr int MLP_OUT = MLP_BLOCK;

mk_status sgemm(bool ta, bool tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B, int ldb,
                float beta, float* C, int ldc, hipStream_t st);

struct Mlp {
  Ctx* ctx = nullptr;
  DevBuf params, grads, m1, m2;        // flattened [W0 b0 W1 b1 W2 b2]
  DevBuf act1, act2, act3, d1, d2, d3, part;
  int64_t act_rows = 0;
  int64_t step = 0;
  // training set (rows of 512 features / 256 targets)
  DevBuf DX, DT;
  int64_t rows = 0, cap = 0;
  DevBuf bx, bt, bidx;                 // minibatch staging
  int64_t brows = 0;

  static int64_t param_count();
  float* W(int l) const;
  float* b(int l) const;
  float* dW(int l) const;
  float* db(int l) const;
  mk_status init(uint64_t seed, hipStream_t st);
  mk_status ensure_act(int64_t rows);
  // Y (rows x 256) = MLP(X); Y = null keeps the output in act3
  mk_status forward(const float* X, int64_t rows, float* Y, hipStream_t st);
  // gradients of the mean squared error to T into grads (loss on the host if requested)
  mk_status backward(const float* X, const float* T, int64_t rows, float* loss, hipStream_t st);
  mk_status adam(float lr, hipStream_t st);
};

mk_status mlp_features(const double* r, const double* q, const int64_t* blk, int64_t nb, int64_t n, double s_r,
                       double s_q, float* X, hipStream_t st);
mk_status mlp_diagprec(const double* r, const double* diag, double theta, double dmin, int64_t n, double* q,
                       hipStream_t st);
mk_status mlp_scatter(const float* Y, int64_t nb, int64_t n, double* t, hipStream_t st);
mk_status mlp_targets(const double* xstar, double sgn, const double* x, double ov, const int64_t* blk, int64_t nb,
                      int64_t n, double s_t, float* T, hipStream_t st);
mk_status mlp_gather_rows(const float* src, const int64_t* idx, int64_t rows, int len, float* dst, hipStream_t st);

}  // namespace mk
