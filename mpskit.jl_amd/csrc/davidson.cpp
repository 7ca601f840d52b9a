// Davidson eigensolver with a learned preconditioner, and the MLP's C ABI.
//
// The local problems of DMRG/VUMPS (fixedpoint.jl:15-18) are solved in the
// reference by Lanczos.  mk_davidson solves the same lowest-eigenpair problem
// by Davidson iteration on the prepared operator, all vectors and Rayleigh-Ritz
// quantities in float64 on the device; the correction added to the search
// space each step is
//   mode 0: the residual r (Krylov space; Lanczos-equivalent)
//   mode 1: the diagonal (Davidson) correction r / (theta - diag H)
//   mode 2: the learned refinement q (1 + MLP(r, q)) of q = r / (theta - diag H).
// The float64 residual norm |H x - theta x| decides convergence whatever the
// correction; if a learned correction collapses against the search space the
// step falls back to the residual, and if Davidson has not converged after
// maxiter steps the solve continues with mk_eigsolve's Lanczos from the current
// vector (plain Lanczos as fallback).
//
// With collect = 1 and an MLP, (features, ideal correction) pairs of sampled
// blocks of the iterates are appended to the MLP's training set once the
// eigenvector x* is known: target = sgn(x . x*) x* - |x . x*| x, normalised.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "kernels.h"
#include "mlp.h"
#include "ops.h"

namespace mk {

void lanczos_lowest(const std::vector<double>& G, int ld, int k, double& lam, std::vector<double>& s);
mk_status op_diagonal(Op* op, double* diag, hipStream_t st);

namespace {

struct Rec {
  DevBuf x;       // iterate (n doubles)
  DevBuf X;       // features of the sampled blocks (S x 512 floats)
  DevBuf blk;     // sampled block indices (S int64)
  int64_t S = 0;
};

mk_status dot1(const double* a, const double* b, int64_t n, double* dv, double* part, hipStream_t st, double& out) {
  MK_TRY(k_multidot(a, n, 1, b, n, dv, part, st));
  MK_CHECK_HIP(hipMemcpyAsync(&out, dv, 8, hipMemcpyDeviceToHost, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}

}  // namespace
}  // namespace mk

using namespace mk;

extern "C" {

mk_status mk_mlp_create(mk_ctx h, uint64_t seed, int64_t capacity, mk_mlp* out) {
  MK_REQUIRE(h != nullptr && out != nullptr, MK_BAD_ARG, "mlp: null argument");
  MK_REQUIRE(capacity >= 0, MK_BAD_ARG, "mlp: negative capacity");
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  Mlp* m = new Mlp();
  m->ctx = ctx;
  mk_status s = m->init(seed, ctx->stream);
  if (s == MK_OK && capacity > 0) {
    s = m->DX.alloc((size_t)capacity * MLP_IN * 4);
    if (s == MK_OK) s = m->DT.alloc((size_t)capacity * MLP_OUT * 4);
    m->cap = capacity;
  }
  if (s != MK_OK) {
    delete m;
    return s;
  }
  ctx->retain();
  *out = reinterpret_cast<mk_mlp>(m);
  return MK_OK;
}

mk_status mk_mlp_free(mk_mlp h) {
  if (!h) return MK_OK;
  Mlp* m = reinterpret_cast<Mlp*>(h);
  Ctx* ctx = m->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  delete m;
  ctx->release();
  return MK_OK;
}

mk_status mk_mlp_param_count(int64_t* n) {
  MK_REQUIRE(n != nullptr, MK_BAD_ARG, "mlp: null argument");
  *n = Mlp::param_count();
  return MK_OK;
}

mk_status mk_mlp_get_params(mk_mlp h, float* host) {
  MK_REQUIRE(h != nullptr && host != nullptr, MK_BAD_ARG, "mlp: null argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  MK_CHECK_HIP(hipMemcpyAsync(host, m->params.ptr, (size_t)Mlp::param_count() * 4, hipMemcpyDeviceToHost,
                              m->ctx->stream));
  MK_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return MK_OK;
}

mk_status mk_mlp_set_params(mk_mlp h, const float* host) {
  MK_REQUIRE(h != nullptr && host != nullptr, MK_BAD_ARG, "mlp: null argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  MK_CHECK_HIP(hipMemcpyAsync(m->params.ptr, host, (size_t)Mlp::param_count() * 4, hipMemcpyHostToDevice,
                              m->ctx->stream));
  MK_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return MK_OK;
}

mk_status mk_mlp_forward(mk_mlp h, const float* X, int64_t rows, float* Y) {
  MK_REQUIRE(h != nullptr && X != nullptr && Y != nullptr && rows >= 0, MK_BAD_ARG, "mlp forward: bad argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  MK_CHECK_HIP(hipSetDevice(m->ctx->device));
  return m->forward(X, rows, Y, m->ctx->stream);
}

mk_status mk_mlp_gradient(mk_mlp h, const float* X, const float* T, int64_t rows, float* grads_host, float* loss) {
  MK_REQUIRE(h != nullptr && X != nullptr && T != nullptr && rows > 0, MK_BAD_ARG, "mlp gradient: bad argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  MK_CHECK_HIP(hipSetDevice(m->ctx->device));
  MK_TRY(m->backward(X, T, rows, loss, m->ctx->stream));
  if (grads_host) {
    MK_CHECK_HIP(hipMemcpyAsync(grads_host, m->grads.ptr, (size_t)Mlp::param_count() * 4, hipMemcpyDeviceToHost,
                                m->ctx->stream));
    MK_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  }
  return MK_OK;
}

mk_status mk_mlp_dataset_size(mk_mlp h, int64_t* rows) {
  MK_REQUIRE(h != nullptr && rows != nullptr, MK_BAD_ARG, "mlp: null argument");
  *rows = reinterpret_cast<Mlp*>(h)->rows;
  return MK_OK;
}

mk_status mk_mlp_eval(mk_mlp h, int64_t row_begin, int64_t row_end, float* loss) {
  MK_REQUIRE(h != nullptr && loss != nullptr, MK_BAD_ARG, "mlp eval: null argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  MK_REQUIRE(0 <= row_begin && row_begin < row_end && row_end <= m->rows, MK_BAD_ARG,
             "mlp eval: rows [%lld, %lld) outside the %lld collected", (long long)row_begin, (long long)row_end,
             (long long)m->rows);
  MK_CHECK_HIP(hipSetDevice(m->ctx->device));
  const float* X = static_cast<const float*>(m->DX.ptr) + row_begin * MLP_IN;
  const float* T = static_cast<const float*>(m->DT.ptr) + row_begin * MLP_OUT;
  return m->backward(X, T, row_end - row_begin, loss, m->ctx->stream);
}

mk_status mk_mlp_train(mk_mlp h, int steps, int batch, float lr, uint64_t seed, int64_t row_begin, int64_t row_end,
                       float* first_loss, float* last_loss) {
  MK_REQUIRE(h != nullptr && steps >= 0 && batch > 0, MK_BAD_ARG, "mlp train: bad argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  if (row_end < 0) row_end = m->rows;
  MK_REQUIRE(0 <= row_begin && row_begin < row_end && row_end <= m->rows, MK_BAD_ARG,
             "mlp train: rows [%lld, %lld) outside the %lld collected", (long long)row_begin, (long long)row_end,
             (long long)m->rows);
  hipStream_t st = m->ctx->stream;
  MK_CHECK_HIP(hipSetDevice(m->ctx->device));
  const int64_t B = std::min<int64_t>(batch, row_end - row_begin);
  if (m->brows < B) {
    MK_TRY(m->bx.alloc((size_t)B * MLP_IN * 4));
    MK_TRY(m->bt.alloc((size_t)B * MLP_OUT * 4));
    MK_TRY(m->bidx.alloc((size_t)B * 8));
    m->brows = B;
  }
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int64_t> pick(row_begin, row_end - 1);
  std::vector<int64_t> idx((size_t)B);
  for (int it = 0; it < steps; ++it) {
    for (auto& v : idx) v = pick(rng);
    MK_CHECK_HIP(hipMemcpyAsync(m->bidx.ptr, idx.data(), (size_t)B * 8, hipMemcpyHostToDevice, st));
    MK_TRY(mlp_gather_rows(static_cast<const float*>(m->DX.ptr), static_cast<const int64_t*>(m->bidx.ptr), B, MLP_IN,
                           static_cast<float*>(m->bx.ptr), st));
    MK_TRY(mlp_gather_rows(static_cast<const float*>(m->DT.ptr), static_cast<const int64_t*>(m->bidx.ptr), B, MLP_OUT,
                           static_cast<float*>(m->bt.ptr), st));
    float loss = 0.f;
    const bool want = (it == 0 && first_loss) || (it == steps - 1 && last_loss);
    MK_TRY(m->backward(static_cast<const float*>(m->bx.ptr), static_cast<const float*>(m->bt.ptr), B,
                       want ? &loss : nullptr, st));
    MK_TRY(m->adam(lr, st));
    if (it == 0 && first_loss) *first_loss = loss;
    if (it == steps - 1 && last_loss) *last_loss = loss;
    // the host index buffer is reused next step: wait for the copy
    MK_CHECK_HIP(hipStreamSynchronize(st));
  }
  return MK_OK;
}

mk_status mk_davidson(mk_op h, double* x, int maxdim, int maxiter, double tol, mk_mlp hm, int mode, int collect,
                      double* eigval, double* resnorm, int* numops, int* converged) {
  MK_REQUIRE(h != nullptr && x != nullptr, MK_BAD_ARG, "davidson: null argument");
  MK_REQUIRE(maxdim >= 2 && maxdim <= 256, MK_BAD_ARG, "davidson: maxdim %d not in [2, 256]", maxdim);
  MK_REQUIRE(maxiter >= 1 && tol >= 0.0, MK_BAD_ARG, "davidson: bad maxiter / tol");
  MK_REQUIRE(mode >= 0 && mode <= 2, MK_BAD_ARG, "davidson: mode %d", mode);
  MK_REQUIRE(mode < 2 || hm != nullptr, MK_BAD_ARG, "davidson: mode 2 needs an MLP");
  Op* op = reinterpret_cast<Op*>(h);
  Mlp* mlp = reinterpret_cast<Mlp*>(hm);
  MK_REQUIRE(op->x_elems == op->y_elems, MK_BAD_SHAPE, "davidson: operator is not square");
  Ctx* ctx = op->ctx;
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t n = op->x_elems;
  const int64_t nb = cdiv(n, (int64_t)MLP_BLOCK);
  const bool want_diag = mode >= 1 || (collect && mlp);
  DevBuf Vb, AVb, rb, tb, xb, axb, db, cb, pb, Xb, Yb;
  MK_TRY(Vb.alloc((size_t)n * maxdim * 8));
  MK_TRY(AVb.alloc((size_t)n * maxdim * 8));
  MK_TRY(rb.alloc((size_t)n * 8));
  MK_TRY(tb.alloc((size_t)n * 8));
  MK_TRY(xb.alloc((size_t)n * 8));
  MK_TRY(axb.alloc((size_t)n * 8));
  MK_TRY(cb.alloc((size_t)(maxdim + 4) * 8));
  MK_TRY(pb.alloc((size_t)std::max<int64_t>(1, cdiv(n, 256 * 16)) * (maxdim + 4) * 8));
  double* V = static_cast<double*>(Vb.ptr);
  double* AV = static_cast<double*>(AVb.ptr);
  double* r = static_cast<double*>(rb.ptr);
  double* t = static_cast<double*>(tb.ptr);
  double* xv = static_cast<double*>(xb.ptr);
  double* ax = static_cast<double*>(axb.ptr);
  double* cdev = static_cast<double*>(cb.ptr);
  double* part = static_cast<double*>(pb.ptr);
  double* diag = nullptr;
  if (want_diag) {
    MK_TRY(db.alloc((size_t)n * 8));
    diag = static_cast<double*>(db.ptr);
    MK_TRY(op_diagonal(op, diag, st));
  }
  if (mode == 2) {
    MK_TRY(Xb.alloc((size_t)nb * MLP_IN * 4));
    MK_TRY(Yb.alloc((size_t)nb * MLP_OUT * 4));
  }
  auto col = [&](double* M, int j) { return M + (int64_t)j * n; };
  std::vector<double> hc;
  auto dots = [&](const double* Vm, int k, const double* u) -> mk_status {
    MK_TRY(k_multidot(Vm, n, k, u, n, cdev, part, st));
    hc.resize(k);
    MK_CHECK_HIP(hipMemcpyAsync(hc.data(), cdev, (size_t)k * 8, hipMemcpyDeviceToHost, st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
    return MK_OK;
  };
  // v0 = x / |x|
  double nrm = 0.0;
  MK_TRY(dot1(x, x, n, cdev, part, st, nrm));
  MK_REQUIRE(std::isfinite(nrm) && nrm > 0.0, MK_BAD_ARG, "davidson: zero or non-finite start vector");
  MK_TRY(k_axpby(1.0 / std::sqrt(nrm), x, 0.0, col(V, 0), n, st));
  MK_TRY(op_apply(op, col(V, 0), col(AV, 0), 1.0, 0.0, MK_FP64));
  int ops = 1, k = 1;
  const int ldh = maxdim;
  std::vector<double> Hk((size_t)ldh * ldh, 0.0), s;
  MK_TRY(dots(col(V, 0), 1, col(AV, 0)));
  Hk[0] = hc[0];
  double theta = hc[0], res = INFINITY;
  bool conv = false;
  std::vector<Rec> recs;
  std::mt19937_64 rng(0x5eed + (uint64_t)n);
  const int max_rec = 16;
  for (int it = 0;; ++it) {
    double lam;
    lanczos_lowest(Hk, ldh, k, lam, s);
    theta = lam;
    MK_CHECK_HIP(hipMemcpyAsync(cdev, s.data(), (size_t)k * 8, hipMemcpyHostToDevice, st));
    MK_TRY(k_gemv(V, n, k, cdev, xv, n, st));
    MK_TRY(k_gemv(AV, n, k, cdev, ax, n, st));
    MK_TRY(k_axpby(1.0, ax, 0.0, r, n, st));
    MK_TRY(k_axpby(-theta, xv, 1.0, r, n, st));
    double rr = 0.0;
    MK_TRY(dot1(r, r, n, cdev, part, st, rr));
    res = std::sqrt(std::max(rr, 0.0));
    MK_REQUIRE(std::isfinite(theta) && std::isfinite(res), MK_NONFINITE, "davidson: non-finite Ritz pair");
    if (res <= tol) {
      conv = true;
      break;
    }
    if (it >= maxiter) break;
    // correction
    if (mode >= 1 || (collect && mlp)) MK_TRY(mlp_diagprec(r, diag, theta, 1e-2, n, t, st));
    double qq = 0.0;
    if (mode == 2 || (collect && mlp)) MK_TRY(dot1(t, t, n, cdev, part, st, qq));
    const double scale = std::sqrt((double)n / MLP_BLOCK);
    const double s_r = scale / res, s_q = qq > 0.0 ? scale / std::sqrt(qq) : 0.0;
    if (collect && mlp && (int)recs.size() < max_rec && mlp->cap > 0) {
      Rec rc;
      rc.S = std::min<int64_t>(nb, 64);
      std::vector<int64_t> bi((size_t)rc.S);
      std::uniform_int_distribution<int64_t> pick(0, nb - 1);
      for (auto& v : bi) v = pick(rng);
      MK_TRY(rc.x.alloc((size_t)n * 8));
      MK_TRY(rc.X.alloc((size_t)rc.S * MLP_IN * 4));
      MK_TRY(rc.blk.alloc((size_t)rc.S * 8));
      MK_CHECK_HIP(hipMemcpyAsync(rc.blk.ptr, bi.data(), (size_t)rc.S * 8, hipMemcpyHostToDevice, st));
      MK_CHECK_HIP(hipMemcpyAsync(rc.x.ptr, xv, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
      MK_TRY(mlp_features(r, t, static_cast<const int64_t*>(rc.blk.ptr), rc.S, n, s_r, s_q,
                          static_cast<float*>(rc.X.ptr), st));
      MK_CHECK_HIP(hipStreamSynchronize(st));
      recs.push_back(std::move(rc));
    }
    if (mode == 0) {
      MK_TRY(k_axpby(1.0, r, 0.0, t, n, st));
    } else if (mode == 2) {
      MK_TRY(mlp_features(r, t, nullptr, nb, n, s_r, s_q, static_cast<float*>(Xb.ptr), st));
      MK_TRY(mlp->forward(static_cast<const float*>(Xb.ptr), nb, static_cast<float*>(Yb.ptr), st));
      MK_TRY(mlp_scatter(static_cast<const float*>(Yb.ptr), nb, n, t, st));
    }
    // restart: keep the Ritz vector
    if (k == maxdim) {
      MK_CHECK_HIP(hipMemcpyAsync(col(V, 0), xv, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
      MK_CHECK_HIP(hipMemcpyAsync(col(AV, 0), ax, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
      std::fill(Hk.begin(), Hk.end(), 0.0);
      Hk[0] = theta;
      k = 1;
    }
    // orthogonalise the correction (two classical Gram-Schmidt passes)
    double tt0 = 0.0;
    MK_TRY(dot1(t, t, n, cdev, part, st, tt0));
    for (int pass = 0; pass < 2; ++pass) {
      MK_TRY(k_multidot(V, n, k, t, n, cdev, part, st));
      MK_TRY(k_multiaxpy(V, n, k, cdev, t, n, st));
    }
    double tt = 0.0;
    MK_TRY(dot1(t, t, n, cdev, part, st, tt));
    if (!(tt > 1e-20 * tt0) || !std::isfinite(tt)) {
      // the correction collapsed against the search space: use the residual
      MK_TRY(k_axpby(1.0, r, 0.0, t, n, st));
      for (int pass = 0; pass < 2; ++pass) {
        MK_TRY(k_multidot(V, n, k, t, n, cdev, part, st));
        MK_TRY(k_multiaxpy(V, n, k, cdev, t, n, st));
      }
      MK_TRY(dot1(t, t, n, cdev, part, st, tt));
      if (!(tt > 0.0)) break;
    }
    MK_TRY(k_axpby(1.0 / std::sqrt(tt), t, 0.0, col(V, k), n, st));
    MK_TRY(op_apply(op, col(V, k), col(AV, k), 1.0, 0.0, MK_FP64));
    ++ops;
    MK_TRY(dots(V, k + 1, col(AV, k)));
    for (int i = 0; i <= k; ++i) Hk[(size_t)i * ldh + k] = Hk[(size_t)k * ldh + i] = hc[i];
    ++k;
  }
  MK_CHECK_HIP(hipMemcpyAsync(x, xv, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  if (!conv) {
    // plain Lanczos from the current vector
    double ev = theta, rn = res;
    int nops = 0, cv = 0;
    MK_TRY(mk_eigsolve(h, x, std::min(maxdim, 30), 200, tol, MK_FP64, &ev, &rn, &nops, &cv));
    theta = ev;
    res = rn;
    ops += nops;
    conv = cv != 0;
  }
  // training pairs from the recorded iterates
  if (collect && mlp && conv && !recs.empty()) {
    for (Rec& rc : recs) {
      if (mlp->rows + rc.S > mlp->cap) break;
      double ov = 0.0;
      MK_TRY(dot1(static_cast<const double*>(rc.x.ptr), x, n, cdev, part, st, ov));
      const double sgn = ov < 0 ? -1.0 : 1.0;
      const double a = std::fabs(ov);
      const double tn = std::sqrt(std::max(1.0 - a * a, 0.0));
      // nearly converged iterates carry no direction information
      if (tn < 1e-4) continue;
      const double s_t = std::sqrt((double)n / MLP_BLOCK) / tn;
      MK_CHECK_HIP(hipMemcpyAsync(static_cast<float*>(mlp->DX.ptr) + mlp->rows * MLP_IN, rc.X.ptr,
                                  (size_t)rc.S * MLP_IN * 4, hipMemcpyDeviceToDevice, st));
      MK_TRY(mlp_targets(x, sgn, static_cast<const double*>(rc.x.ptr), a, static_cast<const int64_t*>(rc.blk.ptr),
                         rc.S, n, s_t, static_cast<float*>(mlp->DT.ptr) + mlp->rows * MLP_OUT, st));
      mlp->rows += rc.S;
    }
  }
  MK_CHECK_HIP(hipStreamSynchronize(st));
  if (eigval) *eigval = theta;
  if (resnorm) *resnorm = res;
  if (numops) *numops = ops;
  if (converged) *converged = conv ? 1 : 0;
  return MK_OK;
}

}  // extern "C"
