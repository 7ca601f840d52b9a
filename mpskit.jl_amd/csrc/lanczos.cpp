// Lowest-eigenpair Krylov solver for prepared (Hermitian) derivative operators.
//
// Replaces fixedpoint(H, x0, :SR, Lanczos(...)) (src/algorithms/fixedpoint.jl:15-18)
// for the local DMRG / VUMPS eigenproblems.  All orthogonalisation, Ritz
// residuals and normalisation are float64 on the device; the host only solves
// the small projected eigenproblem and reads a few scalars per step.
//
// Each restart cycle starts from the current vector x with one float64 apply
// w = H x, which gives the exact Ritz value theta = <x, w> and residual
// r = w - theta x.  The cycle extends span{x, y_2 .. y_p} (x and the next
// lowest Ritz vectors kept from the previous cycle: a thick restart, as
// KrylovKit's Lanczos keeps part of its subspace) by a Krylov space of the
// (possibly mixed-precision) operator started from r, orthonormalised in
// float64 (two classical Gram-Schmidt passes, of which the first covers only
// the last three columns after the first Krylov vector of a cycle: the
// three-term recurrence), and takes the lowest Ritz pair
// of the projected matrix, whose x-row <q, H x> uses the float64 w.  Keeping
// Ritz vectors carries the convergence of the cycles before, which a restart
// from x alone loses: with eigenvalue gaps small against the spectral width
// (the local problems of near-critical VUMPS at large D) a single-vector
// restart stalls.  With a float64 inner operator this is thick-restart
// Lanczos; with an FP32 / BF16x3 inner operator it is iterative refinement:
// only the correction is computed in low precision, so every cycle contracts
// the float64 residual by about the low-precision error and the loop stops on
// the float64 residual.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "gemm.h"
#include "kernels.h"
#include "ops.h"
#include "symeig.h"

namespace mk {
namespace {

// the p lowest eigenpairs of the leading k x k block of G (row-major, stride
// ld): eigenvalues ascending in lam, eigenvectors as columns of S (k x p,
// column-major)
void lowest_p(const std::vector<double>& G, int ld, int k, int p, std::vector<double>& lam,
              std::vector<double>& S) {
  std::vector<double> A((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) A[(size_t)i * k + j] = 0.5 * (G[(size_t)i * ld + j] + G[(size_t)j * ld + i]);
  std::vector<double> ev, Q;
  sym_eig(std::move(A), k, ev, Q);   // ascending
  lam.resize(p);
  S.resize((size_t)k * p);
  for (int c = 0; c < p; ++c) {
    lam[c] = ev[c];
    for (int i = 0; i < k; ++i) S[(size_t)c * k + i] = Q[(size_t)i * k + c];
  }
}

// lowest eigenpair of the leading k x k block of G (row-major, stride ld)
void lowest(const std::vector<double>& G, int ld, int k, double& lam, std::vector<double>& s) {
  std::vector<double> A((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) A[(size_t)i * k + j] = 0.5 * (G[(size_t)i * ld + j] + G[(size_t)j * ld + i]);
  std::vector<double> ev, Q;
  sym_eig(std::move(A), k, ev, Q);   // ascending
  lam = ev[0];
  s.resize(k);
  for (int i = 0; i < k; ++i) s[i] = Q[(size_t)i * k];
}

struct Dev {
  Ctx* ctx;
  double* coef;     // device coefficient vector (length >= kmax)
  double* partial;  // multidot partials
  // <V_j, u> for j < k, copied to host
  mk_status dots(const double* V, int64_t n, int k, const double* u, std::vector<double>& out) {
    MK_TRY(k_multidot(V, n, k, u, n, coef, partial, ctx->stream));
    out.resize(k);
    MK_CHECK_HIP(hipMemcpyAsync(out.data(), coef, (size_t)k * 8, hipMemcpyDeviceToHost, ctx->stream));
    MK_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MK_OK;
  }
  // u -= V c with c the device coefficients of the last dots() call
  mk_status subtract(const double* V, int64_t n, int k, double* u) {
    return k_multiaxpy(V, n, k, coef, u, n, ctx->stream);
  }
  mk_status norm(const double* u, int64_t n, double& nrm) {
    std::vector<double> t;
    MK_TRY(dots(u, n, 1, u, t));
    nrm = std::sqrt(std::max(t[0], 0.0));
    return MK_OK;
  }
};

}  // namespace

void lanczos_lowest(const std::vector<double>& G, int ld, int k, double& lam, std::vector<double>& s) {
  lowest(G, ld, k, lam, s);
}

void lanczos_lowest_p(const std::vector<double>& G, int ld, int k, int p, std::vector<double>& lam,
                      std::vector<double>& S) {
  lowest_p(G, ld, k, p, lam, S);
}

// inner Krylov stopping factor for the mixed modes: stop the low-precision
// cycle once its estimated residual is below eta * (float64 residual at the
// start of the cycle), i.e. aim for a few digits of contraction per cycle.
// MK_INNER_ETA overrides the factor (read once).
static double inner_eta(int precision) {
  static const double eta = [] {
    const char* e = getenv("MK_INNER_ETA");
    const double v = e ? atof(e) : 0.0;
    return v > 0.0 && v < 1.0 ? v : 1e-3;
  }();
  return precision == MK_FP64 ? 0.0 : eta;
}

// Krylov steps after the first of a cycle take their first orthogonalisation
// pass on the last three basis columns (the three-term recurrence) instead of
// the whole basis; MK_LANCZOS_CGS=full restores two full passes (read once).
static bool local_first_pass() {
  static const bool on = [] {
    const char* e = getenv("MK_LANCZOS_CGS");
    return !(e && std::string(e) == "full");
  }();
  return on;
}

bool lanczos_local_first_pass() { return local_first_pass(); }

}  // namespace mk

using namespace mk;

// Two classical Gram-Schmidt passes of u against the k <= 32 columns of V
// (KrylovKit's ClassicalGramSchmidt2, the orthogonalisation of the Arnoldi,
// Lanczos and GMRES solvers of the reference): u <- u - V (c1 + c2) with
// c1 = V^T u and c2 = V^T (u - V c1), then |u|.  h[0..k) = c1 + c2 and
// h[k] = |u| on the host.  Three reads of V, one host sync.
extern "C" mk_status mk_krylov_orthogonalize(mk_ctx h, const double* V, int64_t ldv, int k, double* u, int64_t n,
                                             double* hout) {
  MK_REQUIRE(h != nullptr && V != nullptr && u != nullptr && hout != nullptr, MK_BAD_ARG,
             "krylov_orthogonalize: null argument");
  MK_REQUIRE(k >= 1 && k <= 32, MK_BAD_ARG, "krylov_orthogonalize: k = %d not in [1, 32]", k);
  MK_REQUIRE(n >= 1 && ldv >= n, MK_BAD_SHAPE, "krylov_orthogonalize: n = %lld, ldv = %lld", (long long)n,
             (long long)ldv);
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  void* mem = nullptr;
  const size_t np = (size_t)std::max<int64_t>((int64_t)cgs_blocks(n) * 34, std::max<int64_t>(1, cdiv(n, 256 * 16)) * 34);
  MK_TRY(ctx->small((3 * 32 + 2 + np) * 8, &mem));
  double* cd = static_cast<double*>(mem);
  double* part = cd + 3 * 32 + 2;
  MK_TRY(k_multidot(V, ldv, k, u, n, cd, part, st));
  MK_TRY(k_cgs_update_dot(V, ldv, k, cd, u, n, cd + 32, part, st));
  MK_TRY(k_cgs_update_norm(V, ldv, k, cd + 32, u, u, n, cd + 64, part, st));
  std::vector<double> hv(66);
  MK_CHECK_HIP(hipMemcpyAsync(hv.data(), cd, hv.size() * 8, hipMemcpyDeviceToHost, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < k; ++i) hout[i] = hv[i] + hv[32 + i];
  hout[k] = std::sqrt(std::max(hv[64], 0.0));
  return MK_OK;
}

extern "C" mk_status mk_eigsolve(mk_op h, double* x, int krylovdim, int maxiter, double tol, int precision,
                                 double* eigval, double* resnorm, int* numops, int* converged) {
  MK_REQUIRE(h != nullptr && x != nullptr, MK_BAD_ARG, "eigsolve: null argument");
  MK_REQUIRE(krylovdim >= 2 && krylovdim <= 512, MK_BAD_ARG, "eigsolve: krylovdim %d not in [2, 512]", krylovdim);
  MK_REQUIRE(maxiter >= 1, MK_BAD_ARG, "eigsolve: maxiter must be >= 1");
  MK_REQUIRE(tol >= 0.0, MK_BAD_ARG, "eigsolve: negative tolerance");
  Op* op = reinterpret_cast<Op*>(h);
  MK_REQUIRE(op->x_elems == op->y_elems, MK_BAD_SHAPE, "eigsolve: operator is not square");
  MK_REQUIRE(op->precisions & (1 << precision), MK_BAD_ARG, "eigsolve: precision %d not prepared", precision);
  Ctx* ctx = op->ctx;
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t n = op->x_elems;
  const int K = krylovdim;                        // basis columns: x, kept Ritz vectors, Krylov vectors
  // Ritz vectors kept at a restart (x included).  A mixed-precision cycle costs
  // one float64 apply besides its low-precision Krylov steps, so it keeps fewer
  // (krylovdim / 4) and runs longer cycles.
  const int pmax = std::max(1, precision == MK_FP64 ? krylovdim / 2 : krylovdim / 4);

  // basis Q = [x, y_2..y_p, v_0, ...] (column-major n x K), then pmax columns for
  // the new Ritz vectors; w = H x, u = work
  DevBuf Qb, wb, ub, cb, pb, sb;
  MK_TRY(Qb.alloc((size_t)n * (K + pmax) * 8));
  MK_TRY(wb.alloc((size_t)n * 8));
  MK_TRY(ub.alloc((size_t)n * 8));
  // coefficients: generic dots at [0, K + 4); fused passes at [FC, FC + 2 * 32 + 2)
  constexpr int FC = 1024;
  MK_TRY(cb.alloc((size_t)(FC + 2 * 32 + 2) * 8));
  MK_TRY(sb.alloc((size_t)K * pmax * 8));
  const int64_t nb = std::max<int64_t>(1, cdiv(n, 256 * 16));
  MK_TRY(pb.alloc((size_t)std::max<int64_t>(nb * (K + 4), (int64_t)cgs_blocks(n) * 34) * 8));
  double* Q = static_cast<double*>(Qb.ptr);
  double* w = static_cast<double*>(wb.ptr);
  double* u = static_cast<double*>(ub.ptr);
  auto col = [&](int j) { return Q + (int64_t)j * n; };
  Dev dv{ctx, static_cast<double*>(cb.ptr), static_cast<double*>(pb.ptr)};

  int ops = 0;
  MK_CHECK_HIP(hipMemcpyAsync(col(0), x, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  double nrm = 0.0;
  MK_TRY(dv.norm(col(0), n, nrm));
  MK_REQUIRE(std::isfinite(nrm), MK_NONFINITE, "eigsolve: start vector is not finite");
  MK_REQUIRE(nrm > 0.0, MK_BAD_ARG, "eigsolve: zero start vector");
  MK_TRY(k_axpby(1.0 / nrm, col(0), 0.0, col(0), n, st));

  double theta = 0.0, res = INFINITY;
  double est_last = INFINITY;
  bool conv = false;
  const int ldg = K;
  int p = 1;                        // kept vectors [x, y_2..y_p] at the front of Q
  std::vector<double> lamk(1, 0.0); // Ritz values of the kept vectors
  std::vector<double> G((size_t)ldg * ldg), c, c2, s, cx, lamv, S;
  for (int outer = 0;; ++outer) {
    // float64 mode: accept the Ritz estimate of the previous cycle, as
    // KrylovKit does, which saves one apply
    if (precision == MK_FP64 && outer > 0 && est_last <= tol) {
      res = est_last;
      conv = true;
      break;
    }
    // float64 Ritz value and residual of x
    MK_TRY(op_apply(op, col(0), w, 1.0, 0.0, MK_FP64));
    ++ops;
    MK_TRY(dv.dots(Q, n, p, w, cx));          // <x, w> and <y_i, w>
    theta = cx[0];
    MK_TRY(k_axpby(1.0, w, 0.0, col(p), n, st));
    MK_TRY(k_axpby(-theta, col(0), 1.0, col(p), n, st));
    // re-orthogonalise the residual against x (float64): its norm is the residual of x
    MK_TRY(dv.dots(col(0), n, 1, col(p), c));
    MK_TRY(dv.subtract(col(0), n, 1, col(p)));
    MK_TRY(dv.norm(col(p), n, res));
    MK_REQUIRE(std::isfinite(theta) && std::isfinite(res), MK_NONFINITE, "eigsolve: non-finite Ritz pair");
    if (res <= tol) {
      conv = true;
      break;
    }
    if (outer == maxiter) break;
    // the Krylov space starts from the part of r outside the kept Ritz vectors
    double rperp = res;
    if (p > 1) {
      for (int pass = 0; pass < 2; ++pass) {
        MK_TRY(dv.dots(Q, n, p, col(p), c));
        MK_TRY(dv.subtract(Q, n, p, col(p)));
      }
      MK_TRY(dv.norm(col(p), n, rperp));
    }
    if (!(rperp > 1e-14 * std::max(std::fabs(theta), res))) {
      // r lies in the kept space: restart from x alone
      p = 1;
      MK_TRY(k_axpby(1.0, w, 0.0, col(1), n, st));
      MK_TRY(k_axpby(-theta, col(0), 1.0, col(1), n, st));
      MK_TRY(dv.dots(col(0), n, 1, col(1), c));
      MK_TRY(dv.subtract(col(0), n, 1, col(1)));
      rperp = res;
    }
    MK_TRY(k_axpby(1.0 / rperp, col(p), 0.0, col(p), n, st));

    std::fill(G.begin(), G.end(), 0.0);
    G[0] = theta;
    for (int i = 1; i < p; ++i) {
      G[(size_t)i * ldg + i] = lamk[i];
      G[i] = G[(size_t)i * ldg] = cx[i];       // <y_i, H x> in float64
    }
    G[p] = G[(size_t)p * ldg] = rperp;         // <v_0, H x> = |r_perp|
    int k = 1;                                 // Krylov vectors v_0..v_{k-1} in the cycle
    const double thresh = std::max(tol, inner_eta(precision) * res);
    const int mmax = K - p;                    // room for Krylov vectors
    for (int j = 0; j < mmax; ++j) {
      const int cj = p + j;                    // column of v_j
      // u = H_p v_j
      MK_TRY(op_apply(op, col(cj), u, 1.0, 0.0, precision));
      ++ops;
      // two classical Gram-Schmidt passes against [x, y.., v_0..v_j]
      double beta = 0.0, uw = 0.0;
      const int kq = cj + 1;
      if (kq <= 32) {
        // fused: c1 = Q^T u; u -= Q c1 with c2 = Q^T u in the same read of Q;
        // u -= Q c2 with |u|^2 and <w, u>; one readback for the step
        double* cd = dv.coef + FC;
        if (j >= 1 && local_first_pass()) {
          // first pass on the last three columns only.  For j >= 1, H v_j lies
          // in span{v_{j-1}, v_j, v_{j+1}} in exact arithmetic (x and the kept
          // Ritz vectors couple to v_0 alone), so the three-term recurrence
          // removes the large components; the full pass below removes what is
          // left, which is small, so the pair is as orthogonal as two full
          // passes (twice is enough) for one read of the columns instead of two
          MK_CHECK_HIP(hipMemsetAsync(cd, 0, (size_t)kq * 8, st));
          const int lo = std::max(0, cj - 2);
          MK_TRY(k_multidot(col(lo), n, cj + 1 - lo, u, n, cd + lo, dv.partial, st));
        } else {
          MK_TRY(k_multidot(Q, n, kq, u, n, cd, dv.partial, st));
        }
        MK_TRY(k_cgs_update_dot(Q, n, kq, cd, u, n, cd + 32, dv.partial, st));
        MK_TRY(k_cgs_update_norm(Q, n, kq, cd + 32, u, w, n, cd + 64, dv.partial, st));
        std::vector<double> hv(2 * 32 + 2);
        MK_CHECK_HIP(hipMemcpyAsync(hv.data(), cd, hv.size() * 8, hipMemcpyDeviceToHost, st));
        MK_CHECK_HIP(hipStreamSynchronize(st));
        c.assign(hv.begin(), hv.begin() + kq);
        c2.assign(hv.begin() + 32, hv.begin() + 32 + kq);
        beta = std::sqrt(std::max(hv[64], 0.0));
        uw = hv[65];
      } else {
        MK_TRY(dv.dots(Q, n, kq, u, c));
        MK_TRY(dv.subtract(Q, n, kq, u));
        MK_TRY(dv.dots(Q, n, kq, u, c2));
        MK_TRY(dv.subtract(Q, n, kq, u));
        MK_TRY(dv.norm(u, n, beta));
        std::vector<double> t;
        MK_TRY(dv.dots(w, n, 1, u, t));
        uw = t[0];
      }
      // projected matrix; the x-row stays float64 (<v_i, w>)
      for (int i = 1; i < kq; ++i) G[(size_t)i * ldg + cj] = G[(size_t)cj * ldg + i] = c[i] + c2[i];
      MK_REQUIRE(std::isfinite(beta), MK_NONFINITE, "eigsolve: non-finite Krylov vector");
      double lam;
      lowest(G, ldg, kq, lam, s);
      est_last = beta * std::fabs(s[cj]);
      k = j + 1;
      if (est_last <= thresh || beta <= 1e-14 * std::fabs(theta) || j + 1 == mmax) break;
      MK_TRY(k_axpby(1.0 / beta, u, 0.0, col(cj + 1), n, st));
      G[(size_t)(cj + 1) * ldg + cj] = G[(size_t)cj * ldg + cj + 1] = beta;
      // <v_{j+1}, w> = <u, w> / beta
      G[(size_t)(cj + 1)] = G[(size_t)(cj + 1) * ldg] = uw / beta;
    }
    // the lowest Ritz vectors of the cycle's space: x and the kept vectors of
    // the next cycle
    const int kk = p + k;
    const int pn = std::min(pmax, kk);
    lowest_p(G, ldg, kk, pn, lamv, S);
    MK_CHECK_HIP(hipMemcpyAsync(sb.ptr, S.data(), S.size() * 8, hipMemcpyHostToDevice, st));
    ProbSpec ps;
    ps.M = n; ps.N = pn; ps.C = col(K); ps.ldc = n;
    ps.add(Q, n, sb.ptr, kk, kk);
    MK_TRY(gemm_f64(&ps, 1, GemmShape{}, st));
    MK_CHECK_HIP(hipMemcpyAsync(col(0), col(K), (size_t)n * pn * 8, hipMemcpyDeviceToDevice, st));
    MK_TRY(dv.norm(col(0), n, nrm));
    MK_TRY(k_axpby(1.0 / nrm, col(0), 0.0, col(0), n, st));
    p = pn;
    lamk = lamv;
    theta = lamv[0];
  }
  MK_CHECK_HIP(hipMemcpyAsync(x, col(0), (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  if (eigval) *eigval = theta;
  if (resnorm) *resnorm = res;
  if (numops) *numops = ops;
  if (converged) *converged = conv ? 1 : 0;
  return MK_OK;
}
