// Eigen-decomposition of a small dense symmetric matrix on the host, for the
// projected matrices of the Krylov solvers (at most a few hundred rows; every
// Krylov step solves one).
//
// Householder reduction to tridiagonal form, then implicit symmetric QR steps
// with Wilkinson shifts (Golub & Van Loan, Matrix Computations, 4th ed.,
// Algorithms 8.3.1 and 8.3.3); the orthogonal transformations are accumulated,
// so the columns of Q are the eigenvectors.  About 10 n^3 flops, against the
// 20-40 n^3 of cyclic Jacobi sweeps.  Header only, no dependencies, so that the
// CPU tests can build it on its own.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace mk {

// A: row-major n x n symmetric matrix (destroyed).  On return ev holds the
// eigenvalues in ascending order and Q (row-major n x n) the eigenvectors as
// columns, in the same order.
inline void sym_eig(std::vector<double> A, int n, std::vector<double>& ev, std::vector<double>& Q) {
  auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  auto q = [&](int i, int j) -> double& { return Q[(size_t)i * n + j]; };
  // symmetrise against round-off in the caller's assembly
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) a(i, j) = a(j, i) = 0.5 * (a(i, j) + a(j, i));
  // 1. tridiagonalisation: A <- H A H, Q <- Q H with H = I - beta v v^T acting on k+1..n-1
  std::vector<double> v(n), s(n);
  for (int k = 0; k + 2 < n; ++k) {
    double nx = 0.0;
    for (int i = k + 1; i < n; ++i) nx += a(i, k) * a(i, k);
    nx = std::sqrt(nx);
    if (nx == 0.0) continue;
    const double alpha = a(k + 1, k) > 0 ? -nx : nx;
    double vv = 0.0;
    for (int i = k + 1; i < n; ++i) {
      v[i] = a(i, k) - (i == k + 1 ? alpha : 0.0);
      vv += v[i] * v[i];
    }
    if (vv == 0.0) continue;
    const double beta = 2.0 / vv;
    // left: rows k+1..n-1 of every column
    for (int j = 0; j < n; ++j) {
      double t = 0.0;
      for (int i = k + 1; i < n; ++i) t += v[i] * a(i, j);
      t *= beta;
      for (int i = k + 1; i < n; ++i) a(i, j) -= t * v[i];
    }
    // right: columns k+1..n-1 of every row; and the accumulated Q
    for (int i = 0; i < n; ++i) {
      double t = 0.0, u = 0.0;
      for (int j = k + 1; j < n; ++j) {
        t += a(i, j) * v[j];
        u += q(i, j) * v[j];
      }
      t *= beta;
      u *= beta;
      for (int j = k + 1; j < n; ++j) {
        a(i, j) -= t * v[j];
        q(i, j) -= u * v[j];
      }
    }
  }
  // keep the tridiagonal part only (the rest is round-off)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (j < i - 1 || j > i + 1) a(i, j) = 0.0;
  // 2. implicit symmetric QR with Wilkinson shifts on the unreduced blocks,
  // each rotation applied as an explicit similarity on the band window
  const double eps = 2.2e-16;
  int m = n - 1;
  int iter = 0;
  while (m > 0 && iter < 60 * n) {
    // deflate from the bottom
    if (std::fabs(a(m, m - 1)) <= eps * (std::fabs(a(m, m)) + std::fabs(a(m - 1, m - 1))) ||
        std::fabs(a(m, m - 1)) < 1e-300) {
      a(m, m - 1) = a(m - 1, m) = 0.0;
      --m;
      continue;
    }
    int l = m - 1;
    while (l > 0 && !(std::fabs(a(l, l - 1)) <= eps * (std::fabs(a(l, l)) + std::fabs(a(l - 1, l - 1))) ||
                      std::fabs(a(l, l - 1)) < 1e-300))
      --l;
    if (l > 0) a(l, l - 1) = a(l - 1, l) = 0.0;
    ++iter;
    // Wilkinson shift from the trailing 2 x 2 block
    const double dd = 0.5 * (a(m - 1, m - 1) - a(m, m));
    const double e2 = a(m, m - 1) * a(m, m - 1);
    const double den = dd + (dd >= 0 ? 1.0 : -1.0) * std::sqrt(dd * dd + e2);
    const double mu = den != 0.0 ? a(m, m) - e2 / den : a(m, m);
    double x = a(l, l) - mu, z = a(l + 1, l);
    for (int k = l; k < m; ++k) {
      const double r = std::hypot(x, z);
      if (r == 0.0) break;
      const double c = x / r, sn = z / r;
      const int lo = std::max(l, k - 1), hi = std::min(m, k + 2);
      for (int j = lo; j <= hi; ++j) {
        const double t1 = a(k, j), t2 = a(k + 1, j);
        a(k, j) = c * t1 + sn * t2;
        a(k + 1, j) = -sn * t1 + c * t2;
      }
      for (int i = lo; i <= hi; ++i) {
        const double t1 = a(i, k), t2 = a(i, k + 1);
        a(i, k) = c * t1 + sn * t2;
        a(i, k + 1) = -sn * t1 + c * t2;
      }
      for (int i = 0; i < n; ++i) {
        const double t1 = q(i, k), t2 = q(i, k + 1);
        q(i, k) = c * t1 + sn * t2;
        q(i, k + 1) = -sn * t1 + c * t2;
      }
      if (k + 1 < m) {
        x = a(k + 1, k);
        z = a(k + 2, k);
      }
    }
  }
  // ascending order
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int i, int j) { return a(i, i) < a(j, j); });
  ev.resize(n);
  std::vector<double> Qs((size_t)n * n);
  for (int c = 0; c < n; ++c) {
    ev[c] = a(idx[c], idx[c]);
    for (int i = 0; i < n; ++i) Qs[(size_t)i * n + c] = q(i, idx[c]);
  }
  Q.swap(Qs);
}

}  // namespace mk
