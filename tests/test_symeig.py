"""The host eigensolver of the Krylov solvers' projected matrices
(``csrc/symeig.h``: Householder tridiagonalisation and implicit QR with
Wilkinson shifts) against numpy's eigh, built on its own with g++: random
dense, tridiagonal (Lanczos), arrowhead plus tridiagonal (thick restart),
degenerate, diagonal and badly scaled matrices, sizes 1 to 200.  Eigenvalues
ascending to 1e-13 of the norm, eigenvectors orthonormal to 1e-13 with
residuals below 1e-13 of the norm (eigenvectors are compared through residuals:
degenerate ones are a basis choice)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HARNESS = r"""
#include "symeig.h"
#include <cstdio>
int main() {
  int n;
  while (std::scanf("%d", &n) == 1) {
    std::vector<double> A((size_t)n * n);
    for (auto& v : A) std::scanf("%lf", &v);
    std::vector<double> ev, Q;
    mk::sym_eig(A, n, ev, Q);
    for (int i = 0; i < n; ++i) std::printf("%.17g\n", ev[i]);
    for (auto v : Q) std::printf("%.17g\n", v);
  }
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("symeig")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(ROOT / "mpskit.jl_amd" / "csrc"), str(src), "-o", str(exe)],
                   check=True)
    return exe


def cases():
    rng = np.random.default_rng(7)
    out = []
    for n in (1, 2, 3, 7, 30, 64, 200):
        A = rng.standard_normal((n, n))
        out.append(("dense", A + A.T))
        T = np.diag(rng.standard_normal(n)) + np.diag(rng.random(n - 1), 1) + np.diag(rng.random(n - 1), -1) \
            if n > 1 else rng.standard_normal((1, 1))
        out.append(("tridiagonal", (T + T.T) / 2))
        if n > 3:
            p = n // 3
            M = np.zeros((n, n))
            M[:p, :p] = np.diag(np.sort(rng.standard_normal(p)))
            M[0, :p] = M[:p, 0] = rng.standard_normal(p)
            M[p:, p:] = T[p:, p:]
            M[:p, p] = M[p, :p] = rng.standard_normal(p)
            out.append(("arrowhead", (M + M.T) / 2))
        out.append(("degenerate", np.diag(np.repeat(np.arange((n + 2) // 3), 3)[:n].astype(float))))
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        ev = np.repeat([-1.0, 2.0], (n + 1) // 2)[:n]
        out.append(("degenerate rotated", Q @ np.diag(ev) @ Q.T))
        out.append(("scaled 1e-8", 1e-8 * (A + A.T)))
        out.append(("scaled 1e8", 1e8 * (A + A.T)))
    return out


def test_sym_eig_against_numpy(harness):
    cs = cases()
    inp = "".join(f"{A.shape[0]}\n" + " ".join(f"{v:.17g}" for v in A.ravel()) + "\n" for _, A in cs)
    out = subprocess.run([str(harness)], input=inp, capture_output=True, text=True, check=True).stdout.split()
    vals = np.array(out, dtype=float)
    o = 0
    for name, A in cs:
        n = A.shape[0]
        ev = vals[o:o + n]
        o += n
        Q = vals[o:o + n * n].reshape(n, n)
        o += n * n
        nrm = max(np.abs(A).max(), 1e-300)
        ref = np.linalg.eigvalsh(A)
        assert np.all(np.diff(ev) >= 0), name
        assert np.abs(ev - ref).max() <= 1e-13 * nrm * max(1, n / 10), (name, n, np.abs(ev - ref).max())
        assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-13 * max(1, n / 10), (name, n)
        assert np.abs(A @ Q - Q * ev).max() <= 1e-13 * nrm * max(1, n / 10), (name, n)
    assert o == vals.size
