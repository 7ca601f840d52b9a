"""Dense finite MPS and MPO Hamiltonians held on the GPU.

``FiniteMPS`` keeps the mixed-canonical form MPSKit's ``FiniteMPS`` exposes
through its gauge views (src/states/finitemps.jl, orthoview.jl): left-orthonormal
``AL[j]`` for j < center, right-orthonormal ``AR[j]`` for j > center and the
center tensor ``AC``.  Tensors are column-major float64 GPU tensors
``A[l, p, r]``.  QR and SVD of the gauge steps run in float64 on the GPU
(torch.linalg -> rocSOLVER).
"""
from __future__ import annotations

import numpy as np
import torch

from . import tensors as T


class FiniteMPOHamiltonian:
    """Dense Jordan-form MPO tensors ``W[g, b, h, k]`` of a finite chain
    (``FiniteMPOHamiltonian``, mpohamiltonian.jl:439-519): first tensor has one
    row, last tensor one column."""

    def __init__(self, Ws):
        self.host = [np.asarray(W, dtype=np.float64) for W in Ws]
        self.W = [T.from_numpy(W) for W in self.host]

    def __len__(self):
        return len(self.W)

    def __getitem__(self, j):
        return self.W[j]

    def left_virtualdim(self, j):
        return self.host[j].shape[0]

    def right_virtualdim(self, j):
        return self.host[j].shape[3]


def _mat(A, rows):
    return T.fuse(A, rows)


_EPS = float(np.finfo(np.float64).eps) / 2   # unit roundoff


def _householder_pos_qr(M):
    Q, R = torch.linalg.qr(M)
    s = torch.sign(torch.diagonal(R))
    s[s == 0] = 1
    return Q * s, s[:, None] * R


def _chol_step(Q, shift=0.0):
    """One Cholesky QR pass: G = Q^T Q (+ shift I) = R^T R, Q <- Q R^-1.  The
    factorisation is the library's blocked Cholesky (``mk_potrf``)."""
    G = (Q.t() @ Q).contiguous()
    if shift:
        G.diagonal().add_(shift)
    if not G.is_cuda:
        # host tensors (the CPU tests of this QR's logic)
        return torch.linalg.cholesky_ex(G)
    from .krylov import potrf_
    return potrf_(G)


def pos_qr(M):
    """Thin QR of an m x n float64 GPU matrix (m >= n) with a positive diagonal
    in R (the qr_compact(; positive = true) of the reference's gauge steps,
    ortho.jl:307-340), by shifted Cholesky QR3 (Fukaya, Kannan, Nakatsukasa,
    Zhang, Yamamoto, SIAM J. Sci. Comput. 42 (2020) A477): a shifted Cholesky
    QR pass, then two plain ones, all GEMM / POTRF / TRSM work.  A factor whose
    Cholesky fails, or whose Q is not orthonormal to 1e-13 per column, is
    recomputed by Householder QR (rocSOLVER geqrf / orgqr).  The positive-R QR
    of a full-rank matrix is unique, so both routes give the same factors."""
    m, n = M.shape
    if m < n or n == 0:
        return _householder_pos_qr(M)
    fro2 = float(torch.sum(M * M))
    if fro2 == 0.0 or not np.isfinite(fro2):
        return _householder_pos_qr(M)
    shift = 11.0 * (m * n + n * (n + 1)) * _EPS * fro2
    Q = M
    R = None
    for step in range(3):
        L, info = _chol_step(Q, shift if step == 0 else 0.0)
        if int(info) != 0:
            return _householder_pos_qr(M)
        Rk = L.t()
        Q = torch.linalg.solve_triangular(Rk, Q, upper=True, left=False)
        R = Rk if R is None else Rk @ R
    E = Q.t() @ Q
    E.diagonal().sub_(1.0)
    if float(torch.linalg.matrix_norm(E)) > 1e-13 * max(n, 1) ** 0.5:
        return _householder_pos_qr(M)
    return Q, torch.triu(R)


def qr_left(A, householder=False):
    """A[l,p,r] = Q[l,p,m] R[m,r] with positive diagonal R (qr_compact(positive=true));
    ``householder`` skips the Cholesky QR attempt (for inputs known to be rank deficient)."""
    l, p, r = A.shape
    Q, R = (_householder_pos_qr if householder else pos_qr)(_mat(A, 2))
    return T.as_colmajor(T.unfuse(Q, (l, p, Q.shape[1]))), R


def lq_right(A, householder=False):
    """A[l,p,r] = L[l,m] Q[m,p,r] with Q right-orthonormal; ``householder`` as in qr_left."""
    l, p, r = A.shape
    Qt, Rt = (_householder_pos_qr if householder else pos_qr)(_mat(A, 1).t())
    return Rt.t(), T.as_colmajor(T.unfuse(Qt.t(), (Qt.shape[1], p, r)))


class FiniteMPS:
    """Mixed-canonical finite MPS on the GPU."""

    def __init__(self, tensors, center: int | None = None):
        L = len(tensors)
        self.L = L
        A = [T.as_colmajor(t) for t in tensors]
        # left-canonicalise everything, then move the center back (finitemps.jl:229-254)
        self.AL = [None] * L
        self.AR = [None] * L
        R = torch.ones((1, 1), dtype=torch.float64, device=A[0].device)
        for j in range(L - 1):
            Aj = T.as_colmajor(torch.einsum("lm,mpr->lpr", R, A[j]))
            self.AL[j], R = qr_left(Aj)
        AC = T.as_colmajor(torch.einsum("lm,mpr->lpr", R, A[L - 1]))
        AC = AC / torch.linalg.vector_norm(AC)
        self.AC = T.as_colmajor(AC)
        self.center = L - 1
        if center is not None:
            self.move_center(center)

    @classmethod
    def random(cls, L, d, Dmax, seed=0, device=None):
        """i.i.d. N(0,1) tensors with bonds min(d^k, d^(L-k), Dmax) from a seeded generator,
        normalised left to right (finitemps.jl:256-283)."""
        rng = np.random.default_rng(seed)
        dims = [min(d ** k, d ** (L - k), Dmax) for k in range(L + 1)]
        tensors = [T.from_numpy(rng.standard_normal((dims[k], d, dims[k + 1])), device) for k in range(L)]
        return cls(tensors)

    def bond_dims(self):
        dims = [1]
        for j in range(self.L):
            dims.append(self.tensor_at(j).shape[2])
        return dims

    def tensor_at(self, j):
        if j < self.center:
            return self.AL[j]
        if j > self.center:
            return self.AR[j]
        return self.AC

    def move_center(self, target):
        while self.center < target:
            j = self.center
            self.AL[j], R = qr_left(self.AC)
            self.AC = T.as_colmajor(torch.einsum("lm,mpr->lpr", R, self.AR[j + 1]))
            self.AR[j + 1] = None
            self.center += 1
        while self.center > target:
            j = self.center
            Lm, self.AR[j] = lq_right(self.AC)
            self.AC = T.as_colmajor(torch.einsum("lpm,mr->lpr", self.AL[j - 1], Lm))
            self.AL[j - 1] = None
            self.center -= 1

    def AC2(self, pos):
        """Two-site center at (pos, pos+1) with the center at pos (:ACAR) or pos+1 (:ALAC),
        ``AC2(psi, site; kind)`` (finitemps.jl:376-384): x[f,h,i,j]."""
        if self.center == pos:
            A, B = self.AC, self.AR[pos + 1]
        elif self.center == pos + 1:
            A, B = self.AL[pos], self.AC
        else:
            raise ValueError("AC2: center must sit on the bond")
        return T.as_colmajor(torch.einsum("fhm,mjr->fhrj", A, B))

    def to_dense(self):
        psi = self.tensor_at(0)[0]
        for j in range(1, self.L):
            A = self.tensor_at(j)
            psi = torch.einsum("Pm,mpr->Ppr", psi, A).reshape(-1, A.shape[2])
        return psi[:, 0]
