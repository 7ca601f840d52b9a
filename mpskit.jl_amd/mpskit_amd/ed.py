"""exact_diagonalization of a FiniteMPOHamiltonian on the HIP path.

Mirrors src/algorithms/ED.jl:1-90 (paths relative to the MPSKit.jl tree): the
chain is written as an MPS of maximal rank.  Isometries fuse the physical legs
from the left edge up to the middle site and from the right edge down to it, so
the one-site derivative at the middle site acts on the full Hilbert space and
its lowest eigenvectors are the exact eigenstates.  The fusing tensors start
as identities (fused index l + D_l p on the left, p + d r on the right) and
are brought to mixed-canonical form, which keeps them full-rank isometries.
The lowest state comes from the library's eigensolver
(``mk_eigsolve``); further states come from the same derivative with the
states already found moved to the top of the spectrum, P H P + sigma U U^T with
P = 1 - U U^T and sigma above the largest eigenvalue (a Lanczos estimate of it,
plus a margin).  Shifting them to the bottom would fail when the ground energy
lies near zero or the gap is large, and projecting alone is not enough:
round-off leaves components along U in the Krylov vectors, and Lanczos's
recurrence amplifies a component whose eigenvalue lies below the rest (the
found states of P H P sit at 0) until the solver returns a state already
found.  At the top of the spectrum the same components are harmless to a
search for the lowest eigenvalue.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import tensors as T
from .derivatives import JordanMPO_AC_Hamiltonian, prepare_operator
from .environments import FiniteEnvironments
from .excited import lanczos_min
from .states import FiniteMPS


def _fusing_tensors(H, d, seed):
    L = len(H)
    mid = L >> 1
    tensors = [None] * L
    Dl = 1
    for i in range(mid):
        A = np.zeros((Dl, d, Dl * d))
        for p in range(d):
            for l in range(Dl):
                A[l, p, l + Dl * p] = 1.0
        tensors[i] = A
        Dl *= d
    Dr = 1
    for i in range(L - 1, mid, -1):
        A = np.zeros((d * Dr, d, Dr))
        for p in range(d):
            for r in range(Dr):
                A[p + d * r, p, r] = 1.0
        tensors[i] = A
        Dr *= d
    tensors[mid] = np.random.default_rng(seed).standard_normal((Dl, d, Dr))
    return [T.from_numpy(t) for t in tensors], mid


def _project(v, found):
    """v - U U^T v for the orthonormal found states U (two Gram-Schmidt passes)."""
    for _ in range(2):
        for u in found:
            v = v - torch.dot(u.reshape(-1), v.reshape(-1)) * u
    return v


def _above_spectrum(op, x0, e0):
    """A shift above the spectrum of op: the largest Ritz value of a short
    Lanczos run on -op (a lower bound of the largest eigenvalue, close to it
    since it is extreme), plus a margin of the spread and 1."""
    lam, _, _ = lanczos_min(lambda v: -op(T.as_colmajor(v)), T.as_colmajor(x0), tol=1e-6, krylovdim=30, maxiter=3)
    top = -lam
    return top + 0.1 * abs(top - e0) + 1.0


def exact_diagonalization(H, num=1, tol=1e-12, krylovdim=30, maxiter=200, seed=0):
    """The ``num`` lowest eigenpairs of a finite chain as (values, FiniteMPS
    states) (exact_diagonalization, ED.jl:1-90, which = :SR).  All states share
    the fused basis of one maximal-rank MPS and differ in the middle tensor."""
    d = int(H.host[0].shape[1])
    tensors, mid = _fusing_tensors(H, d, seed)
    base = FiniteMPS(tensors, center=mid)
    envs = FiniteEnvironments(base, H)
    op = prepare_operator(JordanMPO_AC_Hamiltonian(envs.leftenv(mid, base), H[mid], envs.rightenv(mid, base)))
    vals, states, found = [], [], []
    sigma = None
    try:
        for _ in range(num):
            if not found:
                lam, x, _ = op.eigsolve(base.AC, krylovdim=krylovdim, maxiter=maxiter, tol=tol)
            else:
                if sigma is None:
                    sigma = _above_spectrum(op, base.AC, vals[0])

                def apply(v):
                    y = _project(op(T.as_colmajor(_project(v, found))), found)
                    for u in found:
                        y = y + (sigma * torch.dot(u.reshape(-1), v.reshape(-1))) * u
                    return y

                start = _project(base.AC, found)
                lam, x, _ = lanczos_min(apply, T.as_colmajor(start), tol=tol, krylovdim=krylovdim, maxiter=maxiter)
                x = _project(x, found)
            x = T.as_colmajor(x / torch.linalg.vector_norm(x))
            found.append(x)
            vals.append(float(lam))
            st = copy.copy(base)
            st.AL, st.AR, st.AC = list(base.AL), list(base.AR), x
            states.append(st)
    finally:
        op.free()
    return vals, states
