"""Infinite DMRG (IDMRG, IDMRG2) on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  IDMRG settings           src/algorithms/groundstate/idmrg.jl:13-31
  driver / convergence     src/algorithms/groundstate/idmrg.jl:94-170
  one left + right sweep   src/algorithms/groundstate/idmrg.jl:196-235
  two-site sweep (IDMRG2)  src/algorithms/groundstate/idmrg.jl:237-338

The unit cell is swept left to right and back: every site solves its one-site
problem with the Jordan-form dAC operator of the current environments (the
HIP eigensolver, any apply precision with float64 residuals), re-orthogonalises
with a positive QR (left sweep) or LQ (right sweep), and pushes its
environment one site further with one MPO transfer application.  The
environments are not re-solved as fixed points between sweeps; they grow with
every sweep, as in the reference.  Convergence is the change of the bond
matrix C at the cell boundary between sweeps.  The result is regauged into a
uniform MPS from the right-orthonormal tensors, with environments recomputed
as fixed points.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import torch

from . import tensors as T
from .derivatives import JordanMPO_AC2_Hamiltonian, JordanMPO_AC_Hamiltonian, prepare_operator
from .states import lq_right, qr_left
from .transfer import transfer_left, transfer_right
from .vumps import (InfiniteEnvironments, InfiniteHamiltonian, InfiniteMPS, _norm, calc_galerkin,
                    energy_per_site)


@dataclass
class IDMRG:
    """``IDMRG`` settings (idmrg.jl:13-31); ``precision`` applies to the local
    eigensolver applies.  The local tolerance follows the global error as the
    reference's adaptive eigensolver does: clamp(tol_factor * eps, tol_min, tol_max)."""
    tol: float = 1e-10
    maxiter: int = 200
    verbosity: int = 1
    precision: str = "fp64"
    tol_factor: float = 0.1
    tol_min: float = 1e-13
    tol_max: float = 1e-4
    krylovdim: int = 20


def _solve(GL, W, GR, AC, tol, alg, precs):
    op = prepare_operator(JordanMPO_AC_Hamiltonian(GL, W, GR), precisions=precs)
    E, x, info = op.eigsolve(AC, krylovdim=alg.krylovdim, maxiter=20, tol=tol, precision=alg.precision)
    op.free()
    return E, x, info


def find_groundstate_idmrg(psi: InfiniteMPS, H: InfiniteHamiltonian, alg: IDMRG,
                           envs: InfiniteEnvironments | None = None):
    """``find_groundstate(psi::InfiniteMPS, H, IDMRG())`` (idmrg.jl:94-170).

    Returns (psi, envs, eps, history); history holds per iteration the change of
    the boundary bond matrix, the energy change per site of the growing
    effective problem, the wall time and the applies."""
    N = psi.N
    envs = envs or InfiniteEnvironments(psi, H, tol=1e-12)
    AL, AR, C, AC = list(psi.AL), list(psi.AR), list(psi.C), list(psi.AC)
    GL, GR = list(envs.GLs), list(envs.GRs)
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    eps = calc_galerkin(psi, H, envs)
    E_old = None
    history = []
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tol = min(max(alg.tol_factor * eps, alg.tol_min), alg.tol_max)
        C_old = C[N - 1]
        numops = 0
        # left to right (idmrg.jl:202-217)
        for pos in range(N):
            _, AC[pos], info = _solve(GL[pos], H.W[pos], GR[pos], AC[pos], tol, alg, precs)
            numops += info["numops"]
            AL[pos], C[pos] = qr_left(AC[pos])
            GL[(pos + 1) % N] = transfer_left(GL[pos], H.W[pos], AL[pos], AL[pos])
        # right to left (idmrg.jl:219-232)
        E = 0.0
        for pos in range(N - 1, -1, -1):
            E, AC[pos], info = _solve(GL[pos], H.W[pos], GR[pos], AC[pos], tol, alg, precs)
            numops += info["numops"]
            Lm, AR[pos] = lq_right(AC[pos])
            C[(pos - 1) % N] = Lm
            GR[(pos - 1) % N] = transfer_right(GR[pos], H.W[pos], AR[pos], AR[pos])
        # error criterion and energy change (idmrg.jl:142-160)
        if C[N - 1].shape == C_old.shape:
            eps = float(_norm(C[N - 1] - C_old))
        else:
            k = min(C_old.shape[0], C[N - 1].shape[0])
            eps = float(_norm(C[N - 1][:k, :k] - C_old[:k, :k]))
        dE = float("nan") if E_old is None else (E - E_old) / 2
        E_old = E
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        history.append(dict(iter=it, eps=eps, dE=dE, time=dt, numops=numops))
        if alg.verbosity >= 2:
            print(f"IDMRG {it:3d}: err = {eps:.3e}  dE = {dE:.12f}  time = {dt:.3f}s  applies = {numops}",
                  flush=True)
        if eps <= alg.tol:
            break
    # uniform MPS from the right-orthonormal tensors (idmrg.jl:162-166)
    psi = InfiniteMPS.from_tensors(AR)
    envs = InfiniteEnvironments(psi, H, tol=1e-12)
    if history:
        history[-1]["energy"] = energy_per_site(psi, H, envs)
    return psi, envs, eps, history


# ---------------------------------------------------------------------------
# two-site IDMRG
# ---------------------------------------------------------------------------

@dataclass
class IDMRG2:
    """``IDMRG2`` settings (idmrg.jl:47-68): the two-site update truncates the
    bond to ``trunc_dim`` states (and, if ``trunc_err`` > 0, drops the tail of
    singular values whose 2-norm is below it).  Singular values below
    ``rank_tol`` times the largest are always dropped, so that every bond matrix
    stays invertible for the edge steps."""
    trunc_dim: int = 64
    trunc_err: float = 0.0
    rank_tol: float = 1e-13
    tol: float = 1e-10
    maxiter: int = 200
    verbosity: int = 1
    precision: str = "fp64"
    tol_factor: float = 0.1
    tol_min: float = 1e-13
    tol_max: float = 1e-4
    krylovdim: int = 20


def _ac2(A, B):
    """Two-site tensor x[f,h,r,j] = sum_m A[f,h,m] B[m,j,r]."""
    return T.as_colmajor(torch.einsum("fhm,mjr->fhrj", A, B))


def _svd_trunc(x, alg: IDMRG2):
    """al, c, ar of the truncated SVD of the fused two-site tensor (svd_trunc!,
    idmrg.jl:255-262), with c = diag(S) normalised."""
    Dl, d1, Dr, d2 = x.shape
    U, S, Vh = torch.linalg.svd(T.fuse(x, 2), full_matrices=False)
    k = min(alg.trunc_dim, int(S.numel()), max(1, int((S > alg.rank_tol * S[0]).sum().item())))
    if alg.trunc_err > 0.0:
        tail = torch.sqrt(torch.flip(torch.cumsum(torch.flip(S * S, [0]), 0), [0]))
        k = max(1, min(k, int((tail > alg.trunc_err).sum().item())))
    U, S, Vh = U[:, :k], S[:k], Vh[:k]
    al = T.as_colmajor(T.unfuse(U, (Dl, d1, k)))
    ar = T.as_colmajor(T.unfuse(Vh, (k, Dr, d2)).permute(0, 2, 1))
    return al, torch.diag(S / _norm(S)), ar


def _rdiv(A, C):
    """A C^{-1} on the right bond of A[l,p,r]."""
    l, p, r = A.shape
    X = torch.linalg.solve(C.t(), T.fuse(A, 2).t()).t()
    return T.as_colmajor(T.unfuse(X, (l, p, C.shape[0])))


def _ldiv(C, A):
    """C^{-1} A on the left bond of A[l,p,r]."""
    l, p, r = A.shape
    X = torch.linalg.solve(C, T.fuse(A, 1))
    return T.as_colmajor(T.unfuse(X, (C.shape[1], p, r)))


def find_groundstate_idmrg2(psi: InfiniteMPS, H: InfiniteHamiltonian, alg: IDMRG2,
                            envs: InfiniteEnvironments | None = None):
    """``find_groundstate(psi::InfiniteMPS, H, IDMRG2(...))`` (idmrg.jl:94-170,
    sweep :237-338): two-site updates inside the cell, the cell edge updated
    across the boundary after each half sweep.  Returns (psi, envs, eps, history)."""
    from .vumps import _ctimes, _timesc
    N = psi.N
    if N < 2:
        raise ValueError("IDMRG2: unit cell should be >= 2")
    envs = envs or InfiniteEnvironments(psi, H, tol=1e-12)
    AL, AR, C, AC = list(psi.AL), list(psi.AR), list(psi.C), list(psi.AC)
    GL, GR = list(envs.GLs), list(envs.GRs)
    W = H.W
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    eps = calc_galerkin(psi, H, envs)
    E_old = None
    history = []

    def solve2(gl, w1, w2, gr, x, tol):
        op = prepare_operator(JordanMPO_AC2_Hamiltonian(gl, w1, w2, gr), precisions=precs)
        E, y, info = op.eigsolve(x, krylovdim=alg.krylovdim, maxiter=20, tol=tol, precision=alg.precision)
        op.free()
        return E, y, info["numops"]

    def envs_at(pos):   # GL right of site pos, GR left of site pos + 1 (idmrg.jl transfer_env steps)
        q = (pos + 1) % N
        GL[q] = transfer_left(GL[pos], W[pos], AL[pos], AL[pos])
        GR[pos] = transfer_right(GR[q], W[q], AR[q], AR[q])

    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tol = min(max(alg.tol_factor * eps, alg.tol_min), alg.tol_max)
        numops = 0
        # left to right inside the cell (idmrg.jl:243-262)
        for pos in range(N - 1):
            _, y, n = solve2(GL[pos], W[pos], W[pos + 1], GR[pos + 1], _ac2(AC[pos], AR[pos + 1]), tol)
            numops += n
            al, c, ar = _svd_trunc(y, alg)
            AL[pos], C[pos], AR[pos + 1] = al, c, ar
            AC[pos + 1] = _ctimes(c, ar)
            envs_at(pos)
        # the edge (idmrg.jl:264-290)
        AL[N - 1] = _rdiv(AC[N - 1], C[N - 1])
        AC[0] = _timesc(AL[0], C[0])
        _, y, n = solve2(GL[N - 1], W[N - 1], W[0], GR[0], _ac2(AL[N - 1], AC[0]), tol)
        numops += n
        al, c, ar = _svd_trunc(y, alg)
        AL[N - 1], C[N - 1], AR[0] = al, c, ar
        AC[N - 1] = _timesc(al, c)
        AC[0] = _ctimes(c, ar)
        AL[0] = _rdiv(AC[0], C[0])
        C_old = c
        envs_at(N - 1)
        # right to left inside the cell (idmrg.jl:292-312)
        for pos in range(N - 2, -1, -1):
            _, y, n = solve2(GL[pos], W[pos], W[pos + 1], GR[pos + 1], _ac2(AL[pos], AC[pos + 1]), tol)
            numops += n
            al, c, ar = _svd_trunc(y, alg)
            AL[pos], C[pos], AR[pos + 1] = al, c, ar
            AC[pos] = _timesc(al, c)
            AC[pos + 1] = _ctimes(c, ar)
            envs_at(pos)
        # the edge (idmrg.jl:314-336)
        AC[N - 1] = _ctimes(C[N - 2], AR[N - 1])
        AR[0] = _ldiv(C[N - 1], AC[0])
        E, y, n = solve2(GL[N - 1], W[N - 1], W[0], GR[0], _ac2(AC[N - 1], AR[0]), tol)
        numops += n
        al, c, ar = _svd_trunc(y, alg)
        AL[N - 1], C[N - 1], AR[0] = al, c, ar
        AR[N - 1] = _ldiv(C[N - 2], _timesc(al, c))
        AC[0] = _ctimes(c, ar)
        envs_at(N - 1)
        # error criterion on the common block of the boundary bond (idmrg.jl:142-153)
        k = min(C_old.shape[0], C[N - 1].shape[0])
        eps = float(_norm(C[N - 1][:k, :k] - C_old[:k, :k]))
        dE = float("nan") if E_old is None else (E - E_old) / 2
        if N == 2:
            dE /= 2
        E_old = E
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        history.append(dict(iter=it, eps=eps, dE=dE, time=dt, numops=numops, D=int(C[N - 1].shape[0])))
        if alg.verbosity >= 2:
            print(f"IDMRG2 {it:3d}: err = {eps:.3e}  dE = {dE:.12f}  D = {C[N - 1].shape[0]}  time = {dt:.3f}s  "
                  f"applies = {numops}", flush=True)
        if eps <= alg.tol:
            break
    psi = InfiniteMPS.from_tensors(AR)
    envs = InfiniteEnvironments(psi, H, tol=1e-12)
    if history:
        history[-1]["energy"] = energy_per_site(psi, H, envs)
    return psi, envs, eps, history
