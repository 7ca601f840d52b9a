"""Single-site infinite DMRG (IDMRG) on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  IDMRG settings           src/algorithms/groundstate/idmrg.jl:13-31
  driver / convergence     src/algorithms/groundstate/idmrg.jl:94-170
  one left + right sweep   src/algorithms/groundstate/idmrg.jl:196-235

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

from .derivatives import JordanMPO_AC_Hamiltonian, prepare_operator
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
