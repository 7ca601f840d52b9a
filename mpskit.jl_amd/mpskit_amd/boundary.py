"""Leading boundary MPS of an infinite (non-Hamiltonian) MPO on the HIP path:
VUMPS and VOMPS for ``leading_boundary``.

Mirrors (paths relative to the MPSKit.jl tree):
  InfiniteMPO environments   src/environments/infinite_envs.jl:115-175
                             (dominant left / right eigenvectors of the mixed
                             transfer matrix, pushed through the unit cell;
                             GR of norm 1, GL scaled to <C|dC(C)> = 1)
  leading_boundary           src/algorithms/statmech/leading_boundary.jl,
                             src/algorithms/groundstate/vumps.jl:62-173
                             (dominant_eigsolve with which = :LM)
  VOMPS                      src/algorithms/statmech/vomps.jl:47-149
                             (AC <- dAC(AC), C <- dC(C), regauge)
  expectation value          src/algorithms/expval.jl:183-195
                             (product over the cell of <AC|dAC|AC>)
  Galerkin error             src/algorithms/toolbox.jl:63-79

Every transfer application and every derivative apply is the HIP library's
(``mk_transfer_{left,right}_mpo``, ``mk_prepare_ac_generic`` / ``mk_prepare_c``
and their applies, FP64 or mixed precision).  The dominant eigenvectors come
from the eager restarted Arnoldi of ``vumps.arnoldi_dominant`` over those
applies; in mixed precision a second Arnoldi pass in FP64, started from the
mixed result, decides convergence on the FP64 residual.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from . import tensors as T
from .derivatives import MPO_AC_Hamiltonian, MPO_C_Hamiltonian, prepare_operator
from .transfer import transfer_left, transfer_right
from .vumps import InfiniteMPS, _dot, _norm, arnoldi_dominant, regauge, uniform_rightorth


class InfiniteMPO:
    """Unit cell of dense MPO tensors O[g, b, h, k] (left, out, in, right) with no
    Jordan structure (``InfiniteMPO``, src/operators/mpo.jl)."""

    def __init__(self, Os):
        self.host = [np.asarray(O, dtype=np.float64) for O in Os]
        for i, O in enumerate(self.host):
            nxt = self.host[(i + 1) % len(self.host)]
            if O.ndim != 4 or O.shape[1] != O.shape[2] or O.shape[3] != nxt.shape[0]:
                raise ValueError("InfiniteMPO: site tensors must be (w_left, d, d, w_right) with matching bonds")
        self.O = [T.from_numpy(O) for O in self.host]
        self.N = len(Os)


def _left_cell(v, psi, O, ket):
    for s in range(psi.N):
        v = transfer_left(v, O.O[s], ket.AL[s], psi.AL[s])
    return v


def _right_cell(v, psi, O, ket):
    for s in range(psi.N - 1, -1, -1):
        v = transfer_right(v, O.O[s], ket.AR[s], psi.AR[s])
    return v


def _start(shape, seed):
    return T.from_numpy(np.random.default_rng(seed).random(shape) + 0.5)


class BoundaryEnvironments:
    """GLs[i] (D_bra x w x D_ket): left environment of site i; GRs[i] (D_ket x w x
    D_bra): right environment of site i.  The bra is the state passed to
    ``recalculate``; the ket is ``ket`` (default: the bra itself, as for
    leading_boundary; another state for approximate).  ``lambdas`` holds the cell
    eigenvalues of the last solve."""

    def __init__(self, psi, O, tol=1e-12, krylovdim=20, ket=None):
        self.tol, self.krylovdim, self.ket = tol, krylovdim, ket
        self.GLs = self.GRs = None
        self.lambdas = (None, None)
        self.numops = 0
        self.recalculate(psi, O)

    def recalculate(self, psi, O, tol=None):
        tol = self.tol if tol is None else tol
        ket = self.ket if self.ket is not None else psi
        N = psi.N
        shape_l = (psi.AL[0].shape[0], O.host[0].shape[0], ket.AL[0].shape[0])
        shape_r = (ket.AR[N - 1].shape[2], O.host[N - 1].shape[3], psi.AR[N - 1].shape[2])
        gl0 = self.GLs[0] if self.GLs is not None and tuple(self.GLs[0].shape) == shape_l else _start(shape_l, 1)
        gr0 = (self.GRs[N - 1] if self.GRs is not None and tuple(self.GRs[N - 1].shape) == shape_r
               else _start(shape_r, 2))
        # compute_leftenvs! / compute_rightenvs! (infinite_envs.jl:125-156)
        il, ir = {}, {}
        lam_l, gl = arnoldi_dominant(lambda v: _left_cell(v, psi, O, ket), gl0, tol=tol, krylovdim=self.krylovdim,
                                     relative=True, info=il)
        lam_r, gr = arnoldi_dominant(lambda v: _right_cell(v, psi, O, ket), gr0, tol=tol, krylovdim=self.krylovdim,
                                     relative=True, info=ir)
        self.numops += il["numops"] + ir["numops"]
        GL = [None] * N
        GR = [None] * N
        GL[0] = T.as_colmajor(gl)
        for s in range(1, N):
            GL[s] = transfer_left(GL[s - 1], O.O[s - 1], ket.AL[s - 1], psi.AL[s - 1])
        GR[N - 1] = T.as_colmajor(gr)
        for s in range(N - 2, -1, -1):
            GR[s] = transfer_right(GR[s + 1], O.O[s + 1], ket.AR[s + 1], psi.AR[s + 1])
        # normalize! (infinite_envs.jl:163-175): |GR| = 1, <C_bra|dC(C_ket)> = 1 through GL[i+1]
        for s in range(N):
            GR[s] = GR[s] / _norm(GR[s])
            nxt = (s + 1) % N
            Hc = prepare_operator(MPO_C_Hamiltonian(GL[nxt], GR[s]))
            lam = float(_dot(psi.C[s], Hc(ket.C[s])))
            Hc.free()
            GL[nxt] = GL[nxt] / lam
        self.GLs, self.GRs = GL, GR
        self.lambdas = (lam_l, lam_r)
        return self


def expectation_value_mpo(psi, O, envs):
    """Partition function per unit cell: prod over sites of <AC|GL O GR|AC>
    (contract_mpo_expval, expval.jl:186-195)."""
    val = 1.0
    for s in range(psi.N):
        op = prepare_operator(MPO_AC_Hamiltonian(envs.GLs[s], O.O[s], envs.GRs[s]))
        val *= float(_dot(psi.AC[s], op(psi.AC[s])))
        op.free()
    return val


def calc_galerkin_mpo(psi, O, envs):
    """max over sites of |(1 - AL AL^T) normalize(dAC(AC_ket))| (toolbox.jl:63-79),
    AC_ket the ket's center tensor (psi's own for leading_boundary)."""
    ket = envs.ket if envs.ket is not None else psi
    worst = 0.0
    for s in range(psi.N):
        op = prepare_operator(MPO_AC_Hamiltonian(envs.GLs[s], O.O[s], envs.GRs[s]))
        y = op(ket.AC[s])
        y = y / _norm(y)
        Xm, Ym = T.fuse(psi.AL[s], 2), T.fuse(y, 2)
        worst = max(worst, float(_norm(Ym - Xm @ (Xm.t() @ Ym))))
        op.free()
    return worst


def _dominant(op, x0, tol, precision, krylovdim, stats):
    """Dominant eigenvector of a prepared derivative (fixedpoint(..., :LM));
    mixed precision first, then FP64 until the FP64 residual meets ``tol``."""
    info = {}
    if precision != "fp64":
        _, x0 = arnoldi_dominant(lambda v: op(v, precision=precision), x0, tol=max(tol, 1e-7),
                                 krylovdim=krylovdim, relative=True, info=info)
        stats["mixed_ops"] += info["numops"]
    lam, x = arnoldi_dominant(lambda v: op(v), x0, tol=tol, krylovdim=krylovdim, relative=True, info=info)
    stats["fp64_ops"] += info["numops"]
    return lam, x


@dataclass
class BoundaryVUMPS:
    """VUMPS as ``leading_boundary`` uses it (vumps.jl:23-55 with which = :LM);
    ``precision`` applies to the AC / C eigensolver applies."""
    tol: float = 1e-10
    maxiter: int = 100
    verbosity: int = 1
    precision: str = "fp64"
    tol_factor: float = 0.1
    tol_min: float = 1e-13
    tol_max: float = 1e-4
    krylovdim: int = 20


@dataclass
class VOMPS:
    """Power method (vomps.jl:1-45): AC <- dAC(AC), C <- dC(C) each iteration."""
    tol: float = 1e-10
    maxiter: int = 200
    verbosity: int = 1
    precision: str = "fp64"
    tol_factor: float = 0.1
    tol_min: float = 1e-13
    tol_max: float = 1e-4


def leading_boundary(psi: InfiniteMPS, O: InfiniteMPO, alg):
    """``leading_boundary(psi, O, alg)`` for ``BoundaryVUMPS`` or ``VOMPS``
    (leading_boundary.jl; vumps.jl:62-125; vomps.jl:52-149).  Returns
    (psi, envs, eps, history) with the per-cell value of every iteration."""
    envs = BoundaryEnvironments(psi, O, tol=1e-10)
    eps = calc_galerkin_mpo(psi, O, envs)
    history = []
    vomps = isinstance(alg, VOMPS)
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    stats = dict(mixed_ops=0, fp64_ops=0)
    factor = alg.tol_factor
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tol = min(max(factor * eps, alg.tol_min), alg.tol_max)
        eps_before = eps
        new_AL = []
        for s in range(psi.N):
            nxt = (s + 1) % psi.N
            Hac = prepare_operator(MPO_AC_Hamiltonian(envs.GLs[s], O.O[s], envs.GRs[s]), precisions=precs)
            Hc = prepare_operator(MPO_C_Hamiltonian(envs.GLs[nxt], envs.GRs[s]), precisions=precs)
            if vomps:
                # AC_projection / C_projection (vomps.jl:127-135)
                AC = Hac(psi.AC[s], precision=alg.precision)
                C = Hc(psi.C[s], precision=alg.precision)
                stats["mixed_ops" if alg.precision != "fp64" else "fp64_ops"] += 2
            else:
                # _localupdate_vumps_step! with which = :LM (vumps.jl:156-173)
                _, AC = _dominant(Hac, psi.AC[s], tol, alg.precision, alg.krylovdim, stats)
                _, C = _dominant(Hc, psi.C[s], tol, alg.precision, alg.krylovdim, stats)
            Hac.free()
            Hc.free()
            new_AL.append(regauge(AC, C))
        # gauge_step! (vumps.jl:175-183, vomps.jl:145-148)
        AR, C = uniform_rightorth(new_AL, psi.C[-1], tol=max(tol, 1e-14))
        psi = InfiniteMPS(new_AL, AR, C)
        envs.recalculate(psi, O, tol=max(tol * 0.1, 1e-14))
        eps = calc_galerkin_mpo(psi, O, envs)
        if eps > 0.5 * eps_before:
            # stagnation: the eigenvector error tracks the solver tolerance over the
            # (small) relative gap of dAC, so the tolerance tied to eps must shrink
            factor *= 0.1
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        val = expectation_value_mpo(psi, O, envs)
        history.append(dict(iter=it, value=val, eps=eps, time=dt, **stats))
        if alg.verbosity >= 2:
            name = "VOMPS" if vomps else "VUMPS"
            print(f"{name} {it:3d}: lambda = {val:.15f}  err = {eps:.3e}  time = {dt:.3f}s", flush=True)
        if eps <= alg.tol:
            break
    return psi, envs, eps, history


def approximate_infinite(psi: InfiniteMPS, O: InfiniteMPO, phi: InfiniteMPS, alg: VOMPS):
    """``approximate(psi, (O, phi), VOMPS(...))`` (approximate/vomps.jl:15-110): psi
    approaches O|phi> (up to normalisation) within its bond dimension.  Each
    iteration sets AC = dAC(phi.AC) and C = dC(phi.C) with the mixed
    environments (bra psi, ket phi), regauges and recomputes the environments.
    Returns (psi, envs, eps, history)."""
    envs = BoundaryEnvironments(psi, O, tol=1e-10, ket=phi)
    eps = calc_galerkin_mpo(psi, O, envs)
    history = []
    precs = ("fp64",) if alg.precision == "fp64" else ("fp64", alg.precision)
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tol = min(max(alg.tol_factor * eps, alg.tol_min), alg.tol_max)
        new_AL = []
        for s in range(psi.N):
            nxt = (s + 1) % psi.N
            Hac = prepare_operator(MPO_AC_Hamiltonian(envs.GLs[s], O.O[s], envs.GRs[s]), precisions=precs)
            Hc = prepare_operator(MPO_C_Hamiltonian(envs.GLs[nxt], envs.GRs[s]), precisions=precs)
            AC = Hac(phi.AC[s], precision=alg.precision)
            C = Hc(phi.C[s], precision=alg.precision)
            Hac.free()
            Hc.free()
            new_AL.append(regauge(AC, C))
        AR, C = uniform_rightorth(new_AL, psi.C[-1], tol=max(tol, 1e-14))
        psi = InfiniteMPS(new_AL, AR, C)
        envs.recalculate(psi, O, tol=max(tol * 0.1, 1e-14))
        eps = calc_galerkin_mpo(psi, O, envs)
        torch.cuda.synchronize()
        history.append(dict(iter=it, eps=eps, time=time.perf_counter() - t0))
        if alg.verbosity >= 2:
            print(f"approximate VOMPS {it:3d}: err = {eps:.3e}", flush=True)
        if eps <= alg.tol:
            break
    return psi, envs, eps, history

