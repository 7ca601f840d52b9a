"""Finite-MPS TDVP and TDVP2 in imaginary time on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  TDVP  (one site)   src/algorithms/timestep/tdvp.jl:147-210 (_timestep_finite!)
  TDVP2 (two site)   src/algorithms/timestep/tdvp.jl:265-311 (_timestep2_finite!)
  TDVP, InfiniteMPS  src/algorithms/timestep/tdvp.jl:72-133 (_timestep_infinite: every AC
                     and C evolved, regauge, right orthogonalisation, new environments)
  integrate          src/algorithms/timestep/integrators.jl:21-31 with
                     imaginary_evolution = true: y = exponentiate(H, -dt, y0)

Every local operator is a prepared derivative of the HIP library (Jordan dAC,
dAC2 and the generic dC over the finite environments), so each Krylov step of
the exponential is one HIP apply.  Imaginary-time evolution keeps every tensor
real; real-time evolution needs complex tensors, which this float64 path does
not carry.  As in the reference, the norm is kept unless ``normalize``: the
decaying weight exp(-tau E) of imaginary time stays in the state.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import tensors as T
from .krylov import cgs2
from .algorithms import truncated_split
from .derivatives import JordanMPO_AC2_Hamiltonian, JordanMPO_AC_Hamiltonian, MPO_C_Hamiltonian, prepare_operator
from .environments import FiniteEnvironments
from .states import FiniteMPS, lq_right, qr_left
from .vumps import _flat, _unflat


def exponentiate(op, tau, x0, tol=1e-12, krylovdim=30, maxsteps=256):
    """y = exp(tau H) x0 for a symmetric operator H given by ``op`` (KrylovKit's
    exponentiate with Lanczos, as integrate calls it).  The Krylov basis is one
    GPU matrix with full (two-pass classical Gram-Schmidt) reorthogonalisation
    and one readback per step; exp(tau T_k) e1 comes from the eigendecomposition
    of the small tridiagonal matrix.  A step is accepted when the a-posteriori
    error beta_k |e_k^T exp(tau T_k) e1| is below ``tol`` (relative to |y|);
    otherwise the time step is halved and the integration proceeds in pieces
    (a piece that converged before the Krylov limit doubles the next one).
    Returns (y, info) with info = dict(numops, substeps)."""
    shape = tuple(x0.shape)
    y = _flat(x0).clone()
    n = y.numel()
    kmax = max(1, min(krylovdim, n))
    Vb = torch.empty((kmax + 1, n), dtype=torch.float64, device=y.device)
    remaining, step = float(tau), float(tau)
    ops, substeps = 0, 0
    while remaining != 0.0:
        if substeps >= maxsteps:
            raise RuntimeError(f"exponentiate: no convergence within {maxsteps} substeps")
        step = math.copysign(min(abs(step), abs(remaining)), remaining)
        beta0 = float(torch.linalg.vector_norm(y))
        if beta0 == 0.0:
            break
        Vb[0] = y / beta0
        H = np.zeros((kmax + 1, kmax))
        c = None
        accepted = False
        for j in range(kmax):
            w = _flat(op(_unflat(Vb[j], shape)))
            ops += 1
            hcol = cgs2(Vb, j + 1, w)
            H[: j + 1, j] = hcol[:-1]
            H[j + 1, j] = hcol[-1]
            Tk = H[: j + 1, : j + 1]
            ev, U = np.linalg.eigh(0.5 * (Tk + Tk.T))
            c = U @ (np.exp(step * ev) * U[0, :])
            err = H[j + 1, j] * abs(c[-1]) / max(np.linalg.norm(c), 1e-300)
            if err <= tol or H[j + 1, j] <= 1e-14 * max(abs(ev).max(), 1e-300) or j + 1 == n:
                accepted = True
                break
            Vb[j + 1] = w / H[j + 1, j]
        if not accepted:
            step *= 0.5
            substeps += 1
            continue
        ct = torch.from_numpy(np.ascontiguousarray(c)).to(Vb.device)
        y = beta0 * (Vb[: ct.numel()].t() @ ct)
        remaining -= step
        substeps += 1
        if j + 1 < kmax:
            step *= 2.0
    return _unflat(y, shape), dict(numops=ops, substeps=substeps)


@dataclass
class TDVP:
    """One-site TDVP (tdvp.jl:18-70) with the Lanczos exponential integrator."""
    tol: float = 1e-12
    krylovdim: int = 30


@dataclass
class TDVP2:
    """Two-site TDVP (tdvp.jl:212-251): truncated split of the evolved two-site
    center (kept states ``trunc_dim``, discarded weight ``trunc_err``)."""
    trunc_dim: int = 64
    trunc_err: float = 0.0
    tol: float = 1e-12
    krylovdim: int = 30


def _evolve(op_desc, tau, x, alg, stats):
    op = prepare_operator(op_desc)
    y, info = exponentiate(op, tau, x, tol=alg.tol, krylovdim=alg.krylovdim)
    op.free()
    stats["numops"] += info["numops"]
    return y


def _ac_op(psi, H, envs, i):
    return JordanMPO_AC_Hamiltonian(envs.leftenv(i, psi), H[i], envs.rightenv(i, psi))


def _unit(t, normalize):
    return t / torch.linalg.vector_norm(t) if normalize else t


def timestep(psi: FiniteMPS, H, dt, alg, envs: FiniteEnvironments | None = None, normalize=False):
    """``timestep!(psi, H, t, dt, alg; imaginary_evolution = true)`` for a finite
    MPS: one symmetric sweep evolving psi by exp(-dt H).  Returns (psi, envs, stats)."""
    envs = envs or FiniteEnvironments(psi, H)
    stats = {"numops": 0, "discarded": 0.0}
    if isinstance(alg, TDVP2):
        _sweep2(psi, H, dt, alg, envs, normalize, stats)
    else:
        _sweep1(psi, H, dt, alg, envs, normalize, stats)
    return psi, envs, stats


def _sweep1(psi, H, dt, alg, envs, normalize, stats):
    L = psi.L
    psi.move_center(0)
    # left to right (tdvp.jl:152-173)
    for i in range(L - 1):
        AC = _evolve(_ac_op(psi, H, envs, i), -dt / 2, psi.AC, alg, stats)
        AL, C = qr_left(T.as_colmajor(AC))
        psi.AL[i] = AL
        psi.AR[i] = None
        C = _unit(C, normalize)
        C = _evolve(MPO_C_Hamiltonian(envs.leftenv(i + 1, psi), envs.rightenv(i, psi)), dt / 2,
                    T.as_colmajor(C), alg, stats)
        psi.AC = T.as_colmajor(torch.einsum("lm,mpr->lpr", C, psi.AR[i + 1]))
        psi.AR[i + 1] = None
        psi.center = i + 1
    psi.AC = T.as_colmajor(_evolve(_ac_op(psi, H, envs, L - 1), -dt / 2, psi.AC, alg, stats))
    # right to left (tdvp.jl:179-205)
    for i in range(L - 1, 0, -1):
        AC = _evolve(_ac_op(psi, H, envs, i), -dt / 2, psi.AC, alg, stats)
        C, AR = lq_right(T.as_colmajor(AC))
        psi.AR[i] = AR
        psi.AL[i] = None
        C = _unit(C, normalize)
        C = _evolve(MPO_C_Hamiltonian(envs.leftenv(i, psi), envs.rightenv(i - 1, psi)), dt / 2,
                    T.as_colmajor(C), alg, stats)
        psi.AC = T.as_colmajor(torch.einsum("lpm,mr->lpr", psi.AL[i - 1], C))
        psi.AL[i - 1] = None
        psi.center = i - 1
    psi.AC = T.as_colmajor(_evolve(_ac_op(psi, H, envs, 0), -dt / 2, psi.AC, alg, stats))


def _split2(psi, i, ac2, direction, alg, normalize, stats):
    """svd_trunc! of the evolved two-site center; the singular values go to the
    side that becomes the center (tdvp.jl:274-279, :293-298)."""
    Dl, d1, Dr, d2 = ac2.shape
    A, B, S, eps = truncated_split(T.fuse(ac2, 2), direction, alg.trunc_dim, alg.trunc_err)
    stats["discarded"] = max(stats["discarded"], eps)
    k = A.shape[1]
    if direction == "right":
        psi.AL[i] = T.as_colmajor(T.unfuse(A, (Dl, d1, k)))
        psi.AC = T.as_colmajor(_unit(T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1), normalize))
        psi.AR[i] = psi.AR[i + 1] = None
        psi.center = i + 1
    else:
        psi.AR[i + 1] = T.as_colmajor(T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1))
        psi.AC = T.as_colmajor(_unit(T.unfuse(A, (Dl, d1, k)), normalize))
        psi.AL[i] = psi.AL[i + 1] = None
        psi.center = i


def _sweep2(psi, H, dt, alg, envs, normalize, stats):
    L = psi.L
    psi.move_center(0)
    # left to right (tdvp.jl:270-287)
    for i in range(L - 1):
        op = JordanMPO_AC2_Hamiltonian(envs.leftenv(i, psi), H[i], H[i + 1], envs.rightenv(i + 1, psi))
        ac2 = _evolve(op, -dt / 2, psi.AC2(i), alg, stats)
        _split2(psi, i, T.as_colmajor(ac2), "right", alg, normalize, stats)
        if i != L - 2:
            psi.AC = T.as_colmajor(_evolve(_ac_op(psi, H, envs, i + 1), dt / 2, psi.AC, alg, stats))
    # right to left (tdvp.jl:290-306)
    for i in range(L - 1, 0, -1):
        op = JordanMPO_AC2_Hamiltonian(envs.leftenv(i - 1, psi), H[i - 1], H[i], envs.rightenv(i, psi))
        ac2 = _evolve(op, -dt / 2, psi.AC2(i - 1), alg, stats)
        _split2(psi, i - 1, T.as_colmajor(ac2), "left", alg, normalize, stats)
        if i != 1:
            psi.AC = T.as_colmajor(_evolve(_ac_op(psi, H, envs, i - 1), dt / 2, psi.AC, alg, stats))


def time_evolve(psi: FiniteMPS, H, dt, nsteps, alg, envs=None, normalize=False, callback=None):
    """``time_evolve`` over ``nsteps`` imaginary-time steps of ``dt``
    (src/algorithms/timestep/time_evolve.jl); returns (psi, envs, history)."""
    envs = envs or FiniteEnvironments(psi, H)
    history = []
    for k in range(nsteps):
        psi, envs, stats = timestep(psi, H, dt, alg, envs, normalize)
        history.append(stats)
        if callback is not None:
            callback(k, psi, envs)
    return psi, envs, history


def timestep_infinite(psi, H, dt, alg: TDVP, envs=None):
    """``timestep(psi::InfiniteMPS, H, t, dt, TDVP(); imaginary_evolution = true)``
    (tdvp.jl:91-133, leftorthflag = true): AC[i] <- exp(-dt H_AC) AC[i] and
    C[i] <- exp(-dt H_C) C[i] for every site with the current environments, AL[i]
    = regauge(AC[i], C[i]), then the right-orthonormal form and new environments.
    ``psi`` is a vumps.InfiniteMPS and ``H`` a vumps.InfiniteHamiltonian.
    Returns (psi, envs, stats)."""
    from .vumps import InfiniteEnvironments, InfiniteMPS, regauge, uniform_rightorth
    envs = envs or InfiniteEnvironments(psi, H, tol=1e-12)
    stats = {"numops": 0, "discarded": 0.0}
    new_AL = []
    for s in range(psi.N):
        AC = _evolve(JordanMPO_AC_Hamiltonian(envs.GLs[s], H.W[s], envs.GRs[s]), -dt, psi.AC[s], alg, stats)
        C = _evolve(MPO_C_Hamiltonian(envs.GLs[(s + 1) % psi.N], envs.GRs[s]), -dt, psi.C[s], alg, stats)
        new_AL.append(regauge(T.as_colmajor(AC), T.as_colmajor(C)))
    AR, C = uniform_rightorth(new_AL, psi.C[-1], tol=1e-14)
    psi = InfiniteMPS(new_AL, AR, C)
    envs.recalculate(psi, H, tol=1e-12)
    return psi, envs, stats

