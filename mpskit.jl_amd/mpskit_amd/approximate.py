"""Variational compression of O|phi> into a finite MPS (approximate with DMRG2) on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  approximate!(psi, (O, phi), DMRG2)   src/algorithms/approximate/fvomps.jl:1-36: two-site
                                       sweeps, AC2_projection of O|phi> onto the local space of
                                       psi, truncated split, error = max relative change of AC2
  environments of (O, phi)             src/environments/finite_envs.jl (bra psi, ket phi)

The projection is the generic MPO dAC2 of the HIP library between mixed
environments: bra legs from psi (the approximant), ket legs from phi, built by
the MPO transfers with A != Abar.  phi's tensors are taken once, in whatever
gauge phi is in: the projection is a plain contraction of the network <psi|O|phi>.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import torch

from . import tensors as T
from .algorithms import truncated_split
from .derivatives import MPO_AC2_Hamiltonian, prepare_operator
from .states import FiniteMPS
from .transfer import transfer_left, transfer_right


@dataclass
class ApproximateDMRG2:
    """DMRG2 settings as approximate uses them: kept states ``trunc_dim``,
    discarded weight ``trunc_err``, stop when the largest relative change of a
    two-site tensor over a sweep is below ``tol``."""
    trunc_dim: int = 64
    trunc_err: float = 0.0
    tol: float = 1e-10
    maxiter: int = 20
    verbosity: int = 0


class _MixedEnvironments:
    """GL[j] (D_psi x w x D_phi) and GR[j], lazily rebuilt when psi's AL / AR at a
    site is a different object than the one the cached environment used."""

    def __init__(self, O, phi_tensors):
        L = len(phi_tensors)
        self.O, self.phi = O, phi_tensors
        one = T.zeros((1, 1, 1))
        one.fill_(1.0)
        self.GL = [one] + [None] * L
        self.GR = [None] * L + [one.clone()]
        self.ldep = [None] * L
        self.rdep = [None] * L

    def left(self, pos, psi):
        first = next((j for j in range(pos) if psi.AL[j] is not self.ldep[j]), None)
        if first is not None:
            for j in range(first, pos):
                self.GL[j + 1] = transfer_left(self.GL[j], self.O[j], self.phi[j], psi.AL[j])
                self.ldep[j] = psi.AL[j]
        return self.GL[pos]

    def right(self, pos, psi):
        L = len(self.phi)
        first = next((j for j in range(L - 1, pos, -1) if psi.AR[j] is not self.rdep[j]), None)
        if first is not None:
            for j in range(first, pos, -1):
                self.GR[j] = transfer_right(self.GR[j + 1], self.O[j], self.phi[j], psi.AR[j])
                self.rdep[j] = psi.AR[j]
        return self.GR[pos + 1]


def approximate(psi: FiniteMPS, O, phi: FiniteMPS, alg: ApproximateDMRG2):
    """``approximate!(psi, (O, phi), DMRG2(...))``: psi (modified in place)
    approaches O|phi> within its bond dimensions.  ``O`` is a finite MPO (for
    example a FiniteMPOHamiltonian; sequence of W[g, b, h, k] GPU tensors with
    boundary bonds of dimension 1).  Returns (psi, eps, history)."""
    L = psi.L
    if phi.L != L:
        raise ValueError("approximate: psi and phi have different lengths")
    phi_t = [T.as_colmajor(phi.tensor_at(j)) for j in range(L)]
    envs = _MixedEnvironments(O, phi_t)
    history = []
    eps = 2 * alg.tol
    for it in range(1, alg.maxiter + 1):
        t0 = time.perf_counter()
        eps = 0.0
        for pos, direction in [(p, "right") for p in range(L - 1)] + [(p, "left") for p in range(L - 3, -1, -1)]:
            psi.move_center(pos if direction == "right" else pos + 1)
            old = psi.AC2(pos)
            op = prepare_operator(MPO_AC2_Hamiltonian(envs.left(pos, psi), O[pos], O[pos + 1],
                                                      envs.right(pos + 1, psi)))
            x = T.as_colmajor(torch.einsum("fhm,mjr->fhrj", phi_t[pos], phi_t[pos + 1]))
            y = T.as_colmajor(op(x))
            op.free()
            Dl, d1, Dr, d2 = y.shape
            A, B, _, _ = truncated_split(T.fuse(y, 2), direction, alg.trunc_dim, alg.trunc_err)
            k = A.shape[1]
            new = (A @ B)
            if tuple(old.shape) == tuple(y.shape):
                eps = max(eps, float(torch.linalg.matrix_norm(new - T.fuse(old, 2)) /
                                     torch.linalg.vector_norm(old)))
            else:
                eps = max(eps, 1.0)
            if direction == "right":
                psi.AL[pos] = T.as_colmajor(T.unfuse(A, (Dl, d1, k)))
                psi.AC = T.as_colmajor(T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1))
                psi.AR[pos] = psi.AR[pos + 1] = None
                psi.center = pos + 1
            else:
                psi.AR[pos + 1] = T.as_colmajor(T.unfuse(B, (k, Dr, d2)).permute(0, 2, 1))
                psi.AC = T.as_colmajor(T.unfuse(A, (Dl, d1, k)))
                psi.AL[pos] = psi.AL[pos + 1] = None
                psi.center = pos
        history.append(dict(iter=it, eps=eps, time=time.perf_counter() - t0))
        if alg.verbosity >= 2:
            print(f"approximate DMRG2 {it}: eps = {eps:.3e}", flush=True)
        if eps < alg.tol:
            break
    return psi, eps, history
