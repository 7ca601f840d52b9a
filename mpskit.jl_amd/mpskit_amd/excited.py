"""Excited states of a finite chain by penalised DMRG2 (FiniteExcited) on the HIP path.

Mirrors (paths relative to the MPSKit.jl tree):
  ProjectionDerivativeOperator  src/algorithms/derivatives/projection_derivatives.jl:1-80
                                (two-site projector derivative of |psi_i><psi_i|)
  projection environments       src/environments/finite_envs.jl:83-91 (overlaps of the
                                optimised state with psi_i, MPS transfers)
  FiniteExcited / excitations   src/algorithms/excitation/dmrgexcitation.jl:18-70
                                (DMRG on H + w sum_i |psi_i><psi_i|, perturbed start)

The two-site derivative of the penalty is rank one per projected state:
P_i(x) = v_i <v_i, x> with v_i = GLo_i . AC2(psi_i) . GRo_i, where GLo_i / GRo_i
are the left / right overlaps of the optimised state with psi_i.  The overlaps
are the HIP library's MPS transfers (mk_transfer_{left,right}_mps); the
Hamiltonian part is the prepared Jordan dAC2.  The local eigensolver is a
restarted Lanczos over that sum (full reorthogonalisation, one readback per
Krylov step).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import tensors as T
from .krylov import cgs2
from .algorithms import expectation_value, gauge2
from .derivatives import JordanMPO_AC2_Hamiltonian, prepare_operator
from .environments import FiniteEnvironments
from .states import FiniteMPS
from .transfer import transfer_left, transfer_right
from .vumps import _flat, _unflat


def lanczos_min(apply, x0, tol=1e-12, krylovdim=30, maxiter=50):
    """Lowest eigenpair of a symmetric map (KrylovKit.eigsolve(..., :SR) with
    Lanczos): restarted from the Ritz vector, converged when the residual
    |beta_k s_k| <= tol.  Returns (lambda, x, info)."""
    shape = tuple(x0.shape)
    x = _flat(x0)
    n = x.numel()
    kmax = max(1, min(krylovdim, n))
    Vb = torch.empty((kmax + 1, n), dtype=torch.float64, device=x.device)
    Vb[0] = x / torch.linalg.vector_norm(x)
    ops, res, lam = 0, math.inf, 0.0
    for _ in range(maxiter):
        H = np.zeros((kmax + 1, kmax))
        for j in range(kmax):
            w = _flat(apply(_unflat(Vb[j], shape)))
            ops += 1
            hcol = cgs2(Vb, j + 1, w)
            H[: j + 1, j] = hcol[:-1]
            H[j + 1, j] = hcol[-1]
            Tk = H[: j + 1, : j + 1]
            ev, U = np.linalg.eigh(0.5 * (Tk + Tk.T))
            lam, s = float(ev[0]), U[:, 0]
            res = abs(H[j + 1, j] * s[-1])
            if res <= tol or H[j + 1, j] <= 1e-14 * max(abs(ev).max(), 1e-300) or j + 1 == n:
                break
            Vb[j + 1] = w / H[j + 1, j]
        st = torch.from_numpy(np.ascontiguousarray(s)).to(Vb.device)
        x = Vb[: st.numel()].t() @ st
        x = x / torch.linalg.vector_norm(x)
        if res <= tol or H[j + 1, j] <= 1e-14 * max(abs(ev).max(), 1e-300) or j + 1 == n:
            break
        Vb[0] = x
    return lam, _unflat(x, shape), dict(numops=ops, normres=float(res), converged=res <= tol)


class OverlapEnvironments:
    """Left / right overlaps of the optimised state ``cur`` (first index) with a
    fixed state ``ket`` (second index), recomputed lazily when a tensor of either
    state is a different object than the one the cached overlap was built from,
    as FiniteEnvironments does."""

    def __init__(self, ket: FiniteMPS):
        self.ket = ket
        L = ket.L
        one = T.zeros((1, 1))
        one.fill_(1.0)
        self.GL = [one] + [None] * L
        self.GR = [None] * L + [one.clone()]
        self.ldep = [None] * L
        self.rdep = [None] * L

    @staticmethod
    def _stale(dep, a, b):
        return dep is None or dep[0] is not a or dep[1] is not b

    def left(self, pos, cur):
        k = self.ket
        first = next((j for j in range(pos) if self._stale(self.ldep[j], cur.AL[j], k.AL[j])), None)
        if first is not None:
            for j in range(first, pos):
                self.GL[j + 1] = transfer_left(self.GL[j], k.AL[j], cur.AL[j])
                self.ldep[j] = (cur.AL[j], k.AL[j])
        return self.GL[pos]

    def right(self, pos, cur):
        k, L = self.ket, self.ket.L
        first = next((j for j in range(L - 1, pos, -1) if self._stale(self.rdep[j], cur.AR[j], k.AR[j])), None)
        if first is not None:
            for j in range(first, pos, -1):
                self.GR[j] = transfer_right(self.GR[j + 1], cur.AR[j], k.AR[j])
                self.rdep[j] = (cur.AR[j], k.AR[j])
        return self.GR[pos + 1]

    def projected_ac2(self, pos, cur):
        """v = GLo . AC2(ket at pos) . GRo in the optimised state's local space."""
        k = self.ket
        if k.center not in (pos, pos + 1):
            k.move_center(pos)
        ac2 = k.AC2(pos)
        GL = self.left(pos, cur)
        GR = self.right(pos + 1, cur)
        return T.as_colmajor(torch.einsum("fk,khmj,rm->fhrj", GL, ac2, GR))


@dataclass
class FiniteExcited:
    """FiniteExcited (dmrgexcitation.jl:18-24): DMRG2 on H + weight sum_i |psi_i><psi_i|."""
    weight: float = 10.0
    trunc_dim: int = 64
    trunc_err: float = 1e-10
    maxiter: int = 20
    tol: float = 1e-10
    krylovdim: int = 30
    verbosity: int = 1


def _perturbed_state(psi: FiniteMPS, atol=1e-2, seed=11):
    """_perturbed_state (dmrgexcitation.jl:27-35): every tensor plus noise of norm atol."""
    rng = np.random.default_rng(seed)
    psi.move_center(psi.L - 1)
    tensors = []
    for j in range(psi.L):
        A = psi.AC if j == psi.L - 1 else psi.AL[j]
        noise = T.from_numpy(rng.standard_normal(tuple(A.shape)))
        tensors.append(T.as_colmajor(A + noise * (atol / float(torch.linalg.vector_norm(noise)))))
    return FiniteMPS(tensors)


def find_excited(psi: FiniteMPS, H, states, alg: FiniteExcited):
    """Lowest state of H + w sum_i |psi_i><psi_i| by two-site sweeps (the
    find_groundstate call of _excitations_finite).  Returns (psi, energy, history)."""
    L = psi.L
    envs = FiniteEnvironments(psi, H)
    ovl = [OverlapEnvironments(s) for s in states]
    history = []
    e_prev = math.inf
    for it in range(1, alg.maxiter + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ops = 0
        lam = 0.0
        for pos, direction in [(p, "right") for p in range(L - 1)] + [(p, "left") for p in range(L - 2, -1, -1)]:
            psi.move_center(pos if direction == "right" else pos + 1)
            Heff = prepare_operator(JordanMPO_AC2_Hamiltonian(envs.leftenv(pos, psi), H[pos], H[pos + 1],
                                                              envs.rightenv(pos + 1, psi)))
            vs = [o.projected_ac2(pos, psi) for o in ovl]

            def apply(x, Heff=Heff, vs=vs):
                y = Heff(T.as_colmajor(x))
                for v in vs:
                    y = y + (alg.weight * torch.dot(_flat(v), _flat(x))) * v
                return y

            lam, ac2, info = lanczos_min(apply, psi.AC2(pos), tol=alg.tol * 0.1, krylovdim=alg.krylovdim)
            ops += info["numops"]
            Heff.free()
            gauge2(psi, pos, direction, T.as_colmajor(ac2), alg.trunc_dim, alg.trunc_err)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        history.append(dict(iter=it, value=lam, numops=ops, time=dt))
        if alg.verbosity >= 2:
            print(f"FiniteExcited sweep {it}: value = {lam:.14f}  applies = {ops}  time = {dt:.3f}s", flush=True)
        if abs(lam - e_prev) <= alg.tol:
            break
        e_prev = lam
    E = expectation_value(psi, H, envs)
    return psi, E, history


def excitations(H, alg: FiniteExcited, states, init: FiniteMPS | None = None, num: int = 1):
    """``excitations(H, FiniteExcited(), states; init, num)`` (dmrgexcitation.jl:37-70):
    the ``num`` next states above ``states``, each penalising all found before it.
    The gauge center of the given states moves during the sweeps (the states
    themselves do not change).  Returns (energies, excited states)."""
    states = list(states) if isinstance(states, (list, tuple)) else [states]
    energies, found = [], []
    for _ in range(num):
        start = init if init is not None else _perturbed_state(_copy(states[0]))
        psi, E, _ = find_excited(_copy(start), H, states, alg)
        energies.append(E)
        found.append(psi)
        states = states + [psi]
    return energies, found


def _copy(psi: FiniteMPS):
    return FiniteMPS([T.as_colmajor(psi.tensor_at(j).clone()) for j in range(psi.L)])
