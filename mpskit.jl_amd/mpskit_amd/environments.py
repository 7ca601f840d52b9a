"""Finite MPO environments with lazy recomputation (src/environments/finite_envs.jl).

``GLs[j]`` is the left environment of site j (sites < j contracted) and
``GRs[j+1]`` the right environment of site j (sites > j).  As in the
reference, each cached environment remembers the tensor it was built from and
is recomputed only when the state's tensor at that site is a different object
(``leftenv``/``rightenv``, finite_envs.jl:110-150).  Updates are the MPO
transfer applications of the HIP library (transfer.jl:125-130).
"""
from __future__ import annotations

from . import tensors as T
from .transfer import transfer_left, transfer_right


class FiniteEnvironments:
    def __init__(self, psi, H):
        L = len(H)
        self.H = H
        dev = T.device()
        one = T.zeros((1, 1, 1), dev)
        one.fill_(1.0)
        # leftstart / rightstart: isomorphisms of the trivial boundary spaces (finite_envs.jl:36-47)
        self.GLs = [one] + [None] * L
        self.GRs = [None] * L + [one.clone()]
        self.ldep = [None] * L
        self.rdep = [None] * L
        self.n_updates = 0

    def leftenv(self, pos, psi):
        """Left environment of site ``pos``, updated from AL[0..pos-1] as needed."""
        first = next((j for j in range(pos) if psi.AL[j] is not self.ldep[j]), None)
        if first is not None:
            for j in range(first, pos):
                A = psi.AL[j]
                self.GLs[j + 1] = transfer_left(self.GLs[j], self.H[j], A, A)
                self.ldep[j] = A
                self.n_updates += 1
        return self.GLs[pos]

    def rightenv(self, pos, psi):
        """Right environment of site ``pos``, updated from AR[pos+1..L-1] as needed."""
        L = len(self.H)
        first = next((j for j in range(L - 1, pos, -1) if psi.AR[j] is not self.rdep[j]), None)
        if first is not None:
            for j in range(first, pos, -1):
                A = psi.AR[j]
                self.GRs[j] = transfer_right(self.GRs[j + 1], self.H[j], A, A)
                self.rdep[j] = A
                self.n_updates += 1
        return self.GRs[pos + 1]


def environments(psi, H):
    return FiniteEnvironments(psi, H)
