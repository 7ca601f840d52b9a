"""Quasiparticle environments (src/environments/qp_envs.jl): a finite chain
(:139-162) and an infinite MPS with a Jordan MPO Hamiltonian (:49-137, with
left/right_excitation_transfer_system, exci_transfer_system.jl:1-100) at the
momenta whose phase is real (k = 0, pi); other momenta need complex tensors.

A finite quasiparticle state is sum_j AL[0..j-1] B[j] AR[j+1..L-1] over the
ground state's left- and right-gauged tensors, with excited tensors
B[j] (Dl, d, De, Dr) carrying the excitation leg e.  Its environments are

    lB[0] = 0,    lB[p+1] = lB[p] * T(AR[p], W[p], AL[p]) + GL[p] * T(B[p], W[p], AL[p])
    rB[L-1] = 0,  rB[p-1] = T(AL[p], W[p], AR[p]) * rB[p] + T(B[p], W[p], AR[p]) * GR[p]

on the HIP transfers with an excitation leg (``mk_transfer_left_leg`` /
``mk_transfer_right_leg``, transfer.jl:141-148) and with an excited upper
tensor (``mk_transfer_left_excited`` / ``mk_transfer_right_excited``,
transfer.jl:133-138).  lB[p] (D_bra, w, De, D_ket) lives on the left bond of
site p, rB[p] (D_ket, w, De, D_bra) on its right bond, as ``allocate_GBL`` /
``allocate_GBR`` (abstract_envs.jl:60-86) lay them out.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import tensors as T
from .transfer import regularize_, transfer_left, transfer_right


def ground_state_environments(AL, AR, H):
    """GL[j] / GR[j], the left / right MPO environments of site j for a finite
    ground state given in both gauges (finite_envs.jl:110-150)."""
    L = len(H)
    dev = AL[0].device
    one = T.zeros((1, 1, 1), dev)
    one.fill_(1.0)
    GL = [one]
    for j in range(L - 1):
        GL.append(transfer_left(GL[j], H[j], AL[j], AL[j]))
    GR = [None] * L
    GR[L - 1] = one.clone()
    for j in range(L - 1, 0, -1):
        GR[j - 1] = transfer_right(GR[j], H[j], AR[j], AR[j])
    return GL, GR


def finite_qp_environments(Bs, AL, AR, H, GL=None, GR=None):
    """``environments(exci::FiniteQP, H)`` (qp_envs.jl:139-162) for a
    quasiparticle on a finite ground state (left_gs = right_gs): ``Bs[j]`` the
    excited tensors, ``AL`` / ``AR`` the ground state's tensors in both gauges,
    ``H`` the MPO tensors W[j] (wl, d, d, wr), ``GL`` / ``GR`` its environments
    (computed from AL / AR when not given).  Returns the lists (lB, rB)."""
    L = len(H)
    if len(Bs) != L or len(AL) != L or len(AR) != L:
        raise ValueError("finite_qp_environments: Bs, AL, AR and H must have one entry per site")
    if GL is None or GR is None:
        GL, GR = ground_state_environments(AL, AR, H)
    De = Bs[0].shape[2]
    dev = AL[0].device
    lB = [T.zeros((AL[0].shape[0], H[0].shape[0], De, AR[0].shape[0]), dev)]
    for p in range(L - 1):
        y = transfer_left(lB[p], H[p], AR[p], AL[p])
        y += transfer_left(GL[p], H[p], Bs[p], AL[p])
        lB.append(y)
    rB = [None] * L
    rB[L - 1] = T.zeros((AL[-1].shape[2], H[-1].shape[3], De, AR[-1].shape[2]), dev)
    for p in range(L - 1, 0, -1):
        y = transfer_right(rB[p], H[p], AL[p], AR[p])
        y += transfer_right(GR[p], H[p], Bs[p], AR[p])
        rB[p - 1] = y
    return lB, rB


def level_kinds(H):
    """Kind of each Jordan level of an infinite MPO cell (mpohamiltonian.jl:677-693):
    "identity" for the first and last level and where every site holds the identity
    on the diagonal, "empty" where some site has no diagonal block, else "diagonal"."""
    w = H.w
    kinds = []
    for i in range(w):
        blocks = [W[i, :, :, i] for W in H.host]
        if i == 0 or i == w - 1 or all(np.array_equal(b, np.eye(b.shape[0])) for b in blocks):
            kinds.append("identity")
        elif any(not np.any(b) for b in blocks):
            kinds.append("empty")
        else:
            kinds.append("diagonal")
    return kinds


def _reg(v, Cb):
    """regularize!(v, C', C) of a level slice v[a, e, c] (transfermatrix.jl:79-84):
    v[a,e,c] -= (sum_ij C[j,i] v[j,e,i]) C[a,c], on the library (``mk_regularize``)."""
    out = T.as_colmajor(v.clone())
    return regularize_(out, T.as_colmajor(Cb.t()), Cb)


def _set_level(y, i, v):
    y[:, i] = v
    return y


def infinite_qp_environments(Bs, psi, H, envs, momentum, tol=1e-12):
    """``environments(exci::InfiniteQP, H::InfiniteMPOHamiltonian)`` (qp_envs.jl:49-137)
    for a trivial excitation (left_gs = right_gs = ``psi``, an ``InfiniteMPS``) with
    excited tensors ``Bs[s]`` (Dl, d, De, Dr), ``H`` an ``InfiniteHamiltonian``,
    ``envs`` its ``InfiniteEnvironments`` of ``psi`` and ``momentum`` 0 or pi.
    Returns (lB, rB): lB[s] (D, w, De, D) on the left bond of site s, rB[s] on the
    right one.  The per-level linear systems are solved by GMRES to ``tol`` relative
    to their right-hand sides.  Regularization vectors: C[s] is the bond right of
    site s, and l_RL / r_RL / l_LR / r_LR (infinitemps.jl:356-401) make every
    regularization regularize!(v, C', C) with the C of v's bond."""
    from .vumps import gmres
    if abs(math.sin(momentum)) > 1e-12:
        raise NotImplementedError("quasiparticle environments at momentum %r need complex tensors; "
                                  "the float64 path covers momentum 0 and pi" % (momentum,))
    p = float(math.cos(momentum))
    N, w = psi.N, H.w
    De = Bs[0].shape[2]
    if len(Bs) != N or H.N != N:
        raise ValueError("infinite_qp_environments: one excited tensor and one MPO tensor per site")
    AL, AR, C, W = psi.AL, psi.AR, psi.C, H.W
    GL, GR = envs.GLs, envs.GRs
    kinds = level_kinds(H)
    ids = [i for i in range(1, w - 1) if kinds[i] == "identity"]
    pN = p ** N
    dev = AL[0].device

    def solve(op, b):
        bn = max(float(torch.linalg.vector_norm(b)), 1e-300)
        x, _ = gmres(lambda x: x - pN * op(x), b, T.as_colmajor(b.clone()), tol * bn)
        return x

    # qp_envs.jl:58-75
    lB = [T.zeros((AL[s].shape[0], w, De, AR[s].shape[0]), dev) for s in range(N)]
    for s in range(N):
        y = transfer_left(lB[s], W[s], AR[s], AL[s])
        y += transfer_left(GL[s], W[s], Bs[s], AL[s])
        y /= p
        for i in ids:
            _set_level(y, i, _reg(y[:, i], C[s]))
        lB[(s + 1) % N] = y
    # qp_envs.jl:77-90
    rB = [T.zeros((AL[s].shape[2], w, De, AR[s].shape[2]), dev) for s in range(N)]
    for s in range(N - 1, -1, -1):
        y = transfer_right(rB[s], W[s], AL[s], AR[s])
        y += transfer_right(GR[s], W[s], Bs[s], AR[s])
        y *= p
        for i in ids:
            _set_level(y, i, _reg(y[:, i], C[s - 1]))
        rB[(s - 1) % N] = y
    Cb = C[N - 1]

    # left_excitation_transfer_system (exci_transfer_system.jl:1-48)
    GBL = lB[0]
    found = T.zeros(GBL.shape, dev)
    for i in range(w):
        x = T.as_colmajor(found[:, : i + 1])
        for s in range(N):
            x = transfer_left(x, T.as_colmajor(W[s][: i + 1, :, :, : i + 1]), AR[s], AL[s])
        start = x[:, i] * pN
        if kinds[i] == "identity":
            start = _reg(start, Cb)
        b = T.as_colmajor(start + GBL[:, i])
        if kinds[i] == "identity":
            def op(v):
                for s in range(N):
                    v = transfer_left(T.as_colmajor(v), AR[s], AL[s])
                return _reg(v, Cb)
            b = solve(op, b)
        elif kinds[i] == "diagonal":
            Wd = [T.as_colmajor(W[s][i:i + 1, :, :, i:i + 1]) for s in range(N)]

            def op(v, Wd=Wd):
                v = T.as_colmajor(v[:, None])
                for s in range(N):
                    v = transfer_left(v, Wd[s], AR[s], AL[s])
                return v[:, 0]
            b = solve(op, b)
        found[:, i] = b
    lB[0] = found

    # right_excitation_transfer_system (exci_transfer_system.jl:50-100)
    GBR = rB[N - 1]
    found = T.zeros(GBR.shape, dev)
    for i in range(w - 1, -1, -1):
        x = T.as_colmajor(found[:, i:])
        for s in range(N - 1, -1, -1):
            x = transfer_right(x, T.as_colmajor(W[s][i:, :, :, i:]), AL[s], AR[s])
        start = x[:, 0] * pN
        if kinds[i] == "identity":
            start = _reg(start, Cb)
        b = T.as_colmajor(start + GBR[:, i])
        if kinds[i] == "identity":
            def op(v):
                for s in range(N - 1, -1, -1):
                    v = transfer_right(T.as_colmajor(v), AL[s], AR[s])
                return _reg(v, Cb)
            b = solve(op, b)
        elif kinds[i] == "diagonal":
            Wd = [T.as_colmajor(W[s][i:i + 1, :, :, i:i + 1]) for s in range(N)]

            def op(v, Wd=Wd):
                v = T.as_colmajor(v[:, None])
                for s in range(N - 1, -1, -1):
                    v = transfer_right(v, Wd[s], AL[s], AR[s])
                return v[:, 0]
            b = solve(op, b)
        found[:, i] = b
    rB[N - 1] = found

    # qp_envs.jl:110-136
    cur = lB[0]
    for s in range(N - 1):
        cur = transfer_left(cur, W[s], AR[s], AL[s])
        cur /= p
        for i in ids:
            _set_level(cur, i, _reg(cur[:, i], C[s]))
        lB[s + 1] = lB[s + 1] + cur
    cur = rB[N - 1]
    for s in range(N - 1, 0, -1):
        cur = transfer_right(cur, W[s], AL[s], AR[s])
        cur *= p
        for i in ids:
            _set_level(cur, i, _reg(cur[:, i], C[s - 1]))
        rB[s - 1] = rB[s - 1] + cur
    return lB, rB
