"""Quasiparticle environments of a finite chain (src/environments/qp_envs.jl:139-162).

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

from . import tensors as T
from .transfer import transfer_left, transfer_right


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
