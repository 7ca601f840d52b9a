"""CPU oracle for the quasiparticle, passenger-leg, density-matrix and MPO-pair
transfers of src/transfermatrix/transfer.jl:58-161.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, as the
checker; the product path never imports it.

One function per reference method, each the method's ``@plansor`` index list
restated as an ``einsum`` over float64 arrays with axes ``(codomain...,
domain...)`` (the conventions of ``oracle.oracle``).  An excited
(quasiparticle) tensor is ``A[l, p, e, r]`` (the reference's A[l p; e r]).
The braiding tensor ``tau`` is kept as an explicit array, so each restatement
reads like its ``@plansor`` line: for bosonic (Trivial / U(1)) spaces
``tau[a, b, c, d] = delta(a, d) delta(b, c)``.

Pinning (tests/test_oracle_qp.py): every function is compared, slice by slice
over its excitation / passenger legs, with ``oracle.transfer_left_mps`` /
``transfer_right_mps`` / ``transfer_left_mpo`` / ``transfer_right_mpo``, which
tests/test_oracle.py pins against Kronecker-product matrices and dense
Hamiltonians; the density-matrix MPO transfers are compared with the MPO
transfer of the fused physical leg under ``O (x) 1`` built by explicit loops;
left and right variants are tied by the ring identity
<TL(v), u> = <v, TR(u)>; a planted leg exchange is shown to fail; and the
finite quasiparticle environments are compared with the same contractions
built from dense state blocks and the dense MPO.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "tau", "tl_excited_bond", "tr_excited_bond", "tl_excited_leg", "tr_excited_leg",
    "tl_utility", "tr_utility", "tl_mpdo_utility", "tr_mpdo_utility", "tl_utility_excitation",
    "tr_utility_excitation", "tl_mpo_pair", "tr_mpo_pair", "tl_mpo_excited", "tr_mpo_excited",
    "tl_mpo_excitation_leg", "tr_mpo_excitation_leg", "tl_mpo_mpdo", "tr_mpo_mpdo", "finite_qp_environments",
]


def _es(spec, *ops):
    return np.einsum(spec, *ops, optimize=True)


def tau(n_outer, n_inner):
    """Bosonic braiding tensor tau[a, b, c, d] = delta(a, d) delta(b, c),
    a, d of size n_outer and b, c of size n_inner."""
    t = np.zeros((n_outer, n_inner, n_inner, n_outer))
    for a in range(n_outer):
        for b in range(n_inner):
            t[a, b, b, a] = 1.0
    return t


# -- excited upper tensor ---------------------------------------------------

def tl_excited_bond(v, A, Ab):
    """transfer.jl:58-60: t[-1; -2 -3] := v[1; 2] A[2 3; -2 -3] conj(Ab[1 3; -1])."""
    return _es("mn,nhec,mha->aec", v, A, np.conj(Ab))


def tr_excited_bond(v, A, Ab):
    """transfer.jl:61-63: t[-1; -2 -3] := A[-1 3; -2 1] v[1; 2] conj(Ab[-3 3; 2])."""
    return _es("ahei,ij,bhj->aeb", A, v, np.conj(Ab))


def tl_excited_leg(v, A, Ab):
    """transfer.jl:66-69: t[-1 -2; -3 -4] := v[1 3; 4] A[4 5; -3 -4] tau[3 2; 5 -2] conj(Ab[1 2; -1])."""
    t = tau(v.shape[1], A.shape[1])
    return _es("mgn,npec,ghpG,mha->aGec", v, A, t, np.conj(Ab))


def tr_excited_leg(v, A, Ab):
    """transfer.jl:70-73: t[-1 -2; -3 -4] := A[-1 4; -3 5] tau[-2 3; 4 2] conj(Ab[-4 3; 1]) v[5 2; 1]."""
    t = tau(v.shape[1], A.shape[1])
    return _es("apei,Gkpg,bkj,igj->aGeb", A, t, np.conj(Ab), v)


# -- utility (passenger) legs of the density matrix ---------------------------

def tl_utility(v, A, Ab):
    """transfer.jl:76-79: v[-1 -2; -3] := v[1 2; 4] A[4 5; -3] tau[2 3; 5 -2] conj(Ab[1 3; -1])."""
    t = tau(v.shape[1], A.shape[1])
    return _es("mgn,npc,ghpG,mha->aGc", v, A, t, np.conj(Ab))


def tr_utility(v, A, Ab):
    """transfer.jl:80-83: v[-1 -2; -3] := A[-1 2; 1] tau[-2 4; 2 3] conj(Ab[-3 4; 5]) v[1 3; 5]."""
    t = tau(v.shape[1], A.shape[1])
    return _es("api,Gkpg,bkj,igj->aGb", A, t, np.conj(Ab), v)


def tl_mpdo_utility(v, A, Ab):
    """transfer.jl:85-91: v[-1 -2; -3] := v[1 3; 8] A[8 7 6; -3] tau[3 2; 7 5] tau[5 4; 6 -2]
    conj(Ab[1 2 4; -1])."""
    g, d, dq = v.shape[1], A.shape[1], A.shape[2]
    return _es("mgn,npqc,gkpf,frqG,mkra->aGc", v, A, tau(g, d), tau(g, dq), np.conj(Ab))


def tr_mpdo_utility(v, A, Ab):
    """transfer.jl:92-99: v[-1 -2; -3] := A[-1 4 2; 1] tau[-2 6; 4 5] tau[5 7; 2 3]
    conj(Ab[-3 6 7; 8]) v[1 3; 8]."""
    g, d, dq = v.shape[1], A.shape[1], A.shape[2]
    return _es("apqi,Gkpf,frqg,bkrj,igj->aGb", A, tau(g, d), tau(g, dq), np.conj(Ab), v)


def tl_utility_excitation(v, A, Ab):
    """transfer.jl:101-104: t[-1 -2; -3 -4] := v[1 2; -3 4] A[4 5; -4] tau[2 3; 5 -2] conj(Ab[1 3; -1])."""
    t = tau(v.shape[1], A.shape[1])
    return _es("mgen,npc,ghpG,mha->aGec", v, A, t, np.conj(Ab))


def tr_utility_excitation(v, A, Ab):
    """transfer.jl:105-108: t[-1 -2; -3 -4] := A[-1 2; 1] tau[-2 4; 2 3] conj(Ab[-4 4; 5]) v[1 3; -3 5]."""
    t = tau(v.shape[1], A.shape[1])
    return _es("api,Gkpg,bkj,igej->aGeb", A, t, np.conj(Ab), v)


# -- two MPO tensors ----------------------------------------------------------

def tl_mpo_pair(v, A, B):
    """transfer.jl:111-113: t[-1; -2] := v[1; 2] A[2 3; 4 -2] conj(B[1 3; 4 -1])."""
    return _es("mn,nhfc,mhfa->ac", v, A, np.conj(B))


def tr_mpo_pair(v, A, B):
    """transfer.jl:114-116: t[-1; -2] := A[-1 3; 4 1] conj(B[-2 3; 4 2]) v[1; 2]."""
    return _es("ahfi,bhfj,ij->ab", A, np.conj(B), v)


# -- MPO transfers -----------------------------------------------------------

def tl_mpo_excited(v, O, A, Ab):
    """transfer.jl:133-135: t[-1 -2; -3 -4] := v[4 2; 1] A[1 3; -3 -4] O[2 5; 3 -2] conj(Ab[4 5; -1])."""
    return _es("mgn,npec,gkpG,mka->aGec", v, A, O, np.conj(Ab))


def tr_mpo_excited(v, O, A, Ab):
    """transfer.jl:136-138: t[-1 -2; -3 -4] := A[-1 4; -3 5] O[-2 2; 4 3] conj(Ab[-4 2; 1]) v[5 3; 1]."""
    return _es("apei,Gkpg,bkj,igj->aGeb", A, O, np.conj(Ab), v)


def tl_mpo_excitation_leg(v, O, A, Ab):
    """transfer.jl:141-143: v[-1 -2; -3 -4] := v[4 2; -3 1] A[1 3; -4] O[2 5; 3 -2] conj(Ab[4 5; -1])."""
    return _es("mgen,npc,gkpG,mka->aGec", v, A, O, np.conj(Ab))


def tr_mpo_excitation_leg(v, O, A, Ab):
    """transfer.jl:144-148: v[-1 -2; -3 -4] := A[-1 4; 5] O[-2 2; 4 3] conj(Ab[-4 2; 1]) v[5 3; -3 1]."""
    return _es("api,Gkpg,bkj,igej->aGeb", A, O, np.conj(Ab), v)


def tl_mpo_mpdo(x, O, A, Ab):
    """transfer.jl:150-155: y[-1 -2; -3] := x[1 2; 6] A[6 7 8; -3] O[2 3; 7 5] tau[5 4; 8 -2]
    conj(Ab[1 3 4; -1])."""
    t = tau(O.shape[3], A.shape[2])
    return _es("mgn,npqc,gkpf,frqG,mkra->aGc", x, A, O, t, np.conj(Ab))


def tr_mpo_mpdo(x, O, A, Ab):
    """transfer.jl:156-161: y[-1 -2; -3] := A[-1 4 2; 1] O[-2 6; 4 5] tau[5 7; 2 3]
    conj(Ab[-3 6 7; 8]) x[1 3; 8]."""
    t = tau(O.shape[3], A.shape[2])
    return _es("apqi,Gkpf,frqg,bkrj,igj->aGb", A, O, t, np.conj(Ab), x)


# -- quasiparticle environments ---------------------------------------------

def finite_qp_environments(Bs, ALs, ARs, Ws, GL, GR):
    """Environments of a finite quasiparticle state, qp_envs.jl:139-162 (0-based
    sites): GL[j] / GR[j] are the ground state's left / right environments of
    site j, Bs[j] the excited tensors (Dl, d, De, Dr), ket = bra ground state.
        lB[0] = 0,  lB[p+1] = lB[p] * T(AR[p], W[p], AL[p]) + GL[p] * T(B[p], W[p], AL[p])
        rB[L-1] = 0, rB[p-1] = T(AL[p], W[p], AR[p]) * rB[p] + T(B[p], W[p], AR[p]) * GR[p]
    lB[p] is (D_bra, w, De, D_ket) on the left bond of site p, rB[p]
    (D_ket, w, De, D_bra) on its right bond."""
    L = len(Ws)
    De = Bs[0].shape[2]
    lB = [np.zeros((ALs[0].shape[0], Ws[0].shape[0], De, ARs[0].shape[0]))]
    for p in range(L - 1):
        lB.append(tl_mpo_excitation_leg(lB[p], Ws[p], ARs[p], ALs[p]) + tl_mpo_excited(GL[p], Ws[p], Bs[p], ALs[p]))
    rB = [None] * L
    rB[L - 1] = np.zeros((ALs[-1].shape[2], Ws[-1].shape[3], De, ARs[-1].shape[2]))
    for p in range(L - 1, 0, -1):
        rB[p - 1] = tr_mpo_excitation_leg(rB[p], Ws[p], ALs[p], ARs[p]) + tr_mpo_excited(GR[p], Ws[p], Bs[p], ARs[p])
    return lB, rB
