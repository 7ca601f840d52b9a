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
    "level_kinds", "infinite_qp_environments",
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


def level_kinds(Ws):
    """Kind of each Jordan level of an infinite MPO cell (mpohamiltonian.jl:677-693):
    "identity" for the first and last level and where every site holds the identity
    on the diagonal, "empty" where some site has no diagonal block, else "diagonal"."""
    w = Ws[0].shape[0]
    kinds = []
    for i in range(w):
        blocks = [W[i, :, :, i] for W in Ws]
        if i == 0 or i == w - 1 or all(np.array_equal(b, np.eye(b.shape[0])) for b in blocks):
            kinds.append("identity")
        elif any(not np.any(b) for b in blocks):
            kinds.append("empty")
        else:
            kinds.append("diagonal")
    return kinds


def _reg(v, Cb):
    """regularize!(v, C', C) on a level slice v[a, e, c] (transfermatrix.jl:79-84 with
    lvec = C', rvec = C): v[a,e,c] -= (sum_ij C[j,i] v[j,e,i]) C[a,c]."""
    lam = np.einsum("ji,jei->e", Cb, v)
    return v - lam[None, :, None] * Cb[:, None, :]


def _dense_solve(op, b, coef):
    """x - coef * op(x) = b by a dense solve over the flattened space (the reference
    runs GMRES, KrylovKit.linsolve(A, b, x0, alg, 1, -coef))."""
    n = b.size
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = e - coef * op(e.reshape(b.shape)).ravel()
    return np.linalg.solve(M, b.ravel()).reshape(b.shape)


def infinite_qp_environments(Bs, AL, AR, C, Ws, GL, GR, momentum):
    """Environments of a quasiparticle on an infinite MPS with a Jordan MPO
    Hamiltonian, trivial excitation (left_gs = right_gs), qp_envs.jl:49-137 with
    left/right_excitation_transfer_system (exci_transfer_system.jl:1-100), for the
    momenta whose phase cis(k) is real (k = 0, pi).  0-based sites: C[s] is the bond
    right of site s (AL[s] C[s] = C[s-1] AR[s]); GL[s] / GR[s] the ground state's
    environments of site s; lB[s] lives on the left bond of site s, rB[s] on the
    right one.  l_RL(loc) = C[loc-1], r_RL(loc) = C[loc]', l_LR(loc) = C[loc-1]',
    r_LR(loc) = C[loc] (infinitemps.jl:356-401) make every regularization here
    regularize!(v, C', C) with the C of v's bond."""
    assert abs(np.sin(momentum)) < 1e-12, "real phases only (momentum 0 or pi)"
    p = float(np.cos(momentum))
    N, w, De = len(Ws), Ws[0].shape[0], Bs[0].shape[2]
    kinds = level_kinds(Ws)
    ids = [i for i in range(1, w - 1) if kinds[i] == "identity"]
    pN = p ** N
    # qp_envs.jl:58-75
    lB = [np.zeros((AL[s].shape[0], Ws[s].shape[0], De, AR[s].shape[0])) for s in range(N)]
    for s in range(N):
        y = (tl_mpo_excitation_leg(lB[s], Ws[s], AR[s], AL[s]) + tl_mpo_excited(GL[s], Ws[s], Bs[s], AL[s])) / p
        for i in ids:
            y[:, i] = _reg(y[:, i], C[s])
        lB[(s + 1) % N] = y
    # qp_envs.jl:77-90
    rB = [np.zeros((AL[s].shape[2], Ws[s].shape[3], De, AR[s].shape[2])) for s in range(N)]
    for s in range(N - 1, -1, -1):
        y = (tr_mpo_excitation_leg(rB[s], Ws[s], AL[s], AR[s]) + tr_mpo_excited(GR[s], Ws[s], Bs[s], AR[s])) * p
        for i in ids:
            y[:, i] = _reg(y[:, i], C[s - 1])
        rB[(s - 1) % N] = y
    Cb = C[N - 1]

    # left_excitation_transfer_system (exci_transfer_system.jl:1-48)
    GBL, found = lB[0], np.zeros_like(lB[0])
    for i in range(w):
        x = found[:, : i + 1]
        for s in range(N):
            x = tl_mpo_excitation_leg(x, Ws[s][: i + 1, :, :, : i + 1], AR[s], AL[s])
        start = x[:, i] * pN
        if kinds[i] == "identity":
            start = _reg(start, Cb)
        found[:, i] = start + GBL[:, i]
        if kinds[i] == "identity":
            def op(v):
                for s in range(N):
                    v = tl_utility(v, AR[s], AL[s])
                return _reg(v, Cb)
        elif kinds[i] == "diagonal":
            def op(v, i=i):
                v = v[:, None]
                for s in range(N):
                    v = tl_mpo_excitation_leg(v, Ws[s][i:i + 1, :, :, i:i + 1], AR[s], AL[s])
                return v[:, 0]
        else:
            continue
        found[:, i] = _dense_solve(op, found[:, i], pN)
    lB[0] = found

    # right_excitation_transfer_system (exci_transfer_system.jl:50-100)
    GBR, found = rB[N - 1], np.zeros_like(rB[N - 1])
    for i in range(w - 1, -1, -1):
        x = found[:, i:]
        for s in range(N - 1, -1, -1):
            x = tr_mpo_excitation_leg(x, Ws[s][i:, :, :, i:], AL[s], AR[s])
        start = x[:, 0] * pN
        if kinds[i] == "identity":
            start = _reg(start, Cb)
        found[:, i] = start + GBR[:, i]
        if kinds[i] == "identity":
            def op(v):
                for s in range(N - 1, -1, -1):
                    v = tr_utility(v, AL[s], AR[s])
                return _reg(v, Cb)
        elif kinds[i] == "diagonal":
            def op(v, i=i):
                v = v[:, None]
                for s in range(N - 1, -1, -1):
                    v = tr_mpo_excitation_leg(v, Ws[s][i:i + 1, :, :, i:i + 1], AL[s], AR[s])
                return v[:, 0]
        else:
            continue
        found[:, i] = _dense_solve(op, found[:, i], pN)
    rB[N - 1] = found

    # qp_envs.jl:110-136
    cur = lB[0]
    for s in range(N - 1):
        cur = tl_mpo_excitation_leg(cur, Ws[s], AR[s], AL[s]) / p
        for i in ids:
            cur[:, i] = _reg(cur[:, i], C[s])
        lB[s + 1] = lB[s + 1] + cur
    cur = rB[N - 1]
    for s in range(N - 1, 0, -1):
        cur = tr_mpo_excitation_leg(cur, Ws[s], AL[s], AR[s]) * p
        for i in ids:
            cur[:, i] = _reg(cur[:, i], C[s - 1])
        rB[s - 1] = rB[s - 1] + cur
    return lB, rB
