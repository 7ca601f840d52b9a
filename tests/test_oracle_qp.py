"""Pins of the quasiparticle / passenger-leg transfer restatements (no GPU).

Each variant of transfer.jl:58-161 is reduced, slice by slice over its
excitation or passenger legs, to the MPS / MPO transfers that tests/test_oracle.py
pins against Kronecker-product matrices and dense Hamiltonians; the
density-matrix MPO transfers are compared with the MPO transfer of the fused
physical leg under O (x) 1 built by loops; left and right variants are tied by
the ring identity; a planted leg exchange fails the pins.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import transfer_qp as Q

RNG = np.random.default_rng(58)
Dm, Dn, Da, Dc, d, De, g, wl, wr, dq = 4, 5, 3, 6, 2, 3, 2, 3, 4, 3


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def r(*shape):
    return RNG.standard_normal(shape)


def fuse_phys(A):
    return A.reshape(A.shape[0], A.shape[1] * A.shape[2], A.shape[3], order="F")


def kron_phys(Op, nq):
    """O (x) 1 on a second physical leg: out[w, k + d q, h + d r, v] = O[w,k,h,v] delta(q,r)."""
    a, dd, _, b = Op.shape
    out = np.zeros((a, dd * nq, dd * nq, b))
    for q in range(nq):
        for k in range(dd):
            for h in range(dd):
                out[:, k + dd * q, h + dd * q, :] = Op[:, k, h, :]
    return out


def test_tau_is_the_leg_exchange():
    t = Q.tau(3, 2)
    x = r(3, 2)
    assert np.allclose(np.einsum("ab,abcd->dc", x, t), x)   # y[d, c] = x[d, c]: the legs swap places
    assert t.sum() == 6


def test_excited_bond_transfers_are_slices_of_mps_transfers():
    v, A, Ab = r(Dm, Dn), r(Dn, d, De, Dc), r(Dm, d, Da)
    y = Q.tl_excited_bond(v, A, Ab)
    assert y.shape == (Da, De, Dc)
    for e in range(De):
        assert rel(y[:, e, :], O.transfer_left_mps(v, A[:, :, e, :], Ab)) < 1e-13
    u, Ar, Abr = r(Dn, Dc), r(Dm, d, De, Dn), r(Da, d, Dc)
    yr = Q.tr_excited_bond(u, Ar, Abr)
    assert yr.shape == (Dm, De, Da)
    for e in range(De):
        assert rel(yr[:, e, :], O.transfer_right_mps(u, Ar[:, :, e, :], Abr)) < 1e-13


def test_excited_transfers_with_a_passing_leg():
    v, A, Ab = r(Dm, g, Dn), r(Dn, d, De, Dc), r(Dm, d, Da)
    y = Q.tl_excited_leg(v, A, Ab)
    assert y.shape == (Da, g, De, Dc)
    for j in range(g):
        for e in range(De):
            assert rel(y[:, j, e, :], O.transfer_left_mps(v[:, j, :], A[:, :, e, :], Ab)) < 1e-13
    u, Ar, Abr = r(Dc, g, Da), r(Dm, d, De, Dc), r(Dn, d, Da)
    yr = Q.tr_excited_leg(u, Ar, Abr)
    assert yr.shape == (Dm, g, De, Dn)
    for j in range(g):
        for e in range(De):
            assert rel(yr[:, j, e, :], O.transfer_right_mps(u[:, j, :], Ar[:, :, e, :], Abr)) < 1e-13


def test_utility_leg_transfers():
    v, A, Ab = r(Dm, g, Dn), r(Dn, d, Dc), r(Dm, d, Da)
    y = Q.tl_utility(v, A, Ab)
    u = r(Dc, g, Da)
    yr = Q.tr_utility(u, A, Ab)
    for j in range(g):
        assert rel(y[:, j, :], O.transfer_left_mps(v[:, j, :], A, Ab)) < 1e-13
        assert rel(yr[:, j, :], O.transfer_right_mps(u[:, j, :], A, Ab)) < 1e-13
    # with an excitation leg too (transfer.jl:101-108)
    v4, u4 = r(Dm, g, De, Dn), r(Dc, g, De, Da)
    y4, yr4 = Q.tl_utility_excitation(v4, A, Ab), Q.tr_utility_excitation(u4, A, Ab)
    for j in range(g):
        for e in range(De):
            assert rel(y4[:, j, e, :], O.transfer_left_mps(v4[:, j, e, :], A, Ab)) < 1e-13
            assert rel(yr4[:, j, e, :], O.transfer_right_mps(u4[:, j, e, :], A, Ab)) < 1e-13


def test_density_matrix_and_mpo_pair_transfers_fuse_the_physical_legs():
    A, Ab = r(Dn, d, dq, Dc), r(Dm, d, dq, Da)
    Af, Abf = fuse_phys(A), fuse_phys(Ab)
    v, u = r(Dm, g, Dn), r(Dc, g, Da)
    y, yr = Q.tl_mpdo_utility(v, A, Ab), Q.tr_mpdo_utility(u, A, Ab)
    for j in range(g):
        assert rel(y[:, j, :], O.transfer_left_mps(v[:, j, :], Af, Abf)) < 1e-13
        assert rel(yr[:, j, :], O.transfer_right_mps(u[:, j, :], Af, Abf)) < 1e-13
    vb, ub = r(Dm, Dn), r(Dc, Da)
    assert rel(Q.tl_mpo_pair(vb, A, Ab), O.transfer_left_mps(vb, Af, Abf)) < 1e-13
    assert rel(Q.tr_mpo_pair(ub, A, Ab), O.transfer_right_mps(ub, Af, Abf)) < 1e-13


def test_mpo_excited_and_excitation_leg_transfers():
    Op = r(wl, d, d, wr)
    x, A, Ab = r(Dm, wl, Dn), r(Dn, d, De, Dc), r(Dm, d, Da)
    y = Q.tl_mpo_excited(x, Op, A, Ab)
    assert y.shape == (Da, wr, De, Dc)
    for e in range(De):
        assert rel(y[:, :, e, :], O.transfer_left_mpo(x, Op, A[:, :, e, :], Ab)) < 1e-13
    v, Ar, Abr = r(Dc, wr, Da), r(Dm, d, De, Dc), r(Dn, d, Da)
    yr = Q.tr_mpo_excited(v, Op, Ar, Abr)
    assert yr.shape == (Dm, wl, De, Dn)
    for e in range(De):
        assert rel(yr[:, :, e, :], O.transfer_right_mpo(v, Op, Ar[:, :, e, :], Abr)) < 1e-13
    x4, A3 = r(Dm, wl, De, Dn), r(Dn, d, Dc)
    y4 = Q.tl_mpo_excitation_leg(x4, Op, A3, Ab)
    assert y4.shape == (Da, wr, De, Dc)
    v4, Ar3 = r(Dc, wr, De, Da), r(Dm, d, Dc)
    yr4 = Q.tr_mpo_excitation_leg(v4, Op, Ar3, Abr)
    assert yr4.shape == (Dm, wl, De, Dn)
    for e in range(De):
        assert rel(y4[:, :, e, :], O.transfer_left_mpo(x4[:, :, e, :], Op, A3, Ab)) < 1e-13
        assert rel(yr4[:, :, e, :], O.transfer_right_mpo(v4[:, :, e, :], Op, Ar3, Abr)) < 1e-13


def test_mpo_density_matrix_transfers_act_on_the_first_physical_leg():
    Op = r(wl, d, d, wr)
    A, Ab = r(Dn, d, dq, Dc), r(Dm, d, dq, Da)
    x = r(Dm, wl, Dn)
    Ok = kron_phys(Op, dq)
    assert rel(Q.tl_mpo_mpdo(x, Op, A, Ab), O.transfer_left_mpo(x, Ok, fuse_phys(A), fuse_phys(Ab))) < 1e-13
    v = r(Dc, wr, Da)
    assert rel(Q.tr_mpo_mpdo(v, Op, A, Ab), O.transfer_right_mpo(v, Ok, fuse_phys(A), fuse_phys(Ab))) < 1e-13


def test_ring_identity_ties_left_and_right_excited_transfers():
    """sum_{a,c} TL(v)[a,G,e,c] u[c,G',a] = sum_{m,n} v[m,g,n] TR(u)[n,G',e,m] for the
    MPO transfers with an excited A: both sides are the same closed ring."""
    Op = r(wl, d, d, wr)
    x, A, Ab = r(Dm, wl, Dn), r(Dn, d, De, Dc), r(Dm, d, Da)
    u = r(Dc, wr, Da)
    lhs = np.einsum("abec,cba->e", Q.tl_mpo_excited(x, Op, A, Ab), u)
    rhs = np.einsum("mgn,ngem->e", x, Q.tr_mpo_excited(u, Op, A, Ab))
    assert rel(lhs, rhs) < 1e-12


def test_planted_leg_exchange_is_caught():
    """Swapping the excitation leg with the right bond in the restatement fails the slice pin."""
    v, A, Ab = r(Dm, Dn), r(Dn, d, De, De), r(Dm, d, Da)
    wrong = np.einsum("mn,nhce,mha->aec", v, A, Ab)
    right = Q.tl_excited_bond(v, A, Ab)
    assert rel(wrong, right) > 1e-3
    assert any(rel(wrong[:, e, :], O.transfer_left_mps(v, A[:, :, e, :], Ab)) > 1e-3 for e in range(De))


# -- finite quasiparticle environments against dense blocks --------------------

def _mpo_left(Ws):
    """Sites 0..p-1 of the MPO as T[S_bra, S_ket, w] with the right MPO bond open."""
    T = np.ones((1, 1, 1))
    for W in Ws:
        nb, nh = T.shape[0], T.shape[1]
        T = np.einsum("BHg,gbhk->BbHhk", T, W).reshape(nb * W.shape[1], nh * W.shape[2], W.shape[3])
    return T


def _mpo_right(Ws):
    """Sites p+1..L-1 of the MPO as T[w, S_bra, S_ket] with the left MPO bond open."""
    T = np.ones((1, 1, 1))
    for W in reversed(Ws):
        T = np.einsum("gbhk,kBH->gbBhH", W, T).reshape(W.shape[0], W.shape[1] * T.shape[1], W.shape[2] * T.shape[2])
    return T


def _ket_left(AL, AR, Bt, j, p):
    """Dense block of sites 0..p-1 with B at j < p, AL left of it and AR right of it: K[S, e, r]."""
    M = np.ones((1, 1))
    for s in range(j):
        M = np.einsum("Sl,lpr->Spr", M, AL[s]).reshape(-1, AL[s].shape[2])
    K = np.einsum("Sl,lper->Sper", M, Bt[j])
    K = K.reshape(-1, K.shape[2], K.shape[3])
    for s in range(j + 1, p):
        K = np.einsum("Sel,lpr->Sper", K, AR[s]).reshape(-1, K.shape[1], AR[s].shape[2])
    return K


def _ket_right(AL, AR, Bt, j, p):
    """Dense block of sites p+1..L-1 with B at j > p, AL left of it and AR right of it: R[l, e, S]."""
    L = len(AL)
    Rm = np.ones((1, 1))
    for s in reversed(range(j + 1, L)):
        Rm = np.einsum("lpr,rS->lpS", AR[s], Rm).reshape(AR[s].shape[0], -1)
    R = np.einsum("lper,rS->lepS", Bt[j], Rm)
    R = R.reshape(R.shape[0], R.shape[1], -1)
    for s in reversed(range(p + 1, j)):
        R = np.einsum("lpr,reS->lepS", AL[s], R).reshape(AL[s].shape[0], R.shape[1], -1)
    return R


def test_finite_qp_environments_vs_dense_blocks():
    """lB[p] = <AL block| H_left |sum_{j<p} AL..B_j..AR block> and
    rB[p] = <AR block| H_right |sum_{j>p} AL..B_j..AR block>, built from dense
    state blocks and the dense MPO (qp_envs.jl:139-162)."""
    from oracle import brute as Bf
    from mpskit_amd import models as M
    L, d, De = 5, 2, 2
    rng = np.random.default_rng(139)
    Ws = M.xxz(L, 0.6)
    psi = Bf.random_mps(L, d, 4, rng)
    AL, AR = Bf.left_canonical(psi), Bf.right_canonical(psi)
    Bs = [rng.standard_normal((A.shape[0], d, De, A.shape[2])) for A in AL]
    GL = [np.ones((1, 1, 1))]
    for j in range(L - 1):
        GL.append(O.transfer_left_mpo(GL[-1], Ws[j], AL[j], AL[j]))
    GRh = [None] * (L + 1)
    GRh[L] = np.ones((1, 1, 1))
    for j in range(L - 1, 0, -1):
        GRh[j] = O.transfer_right_mpo(GRh[j + 1], Ws[j], AR[j], AR[j])
    lB, rB = Q.finite_qp_environments(Bs, AL, AR, Ws, GL, GRh[1:])
    assert not np.any(lB[0]) and not np.any(rB[L - 1])
    for p in range(1, L):
        U = Bf.left_block(AL[:p])
        K = sum(_ket_left(AL, AR, Bs, j, p) for j in range(p))
        want = np.einsum("Sa,STw,Ten->awen", U, _mpo_left(Ws[:p]), K)
        assert rel(lB[p], want) < 1e-12
    for p in range(L - 1):
        V = Bf.right_block(AR[p + 1:])
        R = sum(_ket_right(AL, AR, Bs, j, p) for j in range(p + 1, L))
        want = np.einsum("ieS,wTS,bT->iweb", R, _mpo_right(Ws[p + 1:]), V)
        assert rel(rB[p], want) < 1e-12


# -- infinite quasiparticle environments -----------------------------------------

def _uniform_mps(N, d, D, rng):
    """AL (left isometries), C (bond right of each site, C C^T the right fixed point of
    the AL transfer, |C| = 1, not symmetric) and AR = C[s-1]^-1 AL C[s], so
    AL[s] C[s] = C[s-1] AR[s] and AR is right-isometric."""
    AL = []
    for _ in range(N):
        Q, _ = np.linalg.qr(rng.standard_normal((D * d, D)))
        AL.append(Q.reshape(D, d, D, order="F"))

    def E(s, R):                                   # right action of site s: sum_h AL_h R AL_h^T
        return np.einsum("lhr,rs,mhs->lm", AL[s], R, AL[s])

    R = np.eye(D) / D
    for _ in range(5000):
        Rn = R
        for s in range(N - 1, -1, -1):
            Rn = E(s, Rn)
        Rn = Rn / np.trace(Rn)
        if np.linalg.norm(Rn - R) < 1e-15:
            break
        R = Rn
    Rs = [None] * N
    Rs[N - 1] = R
    for s in range(N - 1, 0, -1):
        Rs[s - 1] = E(s, Rs[s])
    C = []
    for Rb in Rs:
        lam, V = np.linalg.eigh(0.5 * (Rb + Rb.T))
        C.append((V * np.sqrt(lam)) @ V.T)
    # a random orthogonal gauge on every bond makes C non-symmetric (C -> C U, AR -> U^T AR U)
    U = [np.linalg.qr(rng.standard_normal((D, D)))[0] for _ in range(N)]
    C = [C[s] @ U[s] for s in range(N)]
    AR = [np.einsum("lm,mpr,rk->lpk", np.linalg.inv(C[s - 1]), AL[s], C[s]) for s in range(N)]
    return AL, AR, C


def _series_left(S, AL, AR, C, Ws, pN, kinds):
    """x = Reg_id(p^N x T_cell) + S by fixed-point iteration of the whole cell map."""
    N, x = len(Ws), np.zeros_like(S)
    for it in range(20000):
        y = x
        for s in range(N):
            y = Q.tl_mpo_excitation_leg(y, Ws[s], AR[s], AL[s])
        y = pN * y
        for i, k in enumerate(kinds):
            if k == "identity":
                y[:, i] = Q._reg(y[:, i], C[N - 1])
        xn = y + S
        if np.linalg.norm(xn - x) <= 1e-15 * max(np.linalg.norm(xn), 1.0):
            return xn, it
        x = xn
    raise AssertionError("series did not converge")


def _series_right(S, AL, AR, C, Ws, pN, kinds):
    N, x = len(Ws), np.zeros_like(S)
    for it in range(20000):
        y = x
        for s in range(N - 1, -1, -1):
            y = Q.tr_mpo_excitation_leg(y, Ws[s], AL[s], AR[s])
        y = pN * y
        for i, k in enumerate(kinds):
            if k == "identity":
                y[:, i] = Q._reg(y[:, i], C[N - 1])
        xn = y + S
        if np.linalg.norm(xn - x) <= 1e-15 * max(np.linalg.norm(xn), 1.0):
            return xn, it
        x = xn
    raise AssertionError("series did not converge")


@pytest.mark.parametrize("model", ["xxz", "exp_ising", "exp_ising_uniform"])
@pytest.mark.parametrize("N", [1, 2])
@pytest.mark.parametrize("momentum", [0.0, np.pi])
def test_infinite_qp_environments_solve_the_cell_recursion(model, N, momentum):
    """The per-level triangular solves (exci_transfer_system.jl) against the fixed point
    of the whole-cell map x = Reg(p^N x T_cell) + S iterated to convergence, where S is
    the cell recursion of the excited tensors from zero (qp_envs.jl:58-90); the
    regularizing C are checked to be the fixed points of the mixed transfers.  Models:
    XXZ (empty interior levels), exponential Ising (a diagonal level) and its lambda = 1
    version (an interior identity level, regularized inside the recursion)."""
    from mpskit_amd import models as M
    rng = np.random.default_rng(N * 10 + int(momentum) + len(model))
    d, D, De = 2, 4, 2
    W = {"xxz": lambda: M.xxz_bulk(0.7), "exp_ising": lambda: M.exp_ising_bulk(1.0, 0.5, 0.8),
         "exp_ising_uniform": lambda: M.exp_ising_bulk(1.0, 1.0, 0.8)}[model]()
    Ws = [W] * N
    w = W.shape[0]
    kinds = Q.level_kinds(Ws)
    if model == "exp_ising_uniform":
        assert "identity" in kinds[1:-1]
    AL, AR, C = _uniform_mps(N, d, D, rng)
    # the regularizations' vectors: C is the left fixed point of x T(AR, AL) and the right one of T(AL, AR)
    x, y = C[N - 1], C[N - 1]
    for s in range(N):
        x = O.transfer_left_mps(x, AR[s], AL[s])
    for s in range(N - 1, -1, -1):
        y = O.transfer_right_mps(y, AL[s], AR[s])
    assert rel(x, C[N - 1]) < 1e-12 and rel(y, C[N - 1]) < 1e-12
    # and C^T the other fixed point of each: the right one of T(AR, AL), the left one of T(AL, AR)
    x, y = C[N - 1].T, C[N - 1].T
    for s in range(N - 1, -1, -1):
        x = O.transfer_right_mps(x, AR[s], AL[s])
    for s in range(N):
        y = O.transfer_left_mps(y, AL[s], AR[s])
    assert rel(x, C[N - 1].T) < 1e-12 and rel(y, C[N - 1].T) < 1e-12
    assert abs(np.trace(C[N - 1].T @ C[N - 1]) - 1.0) < 1e-12
    # regularize! is the projection that removes the fixed point C along its dual C^T:
    # Reg(C) = 0 and tr(Reg(v)_e C^T ... ) pairs to zero with the dual
    v = rng.standard_normal((D, De, D))
    Cb = C[N - 1]
    assert np.abs(Q._reg(np.repeat(Cb[:, None, :], De, axis=1), Cb)).max() < 1e-13
    assert np.abs(np.einsum("aec,ca->e", Q._reg(v, Cb), Cb.T)).max() < 1e-13
    GL = [rng.standard_normal((D, w, D)) for _ in range(N)]
    GR = [rng.standard_normal((D, w, D)) for _ in range(N)]
    Bs = [rng.standard_normal((D, d, De, D)) for _ in range(N)]
    lB, rB = Q.infinite_qp_environments(Bs, AL, AR, C, Ws, GL, GR, momentum)
    p = float(np.cos(momentum))
    ids = [i for i in range(1, w - 1) if kinds[i] == "identity"]
    # S: the cell recursion of the excited tensors from zero
    Sl = np.zeros((D, w, De, D))
    for s in range(N):
        Sl = (Q.tl_mpo_excitation_leg(Sl, Ws[s], AR[s], AL[s]) + Q.tl_mpo_excited(GL[s], Ws[s], Bs[s], AL[s])) / p
        for i in ids:
            Sl[:, i] = Q._reg(Sl[:, i], C[s])
    Sr = np.zeros((D, w, De, D))
    for s in range(N - 1, -1, -1):
        Sr = (Q.tr_mpo_excitation_leg(Sr, Ws[s], AL[s], AR[s]) + Q.tr_mpo_excited(GR[s], Ws[s], Bs[s], AR[s])) * p
        for i in ids:
            Sr[:, i] = Q._reg(Sr[:, i], C[s - 1])
    xl, _ = _series_left(Sl, AL, AR, C, Ws, p ** N, kinds)
    xr, _ = _series_right(Sr, AL, AR, C, Ws, p ** N, kinds)
    assert rel(lB[0], xl) < 1e-10
    assert rel(rB[N - 1], xr) < 1e-10
    # the other sites carry the same environments moved through the cell
    for s in range(N - 1):
        y = (Q.tl_mpo_excitation_leg(lB[s], Ws[s], AR[s], AL[s]) + Q.tl_mpo_excited(GL[s], Ws[s], Bs[s], AL[s])) / p
        for i in ids:
            y[:, i] = Q._reg(y[:, i], C[s])
        assert rel(lB[s + 1], y) < 1e-10
