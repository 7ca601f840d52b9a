"""Quasiparticle, passenger-leg, density-matrix and MPO-pair transfers
(transfer.jl:58-161) on the HIP path against the oracle's restatements
(``oracle.transfer_qp``), FP64 to 1e-12 relative: every method of the
reference's table through ``transfer_left`` / ``transfer_right``, excitation
legs of size 1 (the usual quasiparticle) and larger (the leg exchange of the
excited right transfers), odd bond dimensions, and a D = 512 case checked
through the oracle and the ring identity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import transfer_qp as Q  # noqa: E402

TOL = 1e-12


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev(mk, a):
    return mk.tensors.from_numpy(a)


def host(mk, t):
    return mk.tensors.to_numpy(t)


def rnd(rng, *shape):
    return rng.standard_normal(shape) / np.sqrt(max(shape[0], 1))


SIZES = [(37, 41, 39, 43, 1), (37, 41, 39, 43, 3), (64, 64, 64, 64, 2)]


@pytest.mark.parametrize("Dm,Dn,Dc,Da,De", SIZES)
def test_excited_mps_transfers(mk, Dm, Dn, Dc, Da, De):
    rng = np.random.default_rng(Dm + 7 * De)
    d, g = 2, 3
    print("excited mps transfers", Dm, Dn, Dc, Da, De, flush=True)
    v, A, Ab = rnd(rng, Dm, Dn), rnd(rng, Dn, d, De, Dc), rnd(rng, Dm, d, Da)
    y = host(mk, mk.transfer_left(dev(mk, v), dev(mk, A), dev(mk, Ab)))
    assert rel(y, Q.tl_excited_bond(v, A, Ab)) < TOL
    u, Ar, Abr = rnd(rng, Dn, Dc), rnd(rng, Dm, d, De, Dn), rnd(rng, Da, d, Dc)
    yr = host(mk, mk.transfer_right(dev(mk, u), dev(mk, Ar), dev(mk, Abr)))
    assert yr.shape == (Dm, De, Da)
    assert rel(yr, Q.tr_excited_bond(u, Ar, Abr)) < TOL
    # with a passing leg (transfer.jl:66-73)
    vg = rnd(rng, Dm, g, Dn)
    y = host(mk, mk.transfer_left(dev(mk, vg), dev(mk, A), dev(mk, Ab)))
    assert y.shape == (Da, g, De, Dc)
    assert rel(y, Q.tl_excited_leg(vg, A, Ab)) < TOL
    ug = rnd(rng, Dn, g, Dc)
    yr = host(mk, mk.transfer_right(dev(mk, ug), dev(mk, Ar), dev(mk, Abr)))
    assert yr.shape == (Dm, g, De, Da)
    assert rel(yr, Q.tr_excited_leg(ug, Ar, Abr)) < TOL


@pytest.mark.parametrize("Dm,Dn,Dc,Da,De", SIZES)
def test_utility_leg_transfers(mk, Dm, Dn, Dc, Da, De):
    rng = np.random.default_rng(Dn + 11 * De)
    d, g = 3, 2
    print("utility leg transfers", Dm, Dn, Dc, Da, De, flush=True)
    A, Ab = rnd(rng, Dn, d, Dc), rnd(rng, Dm, d, Da)
    v, u = rnd(rng, Dm, g, Dn), rnd(rng, Dc, g, Da)
    assert rel(host(mk, mk.transfer_left(dev(mk, v), dev(mk, A), dev(mk, Ab))), Q.tl_utility(v, A, Ab)) < TOL
    assert rel(host(mk, mk.transfer_right(dev(mk, u), dev(mk, A), dev(mk, Ab))), Q.tr_utility(u, A, Ab)) < TOL
    v4, u4 = rnd(rng, Dm, g, De, Dn), rnd(rng, Dc, g, De, Da)
    y4 = host(mk, mk.transfer_left(dev(mk, v4), dev(mk, A), dev(mk, Ab)))
    assert y4.shape == (Da, g, De, Dc)
    assert rel(y4, Q.tl_utility_excitation(v4, A, Ab)) < TOL
    yr4 = host(mk, mk.transfer_right(dev(mk, u4), dev(mk, A), dev(mk, Ab)))
    assert yr4.shape == (Dn, g, De, Dm)
    assert rel(yr4, Q.tr_utility_excitation(u4, A, Ab)) < TOL


@pytest.mark.parametrize("Dm,Dn,Dc,Da,De", SIZES[:2])
def test_density_matrix_and_mpo_pair_transfers(mk, Dm, Dn, Dc, Da, De):
    rng = np.random.default_rng(Dc + 13 * De)
    d, dq, g, wl, wr = 2, 2, 2, 3, 4
    print("density matrix transfers", Dm, Dn, Dc, Da, flush=True)
    A, Ab = rnd(rng, Dn, d, dq, Dc), rnd(rng, Dm, d, dq, Da)
    v, u = rnd(rng, Dm, g, Dn), rnd(rng, Dc, g, Da)
    assert rel(host(mk, mk.transfer_left(dev(mk, v), dev(mk, A), dev(mk, Ab))), Q.tl_mpdo_utility(v, A, Ab)) < TOL
    assert rel(host(mk, mk.transfer_right(dev(mk, u), dev(mk, A), dev(mk, Ab))),
               Q.tr_mpdo_utility(u, A, Ab)) < TOL
    vb, ub = rnd(rng, Dm, Dn), rnd(rng, Dc, Da)
    assert rel(host(mk, mk.transfer_left(dev(mk, vb), dev(mk, A), dev(mk, Ab))), Q.tl_mpo_pair(vb, A, Ab)) < TOL
    assert rel(host(mk, mk.transfer_right(dev(mk, ub), dev(mk, A), dev(mk, Ab))), Q.tr_mpo_pair(ub, A, Ab)) < TOL
    Op = rng.standard_normal((wl, d, d, wr))
    x, vr = rnd(rng, Dm, wl, Dn), rnd(rng, Dc, wr, Da)
    assert rel(host(mk, mk.transfer_left(dev(mk, x), dev(mk, Op), dev(mk, A), dev(mk, Ab))),
               Q.tl_mpo_mpdo(x, Op, A, Ab)) < TOL
    assert rel(host(mk, mk.transfer_right(dev(mk, vr), dev(mk, Op), dev(mk, A), dev(mk, Ab))),
               Q.tr_mpo_mpdo(vr, Op, A, Ab)) < TOL


@pytest.mark.parametrize("Dm,Dn,Dc,Da,De", SIZES)
@pytest.mark.parametrize("wl,wr", [(1, 1), (5, 5), (3, 4)])
def test_mpo_excited_and_excitation_leg_transfers(mk, Dm, Dn, Dc, Da, De, wl, wr):
    rng = np.random.default_rng(Da + 17 * De + wl)
    d = 2
    print("mpo excited transfers", Dm, Dn, Dc, Da, De, wl, wr, flush=True)
    Op = rng.standard_normal((wl, d, d, wr))
    Op[rng.random(Op.shape) < 0.4] = 0.0          # MPO sparsity as in a Jordan MPO
    x, A, Ab = rnd(rng, Dm, wl, Dn), rnd(rng, Dn, d, De, Dc), rnd(rng, Dm, d, Da)
    y = host(mk, mk.transfer_left(dev(mk, x), dev(mk, Op), dev(mk, A), dev(mk, Ab)))
    assert y.shape == (Da, wr, De, Dc)
    assert rel(y, Q.tl_mpo_excited(x, Op, A, Ab)) < TOL
    v, Ar, Abr = rnd(rng, Dc, wr, Da), rnd(rng, Dm, d, De, Dc), rnd(rng, Dn, d, Da)
    yr = host(mk, mk.transfer_right(dev(mk, v), dev(mk, Op), dev(mk, Ar), dev(mk, Abr)))
    assert yr.shape == (Dm, wl, De, Dn)
    assert rel(yr, Q.tr_mpo_excited(v, Op, Ar, Abr)) < TOL
    x4, A3 = rnd(rng, Dm, wl, De, Dn), rnd(rng, Dn, d, Dc)
    y4 = host(mk, mk.transfer_left(dev(mk, x4), dev(mk, Op), dev(mk, A3), dev(mk, Ab)))
    assert y4.shape == (Da, wr, De, Dc)
    assert rel(y4, Q.tl_mpo_excitation_leg(x4, Op, A3, Ab)) < TOL
    v4, Ar3 = rnd(rng, Dc, wr, De, Da), rnd(rng, Dm, d, Dc)
    yr4 = host(mk, mk.transfer_right(dev(mk, v4), dev(mk, Op), dev(mk, Ar3), dev(mk, Abr)))
    assert yr4.shape == (Dm, wl, De, Dn)
    assert rel(yr4, Q.tr_mpo_excitation_leg(v4, Op, Ar3, Abr)) < TOL


def test_excited_mpo_transfers_at_D512(mk):
    """D = 512, w = 5, De = 2: the oracle and the ring identity
    sum TL(x)[a,b,e,c] u[c,b,a] = sum x[m,g,n] TR(u)[n,g,e,m]."""
    rng = np.random.default_rng(512)
    D, w, d, De = 512, 5, 2, 2
    print("excited mpo transfers D512", flush=True)
    Op = rng.standard_normal((w, d, d, w))
    x, A, Ab, u = rnd(rng, D, w, D), rnd(rng, D, d, De, D), rnd(rng, D, d, D), rnd(rng, D, w, D)
    y = host(mk, mk.transfer_left(dev(mk, x), dev(mk, Op), dev(mk, A), dev(mk, Ab)))
    yr = host(mk, mk.transfer_right(dev(mk, u), dev(mk, Op), dev(mk, A), dev(mk, Ab)))
    assert rel(y, Q.tl_mpo_excited(x, Op, A, Ab)) < TOL
    assert rel(yr, Q.tr_mpo_excited(u, Op, A, Ab)) < TOL
    lhs = np.einsum("abec,cba->e", y, u)
    rhs = np.einsum("mgn,ngem->e", x, yr)
    assert rel(lhs, rhs) < 1e-11


def test_mpo_extend_entry_point(mk):
    """``mk_mpo_extend``: O (x) 1 on a virtual and a physical leg, and the identity MPO."""
    from mpskit_amd._lib import load
    from mpskit_amd.context import default_context
    rng = np.random.default_rng(3)
    wl, d, wr, ne, nq = 3, 2, 4, 2, 3
    Op = rng.standard_normal((wl, d, d, wr))
    want = np.zeros((wl * ne, d * nq, d * nq, wr * ne))
    for e in range(ne):
        for q in range(nq):
            for k in range(d):
                for h in range(d):
                    want[e * wl:(e + 1) * wl, k + d * q, h + d * q, e * wr:(e + 1) * wr] = Op[:, k, h, :]
    T = mk.tensors
    out = T.empty(want.shape)
    ctx = default_context()
    assert load().mk_mpo_extend(ctx.handle, T.ptr(dev(mk, Op)), wl, d, wr, ne, nq, T.ptr(out)) == 0
    assert np.array_equal(host(mk, out), want)
    ident = T.empty((wl, d, d, wl))
    assert load().mk_mpo_extend(ctx.handle, None, wl, d, wl, 1, 1, T.ptr(ident)) == 0
    eye = np.einsum("ab,kh->akhb", np.eye(wl), np.eye(d))
    assert np.array_equal(host(mk, ident), eye)
    assert load().mk_mpo_extend(ctx.handle, None, wl, d, wr, 1, 1, T.ptr(ident)) != 0   # identity needs wl == wr


@pytest.mark.parametrize("De", [1, 2])
def test_finite_qp_environments(mk, De):
    """``qp.finite_qp_environments`` (qp_envs.jl:139-162) against the oracle on an
    XXZ chain of 8 sites with bonds up to 64 (ground state in both gauges)."""
    from oracle import brute as Bf
    from oracle import oracle as O
    from mpskit_amd import models as M
    print("finite qp environments", De, flush=True)
    L, d = 8, 2
    rng = np.random.default_rng(8 + De)
    Ws = M.xxz(L, 0.6)
    psi = Bf.random_mps(L, d, 64, rng)
    AL, AR = Bf.left_canonical(psi), Bf.right_canonical(psi)
    Bs = [rng.standard_normal((A.shape[0], d, De, A.shape[2])) for A in AL]
    GL = [np.ones((1, 1, 1))]
    for j in range(L - 1):
        GL.append(O.transfer_left_mpo(GL[-1], Ws[j], AL[j], AL[j]))
    GR = [None] * L
    GR[L - 1] = np.ones((1, 1, 1))
    for j in range(L - 1, 0, -1):
        GR[j - 1] = O.transfer_right_mpo(GR[j], Ws[j], AR[j], AR[j])
    lB, rB = Q.finite_qp_environments(Bs, AL, AR, Ws, GL, GR)
    up = lambda xs: [dev(mk, x) for x in xs]  # noqa: E731
    gl, gr = mk.qp.ground_state_environments(up(AL), up(AR), up(Ws))
    for j in range(L):
        assert rel(host(mk, gl[j]), GL[j]) < TOL and rel(host(mk, gr[j]), GR[j]) < TOL
    lb, rb = mk.qp.finite_qp_environments(up(Bs), up(AL), up(AR), up(Ws))
    for p in range(L):
        assert host(mk, lb[p]).shape == lB[p].shape and host(mk, rb[p]).shape == rB[p].shape
        assert rel(host(mk, lb[p]), lB[p]) < TOL if p > 0 else not np.any(host(mk, lb[p]))
        assert rel(host(mk, rb[p]), rB[p]) < TOL if p < L - 1 else not np.any(host(mk, rb[p]))


def _uniform_host(N, d, D, rng):
    """Uniform MPS on the host: AL left isometries, C (bond right of each site,
    non-symmetric) with C C^T the AL transfer's right fixed point, AR = C^-1 AL C."""
    AL = [np.linalg.qr(rng.standard_normal((D * d, D)))[0].reshape(D, d, D, order="F") for _ in range(N)]

    def E(s, R):
        return np.einsum("lhr,rs,mhs->lm", AL[s], R, AL[s])

    R = np.eye(D) / D
    for _ in range(5000):
        Rn = R
        for s in range(N - 1, -1, -1):
            Rn = E(s, Rn)
        Rn = Rn / np.trace(Rn)
        done = np.linalg.norm(Rn - R) < 1e-15
        R = Rn
        if done:
            break
    Rs = [None] * N
    Rs[N - 1] = R
    for s in range(N - 1, 0, -1):
        Rs[s - 1] = E(s, Rs[s])
    C = []
    for s in range(N):
        lam, V = np.linalg.eigh(0.5 * (Rs[s] + Rs[s].T))
        U = np.linalg.qr(rng.standard_normal((D, D)))[0]
        C.append(((V * np.sqrt(lam)) @ V.T) @ U)
    AR = [np.einsum("lm,mpr,rk->lpk", np.linalg.inv(C[s - 1]), AL[s], C[s]) for s in range(N)]
    return AL, AR, C


@pytest.mark.parametrize("model", ["xxz", "exp_ising", "exp_ising_uniform"])
@pytest.mark.parametrize("N,De", [(1, 2), (2, 1)])
@pytest.mark.parametrize("momentum", [0.0, np.pi])
def test_infinite_qp_environments(mk, model, N, De, momentum):
    """``qp.infinite_qp_environments`` (qp_envs.jl:49-137; GMRES on the HIP transfers)
    against the oracle's dense per-level solves, 1e-9 relative: both momenta with a
    real phase, cells of 1 and 2 sites, excitation legs of 1 and 2, XXZ (empty
    interior levels), exponential Ising (a diagonal level) and its lambda = 1 form
    (an interior identity level)."""
    from types import SimpleNamespace
    from mpskit_amd import models as M
    from mpskit_amd import vumps as V
    print("infinite qp environments", model, N, De, momentum, flush=True)
    rng = np.random.default_rng(7 * N + De + len(model))
    d, D = 2, 12
    W = {"xxz": lambda: M.xxz_bulk(0.7), "exp_ising": lambda: M.exp_ising_bulk(1.0, 0.5, 0.8),
         "exp_ising_uniform": lambda: M.exp_ising_bulk(1.0, 1.0, 0.8)}[model]()
    Ws = [W] * N
    w = W.shape[0]
    AL, AR, C = _uniform_host(N, d, D, rng)
    GL = [rng.standard_normal((D, w, D)) for _ in range(N)]
    GR = [rng.standard_normal((D, w, D)) for _ in range(N)]
    Bs = [rng.standard_normal((D, d, De, D)) for _ in range(N)]
    lB, rB = Q.infinite_qp_environments(Bs, AL, AR, C, Ws, GL, GR, momentum)
    up = lambda xs: [dev(mk, x) for x in xs]  # noqa: E731
    psi = V.InfiniteMPS(up(AL), up(AR), up(C))
    envs = SimpleNamespace(GLs=up(GL), GRs=up(GR))
    lb, rb = mk.qp.infinite_qp_environments(up(Bs), psi, V.InfiniteHamiltonian(Ws), envs, momentum)
    for s in range(N):
        assert rel(host(mk, lb[s]), lB[s]) < 1e-9, s
        assert rel(host(mk, rb[s]), rB[s]) < 1e-9, s


def test_infinite_qp_environments_of_a_ground_state_environment(mk):
    """The same on a random uniform MPS with its own infinite environments
    (``InfiniteEnvironments``: per-level GMRES fixed points), TFIM, D = 16, cell 2."""
    from mpskit_amd import models as M
    from mpskit_amd import vumps as V
    print("infinite qp environments, ground-state environments", flush=True)
    rng = np.random.default_rng(11)
    Ws = [M.tfim_bulk(0.8)] * 2
    H = V.InfiniteHamiltonian(Ws)
    psi = V.InfiniteMPS.random(2, 2, 16, seed=4)
    envs = V.InfiniteEnvironments(psi, H)
    Bs = [rng.standard_normal((A.shape[0], 2, 1, A.shape[2])) for A in psi.AL]
    lb, rb = mk.qp.infinite_qp_environments([dev(mk, b) for b in Bs], psi, H, envs, 0.0)
    hostl = lambda xs: [host(mk, x) for x in xs]  # noqa: E731
    lB, rB = Q.infinite_qp_environments(Bs, hostl(psi.AL), hostl(psi.AR), hostl(psi.C), Ws, hostl(envs.GLs),
                                        hostl(envs.GRs), 0.0)
    for s in range(2):
        assert rel(host(mk, lb[s]), lB[s]) < 1e-9 and rel(host(mk, rb[s]), rB[s]) < 1e-9
    with pytest.raises(NotImplementedError):
        mk.qp.infinite_qp_environments([dev(mk, b) for b in Bs], psi, H, envs, 0.3)
