"""Pins of the CPU oracle (no GPU).

Every oracle routine is checked against something that does not share its
index strings: dense Hamiltonians built from Kronecker products, projected on
random mixed-canonical MPS (bra != ket for the generic operators), closed-form
energies, and algebraic identities.  A planted-mistake test shows the pins
catch a transposed leg.
"""
import numpy as np
import pytest

from oracle import brute as B
from oracle import oracle as O
from mpskit_amd import models as M

L = 7
CHAINS = {
    "tfim": (lambda: M.tfim(L, 0.7), lambda: B.dense_tfim(L, 0.7)),
    "xxz": (lambda: M.xxz(L, 0.6), lambda: B.dense_xxz(L, 0.6)),
    "j1j2": (lambda: M.j1j2(L, 1.0, 0.3), lambda: B.dense_j1j2(L, 1.0, 0.3)),
    # range 3 and infinite range: the A blocks of neighbouring sites connect, so the
    # Jordan two-site operator carries the AA term (hamiltonian_derivatives.jl:352-362)
    "heis3": (lambda: M.heisenberg_range(L, [1.0, 0.45, 0.3]), lambda: B.dense_heisenberg_range(L, [1.0, 0.45, 0.3])),
    "expising": (lambda: M.exp_ising(L, 0.8, 0.6, 0.9), lambda: B.dense_exp_ising(L, 0.8, 0.6, 0.9)),
}


def envs(Ws, ket, bra):
    ALk, ARk = B.left_canonical(ket), B.right_canonical(ket)
    ALb, ARb = B.left_canonical(bra), B.right_canonical(bra)
    GL = [np.ones((1, 1, 1))]
    for j in range(len(Ws)):
        GL.append(O.transfer_left_mpo(GL[-1], Ws[j], ALk[j], ALb[j]))
    GR = [None] * (len(Ws) + 1)
    GR[-1] = np.ones((1, 1, 1))
    for j in range(len(Ws) - 1, -1, -1):
        GR[j] = O.transfer_right_mpo(GR[j + 1], Ws[j], ARk[j], ARb[j])
    return (ALk, ARk, ALb, ARb), GL, GR


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.mark.parametrize("name", list(CHAINS))
def test_mpo_matches_kron_hamiltonian(name):
    Ws, H = CHAINS[name][0](), CHAINS[name][1]()
    assert np.abs(B.mpo_to_dense(Ws) - H.toarray()).max() < 1e-13


@pytest.mark.parametrize("name", list(CHAINS))
def test_generic_applies_vs_projected_dense_H(name):
    """P_bra^T H P_ket x == derivative apply, with bra != ket (mpo_derivatives.jl)."""
    rng = np.random.default_rng(11)
    Ws, H = CHAINS[name][0](), CHAINS[name][1]().toarray()
    ket, bra = B.random_mps(L, 2, 6, rng), B.random_mps(L, 2, 6, rng)
    (ALk, ARk, ALb, ARb), GL, GR = envs(Ws, ket, bra)
    for j in range(L - 1):
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 2].shape[0], 2))
        y = O.ac2_apply_generic(GL[j], Ws[j], Ws[j + 1], GR[j + 2], x)
        yb = B.isometry_ac2(ALb, ARb, j, 2).T @ H @ B.isometry_ac2(ALk, ARk, j, 2) @ x.reshape(-1, order="F")
        assert rel(y.reshape(-1, order="F"), yb) < 1e-13
        Lm, Rm = O.prepare_ac2_generic(GL[j], Ws[j], Ws[j + 1], GR[j + 2])
        assert rel(O.apply_prepared_ac2(Lm, Rm, x), y) < 1e-13
    for j in range(L):
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 1].shape[0]))
        y = O.ac_apply_generic(GL[j], Ws[j], GR[j + 1], x)
        yb = B.isometry_ac(ALb, ARb, j, 2).T @ H @ B.isometry_ac(ALk, ARk, j, 2) @ x.reshape(-1, order="F")
        assert rel(y.reshape(-1, order="F"), yb) < 1e-13
        Lm, Rm = O.prepare_ac_generic(GL[j], Ws[j], GR[j + 1])
        assert rel(O.apply_prepared_ac(Lm, Rm, x), y) < 1e-13
    for j in range(1, L):
        x = rng.standard_normal((GL[j].shape[2], GR[j].shape[0]))
        y = O.c_apply(GL[j], GR[j], x)
        yb = B.isometry_c(ALb, ARb, j).T @ H @ B.isometry_c(ALk, ARk, j) @ x.reshape(-1, order="F")
        assert rel(y.reshape(-1, order="F"), yb) < 1e-13
        Lm, Rm = O.prepare_c(GL[j], GR[j])
        assert rel(O.apply_prepared_c(Lm, Rm, x), y) < 1e-13


@pytest.mark.parametrize("name", list(CHAINS))
def test_jordan_applies_vs_projected_dense_H(name):
    """JordanMPO_AC2/AC (hamiltonian_derivatives.jl) with bra == ket, every bond."""
    rng = np.random.default_rng(12)
    Ws, H = CHAINS[name][0](), CHAINS[name][1]().toarray()
    psi = B.random_mps(L, 2, 6, rng)
    (AL, AR, _, _), GL, GR = envs(Ws, psi, psi)
    for j in range(L - 1):
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 2].shape[0], 2))
        P = B.isometry_ac2(AL, AR, j, 2)
        T = O.jordan_ac2_prepare(GL[j], Ws[j], Ws[j + 1], GR[j + 2])
        assert rel(O.jordan_ac2_apply(T, x).reshape(-1, order="F"), P.T @ H @ P @ x.reshape(-1, order="F")) < 1e-13
    for j in range(L):
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 1].shape[0]))
        P = B.isometry_ac(AL, AR, j, 2)
        T = O.jordan_ac_prepare(GL[j], Ws[j], GR[j + 1])
        assert rel(O.jordan_ac_apply(T, x).reshape(-1, order="F"), P.T @ H @ P @ x.reshape(-1, order="F")) < 1e-13


@pytest.mark.parametrize("name", ["heis3", "expising"])
def test_jordan_aa_term_connected_channels(name):
    """The continuing-continuing term over ``_connected_channels``
    (hamiltonian_derivatives.jl:243-270, :352-362): on every bulk bond of a range-3
    chain the channel sets are the pass-through blocks, the AA term is present, and
    dropping it (a plausible omission) breaks P^T H P, so the projected-H pin above
    covers the AA branch."""
    rng = np.random.default_rng(13)
    Ws, H = CHAINS[name][0](), CHAINS[name][1]().toarray()
    psi = B.random_mps(L, 2, 6, rng)
    (AL, AR, _, _), GL, GR = envs(Ws, psi, psi)
    n_aa = 0
    for j in range(L - 1):
        T = O.jordan_ac2_prepare(GL[j], Ws[j], Ws[j + 1], GR[j + 2])
        b1, b2 = O.jordan_split(Ws[j]), O.jordan_split(Ws[j + 1])
        ch = O._connected_channels(b1["A"], b2["A"])
        bulk = 1 <= j and j + 1 <= L - 2
        assert (ch is not None) == bulk and (T["AA"] is not None) == bulk
        if not bulk:
            continue
        n_aa += 1
        rows, mids, cols = ch
        if name == "heis3":          # block 0 -> block 1 -> block 2 of the mid channels
            assert list(rows) == [0, 1, 2] and list(mids) == [3, 4, 5] and list(cols) == [6, 7, 8]
        else:                        # the single channel feeds itself
            assert list(rows) == [0] and list(mids) == [0] and list(cols) == [0]
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 2].shape[0], 2))
        P = B.isometry_ac2(AL, AR, j, 2)
        yb = P.T @ H @ P @ x.reshape(-1, order="F")
        assert rel(O.jordan_ac2_apply(T, x).reshape(-1, order="F"), yb) < 1e-13
        no_aa = dict(T, AA=None)
        assert rel(O.jordan_ac2_apply(no_aa, x).reshape(-1, order="F"), yb) > 1e-3
    assert n_aa == L - 3


def test_jordan_aa_term_on_random_environments():
    """AA with i.i.d. environments (the reference benchmark's setup, e.g. the
    heisenberg_nnn / coulomb specs): Jordan == generic with identity corner levels."""
    rng = np.random.default_rng(14)
    for W in (M.heisenberg_range_bulk([1.0, 0.4, 0.25]), M.exp_ising_bulk(1.0, 0.7, 0.5)):
        w, D = W.shape[0], 5
        GL, GR = rng.standard_normal((D, w, D)), rng.standard_normal((D, w, D))
        GL[:, 0, :] = np.eye(D)
        GR[:, -1, :] = np.eye(D)
        T = O.jordan_ac2_prepare(GL, W, W, GR)
        assert T["AA"] is not None
        x = rng.standard_normal((D, 2, D, 2))
        assert rel(O.jordan_ac2_apply(T, x), O.ac2_apply_generic(GL, W, W, GR, x)) < 1e-13


def test_jordan_nn_gemm_form_equals_jordan_apply():
    """The two-GEMM form of the nearest-neighbour operator (the timed CPU baseline)
    equals the term-by-term apply; it refuses operators with A blocks."""
    rng = np.random.default_rng(15)
    for W in (M.tfim_bulk(0.7), M.xxz_bulk(0.6)):
        w, Dl, Dr = W.shape[0], 6, 7
        GL, GR = rng.standard_normal((Dl, w, Dl)), rng.standard_normal((Dr, w, Dr))
        T = O.jordan_ac2_prepare(GL, W, W, GR)
        ICm, BEm, CB = O.jordan_ac2_nn_matrices(T, Dl, 2, Dr, 2)
        x = rng.standard_normal((Dl, 2, Dr, 2))
        assert rel(O.jordan_ac2_apply_nn(ICm, BEm, CB, x), O.jordan_ac2_apply(T, x)) < 1e-13
    Tj = O.jordan_ac2_prepare(GL[:, :1].repeat(8, 1), M.j1j2_bulk(), M.j1j2_bulk(), GR[:, :1].repeat(8, 1))
    with pytest.raises(AssertionError):
        O.jordan_ac2_nn_matrices(Tj, Dl, 2, Dr, 2)


def test_mps_transfers_vs_kronecker_matrices():
    """transfer_left/right for MPS tensors (transfer.jl:18-55) against the transfer
    matrix built from Kronecker products E = sum_h kron(Ab_h, A_h) acting on vec(v),
    with bra != ket and rectangular bonds; a right isometry is a right fixed point."""
    rng = np.random.default_rng(16)
    Dm, Dn, Da, Dc = 4, 5, 3, 6
    A, Ab = rng.standard_normal((Dn, 2, Dc)), rng.standard_normal((Dm, 2, Da))
    # left: y[a,c] = sum v[m,n] A[n,h,c] Ab[m,h,a]  ->  y = sum_h Ab_h^T v A_h
    v = rng.standard_normal((Dm, Dn))
    y = sum(Ab[:, h, :].T @ v @ A[:, h, :] for h in range(2))
    assert rel(O.transfer_left_mps(v, A, Ab), y) < 1e-13
    E = sum(np.kron(A[:, h, :].T, Ab[:, h, :].T) for h in range(2))      # acts on vec_F(v)
    assert rel(E @ v.reshape(-1, order="F"), y.reshape(-1, order="F")) < 1e-13
    # right: y[a,b] = sum A[a,h,i] Ab[b,h,j] v[i,j]  ->  y = sum_h A_h v Ab_h^T
    u = rng.standard_normal((Dc, Da))
    yr = sum(A[:, h, :] @ u @ Ab[:, h, :].T for h in range(2))
    assert rel(O.transfer_right_mps(u, A, Ab), yr) < 1e-13
    Er = sum(np.kron(Ab[:, h, :], A[:, h, :]) for h in range(2))
    assert rel(Er @ u.reshape(-1, order="F"), yr.reshape(-1, order="F")) < 1e-13
    # duality <v, T_R(u)> = <T_L(v), u> (the two are each other's adjoint)
    lhs = np.sum(v.T * O.transfer_right_mps(u, A, Ab))
    rhs = np.sum(O.transfer_left_mps(v, A, Ab) * u.T)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # a right isometry (sum_h A_h A_h^T = 1) is a fixed point of transfer_right
    D = 5
    Q, _ = np.linalg.qr(rng.standard_normal((D * 2, D)))
    Ar = Q.T.reshape(D, 2, D, order="F")
    assert rel(O.transfer_right_mps(np.eye(D), Ar, Ar), np.eye(D)) < 1e-13
    assert rel(O.transfer_right_mps(np.eye(D), Ar.transpose(2, 1, 0), Ar.transpose(2, 1, 0)), np.eye(D)) > 1e-3


@pytest.mark.parametrize("name", list(CHAINS))
def test_jordan_prepare_does_not_change_the_operator(name):
    """prepare_operator!! only regroups terms (hamiltonian_derivatives.jl:109-157,
    :371-468): the constructed operator, whose apply runs every term on its own
    (II, IC, ID, CB, CA, AB, AA, BE, DE, EE and D, I, E, C, B, A), equals the
    folded one, which the projected-H pins above check.  Every bond of the chain,
    boundaries included, so each unfolded term's apply is exercised."""
    rng = np.random.default_rng(17)
    Ws = CHAINS[name][0]()
    psi = B.random_mps(L, 2, 6, rng)
    (AL, AR, _, _), GL, GR = envs(Ws, psi, psi)
    seen = set()
    for j in range(L - 1):
        T0 = O.jordan_ac2_construct(GL[j], Ws[j], Ws[j + 1], GR[j + 2])
        seen |= {k for k, v in T0.items() if v is not None}
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 2].shape[0], 2))
        assert rel(O.jordan_ac2_apply(T0, x), O.jordan_ac2_apply(O.jordan_ac2_fold(T0), x)) < 1e-13
    for j in range(L):
        T0 = O.jordan_ac_construct(GL[j], Ws[j], GR[j + 1])
        x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 1].shape[0]))
        assert rel(O.jordan_ac_apply(T0, x), O.jordan_ac_apply(O.jordan_ac_fold(T0), x)) < 1e-13
    assert {"II", "IC", "CB", "BE", "EE"} <= seen
    if name in ("j1j2", "heis3"):
        assert {"CA", "AB"} <= seen
    if name in ("heis3", "expising"):
        assert "AA" in seen
    if name == "tfim":
        assert {"DE", "ID"} <= seen


def test_jordan_uses_identity_corner_levels():
    """Jordan applies ignore GL's "not started" and GR's "finished" levels (taken
    as identities), so they equal the generic apply with those levels set to 1."""
    rng = np.random.default_rng(3)
    W = M.xxz_bulk(0.8)
    D = 6
    GL, GR = rng.standard_normal((D, 5, D)), rng.standard_normal((D, 5, D))
    x = rng.standard_normal((D, 2, D, 2))
    GLi, GRi = GL.copy(), GR.copy()
    GLi[:, 0, :] = np.eye(D)
    GRi[:, -1, :] = np.eye(D)
    yj = O.jordan_ac2_apply(O.jordan_ac2_prepare(GL, W, W, GR), x)
    assert rel(yj, O.ac2_apply_generic(GLi, W, W, GRi, x)) < 1e-13
    xa = rng.standard_normal((D, 2, D))
    ya = O.jordan_ac_apply(O.jordan_ac_prepare(GL, W, GR), xa)
    assert rel(ya, O.ac_apply_generic(GLi, W, GRi, xa)) < 1e-13


def test_jordan_folding_covers_all_terms_j1j2():
    """J1-J2 has an A block: CA/AB folding paths (hamiltonian_derivatives.jl:379-462)."""
    rng = np.random.default_rng(4)
    W = M.j1j2_bulk(1.0, 0.4)
    D = 5
    GL, GR = rng.standard_normal((D, 8, D)), rng.standard_normal((D, 8, D))
    GL[:, 0, :] = np.eye(D)
    GR[:, -1, :] = np.eye(D)
    T = O.jordan_ac2_prepare(GL, W, W, GR)
    assert T["CA"] is not None and T["AB"] is not None and T["IC"] is None and T["BE"] is None
    x = rng.standard_normal((D, 2, D, 2))
    assert rel(O.jordan_ac2_apply(T, x), O.ac2_apply_generic(GL, W, W, GR, x)) < 1e-13


def test_planted_transposition_is_caught():
    """Swapping the bra/ket legs of GL (a plausible layout slip) fails the pin."""
    rng = np.random.default_rng(5)
    Ws, H = M.xxz(L, 0.6), B.dense_xxz(L, 0.6).toarray()
    ket, bra = B.random_mps(L, 2, 6, rng), B.random_mps(L, 2, 6, rng)
    (ALk, ARk, ALb, ARb), GL, GR = envs(Ws, ket, bra)
    j = 3
    x = rng.standard_normal((GL[j].shape[2], 2, GR[j + 2].shape[0], 2))
    yb = B.isometry_ac2(ALb, ARb, j, 2).T @ H @ B.isometry_ac2(ALk, ARk, j, 2) @ x.reshape(-1, order="F")
    bad = O.ac2_apply_generic(np.transpose(GL[j], (2, 1, 0)), Ws[j], Ws[j + 1], GR[j + 2], x)
    assert rel(bad.reshape(-1, order="F"), yb) > 1e-2


def test_transfer_fixed_point_and_regularize():
    """l_LL * T_LL = l_LL for a left-canonical state (test/states/infinitemps.jl:38-46)
    and regularize! removes the projection (transfermatrix.jl:70-80)."""
    rng = np.random.default_rng(6)
    D = 5
    Q, _ = np.linalg.qr(rng.standard_normal((D * 2, D)))
    A = Q.reshape(D, 2, D, order="F")                 # left isometry
    I = np.eye(D)
    assert rel(O.transfer_left_mps(I, A, A), I) < 1e-13
    v = rng.standard_normal((D, D))
    lvec, rvec = rng.standard_normal((D, D)), rng.standard_normal((D, D))
    lvec /= np.einsum("ij,ji->", lvec, rvec)
    r = O.regularize(v, lvec, rvec)
    assert abs(np.einsum("ij,ji->", lvec, r)) < 1e-12
    v3 = rng.standard_normal((D, 3, D))
    r3 = O.regularize(v3, lvec, rvec)
    assert np.abs(np.einsum("ij,jgi->g", lvec, r3)).max() < 1e-12


def test_tfim_closed_form_energy():
    for Lc in (1, 2, 5, 9):
        for g in (0.0, 0.5, 1.0, 1.7):
            e = np.linalg.eigvalsh(B.dense_tfim(Lc, g).toarray())[0]
            assert abs(e - B.tfim_open_energy(Lc, g)) < 1e-10


def test_majumdar_ghosh_dimer_energy():
    """J2/J1 = 1/2 (Majumdar-Ghosh) open chain of even length L: the nearest-neighbour
    dimer product state is a ground state with energy -3/4 per singlet, -3L/8."""
    e = np.linalg.eigvalsh(B.dense_j1j2(8, 1.0, 0.5).toarray())[0]
    assert abs(e - (-0.375 * 8)) < 1e-10


def test_fixtures_reproduced_by_oracle():
    data = np.load(__import__("pathlib").Path(__file__).with_name("golden") / "apply_vectors.npz")
    cases = sorted({k.split("/")[0] for k in data.files})
    assert len(cases) > 30
    for c in cases:
        g = {k.split("/")[1]: data[k] for k in data.files if k.startswith(c + "/")}
        kind = c.split("_")[0]
        if kind == "ac2gen":
            y = O.ac2_apply_generic(g["GL"], g["W1"], g["W2"], g["GR"], g["x"])
        elif kind == "acgen":
            y = O.ac_apply_generic(g["GL"], g["W"], g["GR"], g["x"])
        elif kind == "c":
            y = O.c_apply(g["GL"], g["GR"], g["x"])
        elif kind in ("ac2jor", "ac2chain"):
            y = O.jordan_ac2_apply(O.jordan_ac2_prepare(g["GL"], g["W1"], g["W2"], g["GR"]), g["x"])
        elif kind in ("acjor", "acchain"):
            y = O.jordan_ac_apply(O.jordan_ac_prepare(g["GL"], g["W"], g["GR"]), g["x"])
        elif kind == "tlmps":
            y = O.transfer_left_mps(g["v"], g["A"], g["Ab"])
        elif kind == "trmps":
            y = O.transfer_right_mps(g["v"], g["A"], g["Ab"])
        elif kind == "tlmpo":
            y = O.transfer_left_mpo(g["x"], g["O"], g["A"], g["Ab"])
        elif kind == "trmpo":
            y = O.transfer_right_mpo(g["v"], g["O"], g["A"], g["Ab"])
        else:
            raise AssertionError(c)
        assert rel(y, g["y"]) < 1e-13, c


def test_tfim_infinite_closed_form_vs_open_chains():
    """e0(g) from the dispersion integral equals the L -> infinity limit of the open
    chain closed form (boundary energy cancels in E(2L) - E(L)); g = 1 gives -4/pi."""
    assert abs(B.tfim_infinite_energy(1.0) - B.TFIM_INFINITE_G1) < 1e-12
    for g in (0.3, 0.5, 1.5):
        bulk = (B.tfim_open_energy(800, g) - B.tfim_open_energy(400, g)) / 400
        assert abs(bulk - B.tfim_infinite_energy(g)) < 1e-8, g


def test_onsager_closed_forms_agree_and_pin_the_critical_value():
    """The single- and double-integral forms of Onsager's ln Z per site agree; at
    beta_c = ln(1 + sqrt 2) / 2 the value is ln sqrt 2 + 2 G / pi, which is the
    2.5337 per site the reference's leading_boundary test checks
    (test/algorithms/groundstate.jl:481); beta -> 0 gives ln 2."""
    for beta in (0.05, 0.3, 0.4, 0.5, 0.9):
        assert abs(B.ising2d_log_partition(beta) - B.ising2d_log_partition_double(beta)) < 1e-12
    bc = np.log(1 + np.sqrt(2)) / 2
    crit = B.ising2d_log_partition(bc)
    assert abs(crit - (np.log(np.sqrt(2)) + 2 * B.CATALAN / np.pi)) < 1e-12
    assert abs(np.exp(crit) - 2.5337) < 1e-3
    assert abs(B.ising2d_log_partition(1e-9) - np.log(2)) < 1e-12


def test_tfim_infinite_sx_is_the_energy_derivative():
    """<sigma^x> = -d e0 / dg: the closed form against a central difference of
    Pfeuty's energy, and the limits g -> infinity (1) and g = 0 (0)."""
    for g in (0.3, 0.8, 1.5):
        h = 1e-4
        fd = -(B.tfim_infinite_energy(g + h) - B.tfim_infinite_energy(g - h)) / (2 * h)
        assert abs(B.tfim_infinite_sx(g) - fd) < 1e-8
    assert abs(B.tfim_infinite_sx(1e4) - 1.0) < 1e-7
    assert abs(B.tfim_infinite_sx(0.0)) < 1e-12
