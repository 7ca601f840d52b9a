"""Host-side logic without a GPU: adaptive Krylov budgets, model MPOs."""
import math

import numpy as np
import pytest

from mpskit_amd.algorithms import AdaptiveKrylov
from mpskit_amd import models as M


def test_adaptive_krylov_cold_start_and_stall():
    ak = AdaptiveKrylov()
    tol, kd, mi = ak.adapt_solver(decay_rate=0.0, g_global=1.0)
    assert (kd, mi) == (3, 1) and tol == 1e-2          # cold start: minimal budget, tol clamped to tol_max
    tol, kd, mi = ak.adapt_solver(decay_rate=1.0, g_global=1e-5)
    assert (kd, mi) == (20, 1) and tol == pytest.approx(1e-6)   # stalled bond: maximal budget


def test_adaptive_krylov_predicted_budget():
    """nmatvecs = round(log(g_local / tol) / log(1 / rho)) (dynamictols.jl:224-233)."""
    ak = AdaptiveKrylov()
    tol, kd, mi = ak.adapt_solver(decay_rate=0.5, g_local=1e-2, g_global=1e-6, eps_trunc=1e-4)
    assert tol == pytest.approx(1e-5)                  # max(0.1 * 1e-4, 0.1 * 1e-6)
    assert kd == round(math.log(1e-2 / 1e-5) / math.log(2.0))
    tol, kd, mi = ak.adapt_solver(decay_rate=0.9, g_local=1.0, g_global=1e-12)
    assert tol == 1e-12 and kd == 20


def test_finite_chain_boundaries():
    Ws = M.xxz(6, 0.5)
    assert Ws[0].shape == (1, 2, 2, 5) and Ws[-1].shape == (5, 2, 2, 1)
    assert all(W.shape == (5, 2, 2, 5) for W in Ws[1:-1])
    Wj = M.j1j2_bulk(1.0, 0.24)
    assert Wj.shape == (8, 2, 2, 8)
    # pass-through identities of the next-nearest-neighbour channels in the A block,
    # upper triangular in channel order
    for k in range(3):
        assert np.array_equal(Wj[1 + k, :, :, 4 + k], np.eye(2))
    for g in range(8):
        for k in range(g):
            assert not np.any(Wj[g, :, :, k]), (g, k)


def test_truncated_split_batched_matches_single_and_svd():
    """Batched Gram-eigh splits (size groups, -1 padding) against the single
    split and against the SVD: same kept singular values and discarded weight,
    orthonormal isometries, and A B equal to the SVD's rank-k truncation."""
    import torch
    from mpskit_amd.algorithms import truncated_split, truncated_split_batched
    g = torch.Generator().manual_seed(7)
    shapes = [(40, 36), (64, 64), (24, 50), (61, 30), (64, 20)]
    Ms = []
    for m, n in shapes:
        U, _ = torch.linalg.qr(torch.randn(m, m, generator=g, dtype=torch.float64))
        V, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
        r = min(m, n)
        s = torch.logspace(0, -6, r, dtype=torch.float64)   # distinct, decaying singular values
        Ms.append(U[:, :r] @ torch.diag(s) @ V[:, :r].t())
    for direction in ("right", "left"):
        for trunc_dim, trunc_err in ((16, 0.0), (64, 1e-4)):
            got = truncated_split_batched(Ms, direction, trunc_dim, trunc_err)
            for M, (A, B, S, eps) in zip(Ms, got):
                A1, B1, S1, e1 = truncated_split(M, direction, trunc_dim, trunc_err, method="svd")
                assert S.shape == S1.shape
                assert torch.allclose(S, S1, rtol=1e-6, atol=1e-12)
                # discarded weight from Gram eigenvalues: absolute error ~ eps_mach * s_max^2 each
                assert abs(eps ** 2 - e1 ** 2) <= 600 * 2.3e-16 * float(S1[0]) ** 2
                iso = A if direction == "right" else B.t()
                assert torch.allclose(iso.t() @ iso, torch.eye(iso.shape[1], dtype=torch.float64), atol=1e-12)
                assert torch.allclose(A @ B, A1 @ B1, atol=1e-10)
                S2 = truncated_split(M, direction, trunc_dim, trunc_err, method="eigh")[2]
                assert torch.allclose(S, S2, atol=1e-12)


@pytest.mark.parametrize("L", [3, 4])
def test_classical_ising_mpo_contracts_to_the_torus_partition_function(L):
    """Tr(T_row^L) of the L x L torus built from the MPO tensor equals the
    brute-force sum over all 2^(L*L) spin configurations."""
    from oracle import brute as B
    beta = 0.37
    O = M.classical_ising(beta)
    if L == 3:
        Trow = np.einsum("pbhq,qcir,rdjp->bcdhij", O, O, O).reshape(8, 8)
    else:
        Trow = np.einsum("pbhq,qcir,rdjs,setp->bcdehijt", O, O, O, O).reshape(16, 16)
    Z = np.trace(np.linalg.matrix_power(Trow, L))
    ref = B.ising2d_torus_partition(beta, L)
    assert abs(Z - ref) < 1e-11 * ref


@pytest.mark.parametrize("cond", [1.0, 1e4, 1e8, 1e12])
def test_pos_qr_matches_householder(cond):
    """Shifted Cholesky QR3 gives the unique positive-diagonal QR: the same
    factors as Householder QR with signs fixed, for condition numbers up to
    1e12, with Q orthonormal to rounding."""
    import torch
    from mpskit_amd.states import _householder_pos_qr, pos_qr
    rng = np.random.default_rng(5)
    m, n = 96, 48
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = torch.from_numpy(U @ np.diag(np.logspace(0, -np.log10(cond), n)) @ V.T)
    Q, R = pos_qr(A)
    Qh, Rh = _householder_pos_qr(A)
    assert torch.all(torch.diagonal(R) > 0)
    assert float(torch.linalg.matrix_norm(Q.t() @ Q - torch.eye(n, dtype=torch.float64))) < 1e-13
    assert float(torch.linalg.matrix_norm(Q @ R - A)) < 1e-14 * float(torch.linalg.matrix_norm(A))
    assert float(torch.linalg.matrix_norm(R - Rh)) < 1e-13 * cond ** 0.5
    if cond <= 1e4:
        assert float(torch.linalg.matrix_norm(Q - Qh)) < 1e-10


def test_pos_qr_rank_deficient_and_wide_fall_back():
    import torch
    from mpskit_amd.states import pos_qr
    rng = np.random.default_rng(6)
    A = torch.from_numpy(rng.standard_normal((40, 10)) @ rng.standard_normal((10, 20)))   # rank 10
    Q, R = pos_qr(A)
    assert float(torch.linalg.matrix_norm(Q @ R - A)) < 1e-12 * float(torch.linalg.matrix_norm(A))
    assert float(torch.linalg.matrix_norm(Q.t() @ Q - torch.eye(20, dtype=torch.float64))) < 1e-12
    W = torch.from_numpy(rng.standard_normal((5, 9)))
    Q, R = pos_qr(W)
    assert Q.shape == (5, 5) and float(torch.linalg.matrix_norm(Q @ R - W)) < 1e-13


@pytest.mark.parametrize("direction", ["right", "left"])
@pytest.mark.parametrize("shape", [(24, 40), (40, 24), (32, 32)])
def test_truncated_split_without_truncation_is_exact(direction, shape):
    """Splits that keep every singular direction (trunc_err = 0, trunc_dim >=
    min(m, n)) return an exact factorisation M = A B with the isometry on the
    requested side, also for rank-deficient M."""
    import torch
    from mpskit_amd.algorithms import truncated_split
    g = torch.Generator().manual_seed(3)
    m, n = shape
    for rank in (min(m, n), min(m, n) // 2):
        M = torch.randn(m, rank, generator=g, dtype=torch.float64) @ torch.randn(rank, n, generator=g,
                                                                                 dtype=torch.float64)
        A, B, S, eps = truncated_split(M, direction, 64, 0.0)
        assert eps == 0.0 and A.shape[1] == min(m, n) == B.shape[0]
        assert float(torch.linalg.matrix_norm(A @ B - M)) < 1e-12 * float(torch.linalg.matrix_norm(M))
        iso = A if direction == "right" else B.t()
        k = iso.shape[1]
        assert float(torch.linalg.matrix_norm(iso.t() @ iso - torch.eye(k, dtype=torch.float64))) < 1e-12


def test_rationalize_is_the_smallest_denominator_within_tolerance():
    """rationalize(x, tol) against a brute-force search over denominators, and
    the values Julia's rationalize gives for pi (22//7 at tol 0.01)."""
    from fractions import Fraction
    from mpskit_amd.observables import rationalize
    assert rationalize(np.pi, 0.01) == Fraction(22, 7)
    assert rationalize(-0.97, 0.1) == -1 and rationalize(0.3 / np.pi, 0.1) == 0
    rng = np.random.default_rng(5)
    for x, tol in zip(rng.uniform(-1, 1, 200), 10.0 ** rng.uniform(-4, -0.5, 200)):
        r = rationalize(x, tol)
        assert abs(Fraction(float(x)) - r) <= Fraction(float(tol))
        for q in range(1, r.denominator):
            p = round(x * q)
            assert all(abs(Fraction(float(x)) - Fraction(pp, q)) > Fraction(float(tol)) for pp in (p - 1, p, p + 1))


def test_marek_gap_and_correlation_length_of_a_synthetic_spectrum():
    """marek_gap (transfer_spectrum.jl:136-160) drops the modulus-1 eigenvalue,
    snaps the angles to the simplest fraction of pi within tol_angle and groups
    by the snapped angle.  Angles 0.3, -0.3 and 0.31 (0.096, -0.096 and 0.099 in
    units of pi) all snap to 0 at tol 0.1, so theta = 0 and the first two
    -log|lambda| coincide (delta = 0); -2 snaps to -2/3 pi."""
    from mpskit_amd.observables import approx_angles, marek_gap
    from fractions import Fraction
    lam2 = 0.6 * np.exp(0.3j)
    spec = np.array([1.0, lam2, np.conj(lam2), 0.5 * np.exp(0.31j), 0.55 * np.exp(-2.0j)])
    assert approx_angles(spec[1:], 0.1) == [0, 0, 0, Fraction(-2, 3)]
    eps, delta, theta = marek_gap(spec, tol_angle=0.1)
    assert abs(eps + np.log(0.6)) < 1e-14
    assert theta == 0.0 and delta == 0.0
    # finer tolerance: 0.3 / pi = 0.0955 snaps to 1/10 with tol 0.01; only
    # lambda_2 and 0.5 e^{0.31 i} (0.0987) share it: two eigenvalues, delta = inf
    eps, delta, theta = marek_gap(spec, tol_angle=0.01)
    assert abs(theta - np.pi / 10) < 1e-15 and delta == np.inf and abs(eps + np.log(0.6)) < 1e-14
    # branch cut: angles near -pi snap to -1, mapped to +1
    spec = np.array([1.0, 0.7 * np.exp(-0.97j * np.pi), 0.6 * np.exp(0.98j * np.pi), 0.4 * np.exp(0.99j * np.pi)])
    eps, delta, theta = marek_gap(spec, tol_angle=0.1)
    assert theta == np.pi and abs(eps + np.log(0.7)) < 1e-14
    assert abs(delta - (np.log(0.7) - np.log(0.6))) < 1e-14


def test_mpo_product_is_the_matrix_square():
    """The product MPO used by variance contracts to H @ H (dense Kronecker TFIM)."""
    from oracle import brute as B
    from mpskit_amd.observables import _mpo_product
    Ws = M.tfim(4, 1.3)
    H = B.dense_tfim(4, 1.3).toarray()
    P = [_mpo_product(W, W) for W in Ws]
    assert np.abs(B.mpo_to_dense(P) - H @ H).max() < 1e-12


def test_svdcut_keep_rule():
    import torch
    from mpskit_amd.changebonds import SvdCut, _keep
    S = torch.tensor([1.0, 0.5, 1e-3, 1e-6, 1e-9], dtype=torch.float64)
    assert _keep(S, SvdCut(trunc_dim=3)) == 3
    assert _keep(S, SvdCut(trunc_dim=10, trunc_err=1e-5)) == 3     # tail 2-norm of [1e-6, 1e-9] < 1e-5
    assert _keep(S, SvdCut(trunc_dim=10, trunc_err=0.0)) == 5


def test_infinite_mpo_rejects_inconsistent_bonds():
    """InfiniteMPO checks its unit cell before touching the device."""
    from mpskit_amd.boundary import InfiniteMPO
    with pytest.raises(ValueError):
        InfiniteMPO([np.zeros((2, 2, 2, 3))])            # right bond 3 != left bond 2 of the next site
    with pytest.raises(ValueError):
        InfiniteMPO([np.zeros((2, 2, 3, 2))])            # physical legs 2 and 3


def test_marek_gap_needs_an_eigenvalue_below_one():
    from mpskit_amd.observables import marek_gap
    with pytest.raises(ValueError):
        marek_gap(np.array([1.0, -1.0 + 1e-14]))


def test_argument_checks_before_any_device_work():
    """correlator, approximate and dot reject bad arguments before touching a state."""
    from types import SimpleNamespace
    from mpskit_amd import approximate as Ap, observables as Ob
    with pytest.raises(ValueError):
        Ob.correlator(None, np.eye(2), np.eye(2), 3, [2, 5])        # a j not to the right of i
    with pytest.raises(ValueError):
        Ob.correlator(None, np.eye(2), np.eye(2), 3, [])
    with pytest.raises(ValueError):
        Ap.approximate(SimpleNamespace(L=6), None, SimpleNamespace(L=5), Ap.ApproximateDMRG2())
    with pytest.raises(ValueError):
        Ob.dot(SimpleNamespace(L=4), SimpleNamespace(L=5))


def _cgs2_gemv(Vb, k, w):
    """The two Gram-Schmidt passes of mpskit_amd.krylov.cgs2 as plain torch
    GEMVs (the library kernels behind cgs2 need the GPU; GPU tests check them)."""
    import torch
    Vj = Vb[:k]
    h1 = Vj @ w
    w -= Vj.t() @ h1
    h2 = Vj @ w
    w -= Vj.t() @ h2
    return torch.cat([h1 + h2, torch.linalg.vector_norm(w)[None]]).numpy()


@pytest.fixture
def cpu_layout(monkeypatch):
    """The Krylov solvers' vector layout helper keeps CPU tensors on the CPU
    here, and their orthogonalisation step runs as torch GEMVs."""
    from mpskit_amd import excited, observables, tdvp, vumps
    from mpskit_amd import tensors as T
    monkeypatch.setattr(T, "as_colmajor", lambda t: t)
    for mod in (vumps, excited, tdvp, observables):
        monkeypatch.setattr(mod, "cgs2", _cgs2_gemv)
    return T


def test_krylov_solvers_against_dense_linear_algebra(cpu_layout):
    """arnoldi_dominant, gmres, exponentiate and lanczos_min on dense matrices
    (1-D vectors, where the column-major layout is trivial) against numpy/scipy."""
    import torch
    from scipy.linalg import expm
    from mpskit_amd.excited import lanczos_min
    from mpskit_amd.tdvp import exponentiate
    from mpskit_amd.vumps import arnoldi_dominant, gmres
    rng = np.random.default_rng(21)
    n = 120
    P = torch.from_numpy(rng.random((n, n)))                       # positive: real Perron eigenvalue
    lam, v = arnoldi_dominant(lambda x: P @ x, torch.from_numpy(rng.random(n)), tol=1e-12, relative=True)
    ev = np.linalg.eigvals(P.numpy())
    assert abs(lam - ev[np.argmax(np.abs(ev))].real) < 1e-9 * abs(lam)
    assert float(torch.linalg.vector_norm(P @ v - lam * v)) < 1e-9 * abs(lam)
    A = torch.from_numpy(np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.2)
    b = torch.from_numpy(rng.standard_normal(n))
    x, info = gmres(lambda y: A @ y, b, torch.zeros(n, dtype=torch.float64), 1e-12)
    assert info["converged"] and float(torch.linalg.vector_norm(A @ x - b)) < 1e-11
    S = rng.standard_normal((n, n))
    S = torch.from_numpy((S + S.T) / np.sqrt(n))
    x0 = rng.standard_normal(n)
    y, _ = exponentiate(lambda z: S @ z, -0.7, torch.from_numpy(x0), tol=1e-13)
    ref = expm(-0.7 * S.numpy()) @ x0
    assert np.linalg.norm(y.numpy() - ref) < 1e-11 * np.linalg.norm(ref)
    lam0, u, _ = lanczos_min(lambda z: S @ z, torch.from_numpy(rng.standard_normal(n)), tol=1e-11)
    assert abs(lam0 - np.linalg.eigvalsh(S.numpy())[0]) < 1e-9


def test_complement_columns_span_the_orthogonal_complement():
    """The null-space helper of the finite expansions: orthonormal columns,
    orthogonal to range(M), of dimension m - rank(M) (rank-deficient M too)."""
    import torch
    from mpskit_amd.changebonds import _complement_cols
    g = torch.Generator().manual_seed(2)
    for m, n, r in ((12, 5, 5), (12, 5, 3), (8, 8, 8)):
        M = torch.randn(m, r, generator=g, dtype=torch.float64) @ torch.randn(r, n, generator=g,
                                                                          dtype=torch.float64)
        N = _complement_cols(M)
        assert N.shape == (m, m - r)
        assert float(torch.linalg.matrix_norm(N.t() @ N - torch.eye(m - r, dtype=torch.float64))) < 1e-12
        assert float(torch.linalg.matrix_norm(N.t() @ M)) < 1e-12 * float(torch.linalg.matrix_norm(M))
