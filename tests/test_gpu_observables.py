"""Local expectation values, correlators and entanglement on the HIP path,
against the dense ground state of the Kronecker-built TFIM (exact
diagonalisation) and, for an infinite chain, the closed-form <sigma^x>."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402


@pytest.fixture(scope="module")
def ground(mk_mod):
    m, Ob = mk_mod
    L, g = 10, 1.5
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, g))
    psi = m.FiniteMPS.random(L, 2, 32, seed=0)
    print("observables: ground state L=10", flush=True)
    psi, _, _, _ = m.find_groundstate(psi, H, m.DMRG2(tol=1e-12, maxiter=10, trunc_dim=32, verbosity=0))
    ev, U = np.linalg.eigh(B.dense_tfim(L, g).toarray())
    return psi, U[:, 0], L


@pytest.fixture(scope="module")
def mk_mod():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    from mpskit_amd import observables
    return mpskit_amd, observables


def _dense(op, sites, L):
    M = B.sp.identity(1, format="csr")
    for j in range(L):
        M = B.sp.kron(M, B.sp.csr_matrix(op[sites.index(j)]) if j in sites else B.sp.identity(2), format="csr")
    return M


def test_local_and_two_site_expectation_values(mk_mod, ground):
    _, Ob = mk_mod
    psi, v, L = ground
    for site in (0, 4, 9):
        ref = v @ (B._site_op(B.PX, site, L) @ v)
        assert abs(Ob.expectation_value_local(psi, site, B.PX) - ref) < 1e-10
    ZZ = np.einsum("bh,ck->bchk", B.PZ, B.PZ)
    for site in (0, 5, 8):
        ref = v @ (B._site_op(B.PZ, site, L) @ (B._site_op(B.PZ, site + 1, L) @ v))
        assert abs(Ob.expectation_value_local(psi, site, ZZ) - ref) < 1e-10


def test_correlator_against_dense(mk_mod, ground):
    _, Ob = mk_mod
    psi, v, L = ground
    js = [2, 3, 5, 9]
    got = Ob.correlator(psi, B.PZ, B.PZ, 1, js)
    for j, val in zip(js, got):
        ref = v @ (B._site_op(B.PZ, 1, L) @ (B._site_op(B.PZ, j, L) @ v))
        assert abs(val - ref) < 1e-10, (j, val, ref)


def test_entanglement_spectrum_and_entropy_against_dense(mk_mod, ground):
    _, Ob = mk_mod
    psi, v, L = ground
    for cut in (0, 3, 5, 10):
        s_ref = np.linalg.svd(v.reshape(2 ** cut, 2 ** (L - cut)), compute_uv=False)
        s = Ob.entanglement_spectrum(psi, cut).cpu().numpy()
        k = len(s)
        assert np.allclose(np.sort(s)[::-1], s_ref[:k], atol=1e-10)
        s_ref = s_ref[s_ref >= np.finfo(float).eps]
        S_ref = -np.sum(s_ref ** 2 * np.log(s_ref ** 2))
        assert abs(Ob.entropy(psi, cut) - S_ref) < 1e-9


def test_infinite_tfim_magnetisation_closed_form(mk_mod):
    m, Ob = mk_mod
    from mpskit_amd import vumps as V
    g = 1.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)] * 2)
    psi = V.InfiniteMPS.random(2, 2, 16, seed=4)
    print("observables: infinite tfim", flush=True)
    psi, envs, _, _ = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-11, maxiter=60, verbosity=0))
    for site in (0, 1):
        assert abs(Ob.expectation_value_local(psi, site, B.PX) - B.tfim_infinite_sx(g)) < 1e-9
    # translation invariance of the converged two-site cell and a normalised spectrum
    s0, s1 = Ob.entanglement_spectrum(psi, 0), Ob.entanglement_spectrum(psi, 1)
    assert abs(float((s0 * s0).sum()) - 1.0) < 1e-12
    assert abs(Ob.entropy(psi, 0) - Ob.entropy(psi, 1)) < 1e-8


def test_variance_finite(mk_mod, ground):
    """variance = <H^2> - <H>^2: ~0 for the exact ground state, the dense value
    for a random state (product MPO through the HIP MPO transfers)."""
    m, Ob = mk_mod
    psi, v, L = ground
    H = m.FiniteMPOHamiltonian(m.models.tfim(L, 1.5))
    assert abs(Ob.variance(psi, H)) < 1e-9
    rnd = m.FiniteMPS.random(L, 2, 8, seed=9)
    w = rnd.to_dense().cpu().numpy()
    Hd = B.dense_tfim(L, 1.5).toarray()
    ref = (w @ Hd @ Hd @ w) / (w @ w) - ((w @ Hd @ w) / (w @ w)) ** 2
    assert abs(Ob.variance(rnd, H) - ref) < 1e-9 * max(1.0, abs(ref))


def test_dot_of_two_states(mk_mod, ground):
    m, Ob = mk_mod
    psi, v, L = ground
    other = m.FiniteMPS.random(L, 2, 8, seed=10)
    other.move_center(3)
    w = other.to_dense().cpu().numpy()
    vg = psi.to_dense().cpu().numpy()
    assert abs(Ob.dot(psi, other) - float(vg @ w)) < 1e-12
    assert abs(Ob.dot(other, other) - float(w @ w)) < 1e-12


def test_transfer_spectrum_and_correlation_length(mk_mod):
    """Infinite TFIM paramagnet (g = 1.5): the transfer matrix of the normalised
    MPS has the eigenvalue 1; its next eigenvalue sets the asymptotic decay of
    the connected <sigma^z sigma^z> computed with ``correlator``, and the
    correlation length approaches the exact 1 / ln g from below as D grows."""
    m, Ob = mk_mod
    from mpskit_amd import vumps as V
    g = 1.5
    H = V.InfiniteHamiltonian([m.models.tfim_bulk(g)])
    psi = V.InfiniteMPS.random(1, 2, 24, seed=5)
    print("transfer spectrum tfim g=1.5 D=24", flush=True)
    psi, envs, _, _ = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-11, maxiter=80, verbosity=0))
    ev, res = Ob.transfer_spectrum(psi, howmany=6)
    assert abs(ev[0] - 1.0) < 1e-10 and res[0] < 1e-8
    xi = Ob.correlation_length(psi, howmany=6)
    xi_exact = 1.0 / np.log(g)
    assert 0.8 * xi_exact < xi < xi_exact * (1 + 1e-6)
    # <sigma^z> = 0 in the paramagnet: the correlator is connected; its decay rate
    c = Ob.correlator(psi, B.PZ, B.PZ, 0, [30, 31])
    assert abs(c[1] / c[0] / np.exp(-1.0 / xi) - 1.0) < 1e-2
