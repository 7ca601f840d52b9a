"""approximate(psi, (O, phi), DMRG2) on the HIP path: the compressed state
against the dense O|phi> (oracle/brute.py mpo_to_dense, Kronecker order), with a
random non-symmetric MPO so that a transposed operator or swapped bra/ket legs
in the mixed environments would fail."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    from mpskit_amd import approximate
    return mpskit_amd, approximate


def _random_mpo(L, w, seed):
    rng = np.random.default_rng(seed)
    Ws = []
    for j in range(L):
        wl = 1 if j == 0 else w
        wr = 1 if j == L - 1 else w
        Ws.append(rng.standard_normal((wl, 2, 2, wr)) / np.sqrt(2 * w))
    return Ws


@pytest.mark.parametrize("kind", ["tfim", "random"])
def test_approximate_exact_bond_dimension(mk, kind):
    m, Ap = mk
    L = 8
    Ws = m.models.tfim(L, 1.3) if kind == "tfim" else _random_mpo(L, 3, 1)
    O = m.FiniteMPOHamiltonian(Ws)
    phi = m.FiniteMPS.random(L, 2, 4, seed=2)
    psi = m.FiniteMPS.random(L, 2, 16, seed=3)
    print("approximate", kind, flush=True)
    psi, eps, hist = Ap.approximate(psi, O, phi, Ap.ApproximateDMRG2(trunc_dim=16, tol=1e-12, maxiter=6))
    target = B.mpo_to_dense(Ws) @ phi.to_dense().cpu().numpy()
    got = psi.to_dense().cpu().numpy()
    assert np.linalg.norm(got - target) < 1e-10 * np.linalg.norm(target), (kind, hist[-1])


def test_approximate_truncated_is_close_and_improves_with_D(mk):
    m, Ap = mk
    L = 10
    Ws = _random_mpo(L, 3, 4)
    O = m.FiniteMPOHamiltonian(Ws)
    phi = m.FiniteMPS.random(L, 2, 8, seed=5)
    target = B.mpo_to_dense(Ws) @ phi.to_dense().cpu().numpy()
    errs = []
    for D in (4, 8, 24):
        psi = m.FiniteMPS.random(L, 2, D, seed=6)
        psi, _, _ = Ap.approximate(psi, O, phi, Ap.ApproximateDMRG2(trunc_dim=D, tol=1e-12, maxiter=8))
        errs.append(np.linalg.norm(psi.to_dense().cpu().numpy() - target) / np.linalg.norm(target))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-10, errs
