"""exact_diagonalization on the HIP path: the lowest eigenvalues of an open XXZ
chain (with its degenerate pairs) against numpy's eigvalsh of the
Kronecker-built Hamiltonian, and the states as dense vectors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402


def test_exact_diagonalization_xxz_lowest_levels():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd as m
    m.load()
    from mpskit_amd.ed import exact_diagonalization
    L, delta = 10, 0.7
    H = m.FiniteMPOHamiltonian(m.models.xxz(L, delta))
    print("exact diagonalization xxz L=10", flush=True)
    vals, states = exact_diagonalization(H, num=4, tol=1e-11)
    Hd = B.dense_xxz(L, delta).toarray()
    ev = np.linalg.eigvalsh(Hd)
    assert np.allclose(sorted(vals), ev[:4], atol=1e-9), (vals, ev[:4])
    for lam, st in zip(vals, states):
        v = st.to_dense().cpu().numpy()
        v = v / np.linalg.norm(v)
        assert np.linalg.norm(Hd @ v - lam * v) < 1e-8
        assert st.bond_dims() == [min(2 ** k, 2 ** (L - k)) for k in range(L + 1)]


def test_exact_diagonalization_gapped_ground_state_at_zero():
    """Ground energy near 0 and a gap of about 3 (3 n_i plus a weak field and a
    weak n_i n_{i+1} coupling): a found state shifted by 2 (|E0| + 1) would sit below the
    first excited level, so further states must come from the complement of the
    found ones, not from a shift."""
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd as m
    m.load()
    from mpskit_amd.ed import exact_diagonalization
    L, J, h = 8, 0.4, 0.3
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    n = np.diag([1.0, 0.0])
    onsite = 3.0 * n + h * X
    W = m.models._jordan(2, [n], [J * n], onsite)
    H = m.FiniteMPOHamiltonian(m.models.finite_chain(W, L))
    Hd = np.zeros((2 ** L, 2 ** L))
    for i in range(L):
        Hd += np.kron(np.kron(np.eye(2 ** i), onsite), np.eye(2 ** (L - i - 1)))
    for i in range(L - 1):
        Hd += J * np.kron(np.kron(np.eye(2 ** i), np.kron(n, n)), np.eye(2 ** (L - i - 2)))
    ev = np.linalg.eigvalsh(Hd)
    # a shift by 2 (|E0| + 1) leaves the ground state below E1
    assert abs(ev[0]) < 0.5 and ev[0] + 2.0 * (abs(ev[0]) + 1.0) < ev[1] - 0.3
    print("exact diagonalization gapped chain L=8", flush=True)
    vals, states = exact_diagonalization(H, num=4, tol=1e-11)
    assert np.allclose(vals, ev[:4], atol=1e-9), (vals, ev[:4])
    vecs = []
    for lam, st in zip(vals, states):
        v = st.to_dense().cpu().numpy().ravel()
        v = v / np.linalg.norm(v)
        assert np.linalg.norm(Hd @ v - lam * v) < 1e-8
        vecs.append(v)
    Vm = np.array(vecs)
    assert np.allclose(Vm @ Vm.T, np.eye(4), atol=1e-8)
