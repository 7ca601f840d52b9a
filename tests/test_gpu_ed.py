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
