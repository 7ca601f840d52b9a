"""The Krylov orthogonalisation step shared by the host solvers
(``mk_krylov_orthogonalize`` through ``mpskit_amd.krylov.cgs2``) against the
same two classical Gram-Schmidt passes as torch GEMVs: coefficients and the
orthogonalised vector to 1e-13 relative, orthogonality to the basis, basis
sizes 1 to 32 (the library) and 40 (GEMVs), vector lengths that are not
multiples of the kernels' blocks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


@pytest.mark.parametrize("n", [1, 777, 100003, 2 ** 21 + 5])
def test_cgs2_matches_two_gemv_passes(mk, n):
    import torch
    from mpskit_amd.krylov import cgs2
    g = torch.Generator(device="cuda").manual_seed(n)
    for k in (1, 2, 5, 16, 31, 32, 40):
        if k > n:
            continue
        Q, _ = torch.linalg.qr(torch.randn(n, k, dtype=torch.float64, device="cuda", generator=g))
        Vb = torch.zeros(k + 1, n, dtype=torch.float64, device="cuda")
        Vb[:k] = Q.t()
        w = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
        ref = w.clone()
        h1 = Vb[:k] @ ref
        ref -= Vb[:k].t() @ h1
        h2 = Vb[:k] @ ref
        ref -= Vb[:k].t() @ h2
        hr = torch.cat([h1 + h2, torch.linalg.vector_norm(ref)[None]]).cpu().numpy()
        h = cgs2(Vb, k, w)
        scale = np.abs(hr).max()
        assert np.abs(h - hr).max() <= 1e-13 * scale, (n, k)
        assert float(torch.linalg.vector_norm(w - ref)) <= 1e-13 * float(torch.linalg.vector_norm(ref) + 1e-300) + 1e-15
        assert float((Vb[:k] @ w).abs().max()) <= 1e-13 * max(1.0, float(torch.linalg.vector_norm(w)))
