"""GPU parity of the library's blocked Cholesky (mk_potrf) against torch's
(rocSOLVER) factorisation, and the Cholesky QR of the gauge steps built on it
against Householder QR.  Tolerances: factors to 1e-12 relative (Frobenius),
orthonormality to 1e-13 per column."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd as mk
    mk.load()
    return mk, torch


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 257, 1000, 1024, 2048])
def test_potrf_matches_torch(env, n):
    mk, torch = env
    from mpskit_amd.krylov import potrf_
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    X = torch.randn(n + 7, n, dtype=torch.float64, device="cuda", generator=g)
    G = (X.t() @ X).contiguous()
    ref = torch.linalg.cholesky(G)
    print("potrf n", n, flush=True)
    L, info = potrf_(G.clone())
    torch.cuda.synchronize()
    assert info == 0
    assert float(torch.linalg.matrix_norm(L - ref) / torch.linalg.matrix_norm(ref)) < 1e-12
    assert bool((torch.triu(L, 1) == 0).all())


def test_potrf_reports_the_failing_pivot(env):
    mk, torch = env
    from mpskit_amd.krylov import potrf_
    n = 200
    G = torch.eye(n, dtype=torch.float64, device="cuda")
    G[130, 130] = -1.0
    _, info = potrf_(G)
    assert info == 131
    G = torch.eye(n, dtype=torch.float64, device="cuda")
    G[3, 3] = float("nan")
    _, info = potrf_(G)
    assert info == 4


@pytest.mark.parametrize("m,n", [(2048, 1024), (4096, 2048), (300, 257), (96, 96)])
def test_cholesky_qr_matches_householder(env, m, n):
    mk, torch = env
    from mpskit_amd.states import _householder_pos_qr, pos_qr
    g = torch.Generator(device="cuda")
    g.manual_seed(m + n)
    A = torch.randn(m, n, dtype=torch.float64, device="cuda", generator=g)
    Q, R = pos_qr(A)
    Qh, Rh = _householder_pos_qr(A)
    E = Q.t() @ Q
    E.diagonal().sub_(1.0)
    assert float(torch.linalg.matrix_norm(E)) < 1e-13 * n ** 0.5
    assert float(torch.linalg.matrix_norm(R - Rh) / torch.linalg.matrix_norm(Rh)) < 1e-12
    assert float(torch.linalg.matrix_norm(Q @ R - A) / torch.linalg.matrix_norm(A)) < 1e-13
