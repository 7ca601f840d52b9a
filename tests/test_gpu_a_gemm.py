"""GPU parity of the raw MFMA GEMM kernels (mk_gemm) against torch float64 matmul.

Every contraction of the path reduces to these kernels, so they are checked
first, on tile-aligned and ragged shapes, all transpositions and precisions.
Tolerances are relative Frobenius errors: fp64 1e-13, fp32 1e-5, bf16x3 3e-4.
"""
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-13, "fp32": 1e-5, "bf16x3": 3e-4}
SHAPES = [(1, 16, 16), (16, 16, 16), (64, 64, 64), (67, 45, 131), (128, 256, 96), (300, 257, 129),
          (2048, 2048, 64), (2176, 2048, 200),
          (4096, 4096, 96), (4200, 4100, 130)]       # the last two take the 256x256 mixed kernel


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd as mk
    mk.load()
    return mk, torch


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_vs_torch(env, prec, ta, tb):
    mk, torch = env
    T = mk.tensors
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    for (M, N, K) in SHAPES:
        print(f"gemm {prec} ta={ta} tb={tb} M={M} N={N} K={K}", flush=True)
        A = T.as_colmajor(torch.randn((K, M) if ta else (M, K), dtype=torch.float64, device="cuda", generator=g))
        B = T.as_colmajor(torch.randn((N, K) if tb else (K, N), dtype=torch.float64, device="cuda", generator=g))
        C0 = T.as_colmajor(torch.randn((M, N), dtype=torch.float64, device="cuda", generator=g))
        ref = 1.5 * ((A.t() if ta else A) @ (B.t() if tb else B)) - 0.25 * C0
        C = C0.clone(memory_format=torch.preserve_format)
        assert T.is_colmajor(C)
        mk.gemm(A, B, C, alpha=1.5, beta=-0.25, transa=ta, transb=tb, precision=prec)
        torch.cuda.synchronize()
        err = float(torch.linalg.vector_norm(C - ref) / torch.linalg.vector_norm(ref))
        assert err < TOL[prec], (M, N, K, err)


def test_gemm_beta_zero_ignores_nan_output(env):
    mk, torch = env
    T = mk.tensors
    A = T.as_colmajor(torch.randn(70, 33, dtype=torch.float64, device="cuda"))
    B = T.as_colmajor(torch.randn(33, 90, dtype=torch.float64, device="cuda"))
    for prec in ("fp64", "fp32", "bf16x3"):
        C = T.empty((70, 90))
        C.fill_(float("nan"))
        mk.gemm(A, B, C, beta=0.0, precision=prec)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(C).all()), prec


def test_mfma_output_lane_layout(env):
    """One 16x16x4 block with rank-1 structured operands: C[i,j] = (i+1) * 100 (j+1)
    pins the MFMA output row/column assignment of every precision."""
    mk, torch = env
    T = mk.tensors
    A = torch.zeros(16, 4, dtype=torch.float64, device="cuda")
    A[:, 0] = torch.arange(16, dtype=torch.float64, device="cuda") + 1
    B = torch.zeros(4, 16, dtype=torch.float64, device="cuda")
    B[0, :] = 100 * (torch.arange(16, dtype=torch.float64, device="cuda") + 1)
    ref = A @ B
    for prec in ("fp64", "fp32", "bf16x3"):
        C = mk.gemm(T.as_colmajor(A), T.as_colmajor(B), precision=prec)
        torch.cuda.synchronize()
        assert float((C - ref).abs().max()) == 0.0, prec
