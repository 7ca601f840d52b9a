"""Isolated derivative applies at the benchmark sizes (D = 512 .. 4096) against the oracle.

Workload of the isolated-apply configuration (the reference's AC2 benchmark,
AC2_benchmarks.jl:24-72): XXZ-type Jordan MPO (w = 5, d = 2) with random C/B
blocks and identity corners, dense GL/GR of shape D x 5 x D and x with i.i.d.
Gaussian entries scaled to unit Frobenius norm (seeds 1, 2 for GL, GR; 3 for W;
4 for x).  These sizes run the 256 x 256 mixed GEMM (D >= 2048) with its
segment switching, and a size that is not a multiple of 256 (D = 2056) runs its
edge tiles.

Where the full oracle apply is too slow for the host (generic w = 5 at D = 4096,
1.1e13 flops), the check uses an exact restriction: the derivative applies are
linear in GR's bra index, so the oracle evaluated with GR[:, :, S] returns
exactly the columns S of y (the same identity lets the GPU result be compared
column by column).  The restriction itself is pinned on the CPU
(test_oracle_column_restriction in this file runs without a GPU).

Tolerances (relative Frobenius error against the float64 oracle):
  fp64 1e-12, fp32 1e-5, bf16x3 3e-4.
"""
import numpy as np
import pytest

from oracle import oracle as O

TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 3e-4}
PRECS = ("fp64", "fp32", "bf16x3")


def jordan_w(rng, d=2, nmid=3):
    """Jordan MPO tensor with random C (start) and B (end) blocks, identity corners."""
    w = nmid + 2
    W = np.zeros((w, d, d, w))
    W[0, :, :, 0] = np.eye(d)
    W[-1, :, :, -1] = np.eye(d)
    W[0, :, :, 1:-1] = rng.standard_normal((d, d, nmid))
    W[1:-1, :, :, -1] = rng.standard_normal((nmid, d, d))
    return W


def unit(seed, shape):
    a = np.random.default_rng(seed).standard_normal(shape)
    return a / np.linalg.norm(a)


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_oracle_column_restriction():
    """y[..., S, :] of every derivative apply equals the apply with GR[:, :, S] (CPU)."""
    rng = np.random.default_rng(0)
    D, w = 9, 5
    W = jordan_w(rng)
    GL, GR = rng.standard_normal((D, w, D)), rng.standard_normal((D, w, D))
    S = np.array([0, 3, 4, 8])
    x2 = rng.standard_normal((D, 2, D, 2))
    y = O.ac2_apply_generic(GL, W, W, GR, x2)
    assert rel(O.ac2_apply_generic(GL, W, W, GR[:, :, S], x2), y[:, :, S, :]) < 1e-14
    x1 = rng.standard_normal((D, 2, D))
    y1 = O.ac_apply_generic(GL, W, GR, x1)
    assert rel(O.ac_apply_generic(GL, W, GR[:, :, S], x1), y1[:, :, S]) < 1e-14


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def dev(mk, a):
    return mk.tensors.from_numpy(a)


def host(mk, t):
    import torch
    torch.cuda.synchronize()
    return mk.tensors.to_numpy(t)


def inputs(D, w=5):
    W = jordan_w(np.random.default_rng(3), nmid=w - 2)
    return W, unit(1, (D, w, D)), unit(2, (D, w, D))


def check_all(mk, op, x, y_ref, cols=None, label=""):
    import torch
    xd = dev(mk, x)
    for p in PRECS:
        y = op(xd, precision=p)
        if cols is not None:
            y = y[:, :, torch.as_tensor(cols, device=y.device)]
        e = rel(host(mk, y), y_ref)
        print(f"  {label} {p}: {e:.2e}", flush=True)
        assert e < TOL[p], (label, p, e)
        del y


@pytest.mark.gpu
@pytest.mark.parametrize("D", [512, 1024, 2048, 2056, 4096])
def test_jordan_ac2_bench_sizes(mk, D):
    """The headline apply (Jordan two-site, nearest neighbour) against the full oracle."""
    W, GL, GR = inputs(D)
    x = unit(4, (D, 2, D, 2))
    print("jordan ac2 D", D, flush=True)
    y_ref = O.jordan_ac2_apply(O.jordan_ac2_prepare(GL, W, W, GR), x)
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, W), dev(mk, GR)),
                             precisions=PRECS)
    check_all(mk, op, x, y_ref, label=f"D={D}")
    op.free()


@pytest.mark.gpu
@pytest.mark.parametrize("D", [512, 2048, 4096])
def test_jordan_ac_bench_sizes(mk, D):
    """Single-site Jordan derivative at the same sizes (full oracle)."""
    W, GL, GR = inputs(D)
    x = unit(4, (D, 2, D))
    print("jordan ac D", D, flush=True)
    y_ref = O.jordan_ac_apply(O.jordan_ac_prepare(GL, W, GR), x)
    op = mk.prepare_operator(mk.JordanMPO_AC_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, GR)), precisions=PRECS)
    check_all(mk, op, x, y_ref, label=f"D={D}")
    op.free()


@pytest.mark.gpu
@pytest.mark.parametrize("D", [512, 1024, 2048, 4096])
def test_generic_ac2_bench_sizes(mk, D):
    """Generic (precomputed) w = 5 two-site derivative; D = 4096 through 48 sampled
    output columns (exact restriction of the oracle)."""
    W, GL, GR = inputs(D)
    x = unit(4, (D, 2, D, 2))
    cols = None if D <= 2048 else np.sort(np.random.default_rng(9).choice(D, 48, replace=False))
    print("generic ac2 D", D, flush=True)
    y_ref = O.ac2_apply_generic(GL, W, W, GR if cols is None else GR[:, :, cols], x)
    op = mk.prepare_operator(mk.MPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, W), dev(mk, GR)),
                             precisions=PRECS)
    check_all(mk, op, x, y_ref, cols, label=f"D={D}")
    op.free()


@pytest.mark.gpu
@pytest.mark.parametrize("D", [2048, 4096])
def test_generic_ac_bench_sizes(mk, D):
    """Generic w = 5 single-site derivative; D = 4096 through sampled output columns."""
    W, GL, GR = inputs(D)
    x = unit(4, (D, 2, D))
    cols = None if D <= 2048 else np.sort(np.random.default_rng(10).choice(D, 64, replace=False))
    print("generic ac D", D, flush=True)
    y_ref = O.ac_apply_generic(GL, W, GR if cols is None else GR[:, :, cols], x)
    op = mk.prepare_operator(mk.MPO_AC_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, GR)), precisions=PRECS)
    check_all(mk, op, x, y_ref, cols, label=f"D={D}")
    op.free()
