"""The allocation-free apply (mk_apply_noalloc) against mk_apply through the C ABI.

Both entry points run the same kernels; the allocation-free one takes its
scratch from the library's fixed device array instead of the context arena.
Its results must equal mk_apply's (same kernels, same inputs: bitwise, held
here to 1e-15 relative) and the oracle's to the apply tolerances (fp64 1e-12,
fp32 1e-5, bf16x3 3e-4).  Two calls in a row, the second on a smaller operator,
show that stale scratch contents from a larger call do not leak into a smaller
one; an apply whose scratch exceeds the array fails with MK_BAD_SHAPE before any
work is enqueued.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 3e-4}
PRECS = ("fp64", "fp32", "bf16x3")


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def jordan_case(mk, D, seed):
    rng = np.random.default_rng(seed)
    W = mk.models.xxz_bulk(1.0)
    GL, GR = rng.standard_normal((D, 5, D)) / D, rng.standard_normal((D, 5, D)) / D
    x = rng.standard_normal((D, 2, D, 2))
    T = mk.tensors
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(T.from_numpy(GL), T.from_numpy(W), T.from_numpy(W),
                                                          T.from_numpy(GR)), precisions=PRECS)
    return op, x, O.jordan_ac2_apply(O.jordan_ac2_prepare(GL, W, W, GR), x)


def apply_both(mk, op, x, prec, alpha=1.0, beta=0.0, y0=None):
    import torch
    from mpskit_amd._lib import PREC_NAMES, check, load
    lib = load()
    T = mk.tensors
    xd = T.from_numpy(x)
    y_ref = T.from_numpy(y0) if y0 is not None else T.zeros(op.y_shape)
    y_na = T.from_numpy(y0) if y0 is not None else T.zeros(op.y_shape)
    op.ctx.sync_stream()
    check(lib.mk_apply(op.handle, T.ptr(xd), T.ptr(y_ref), alpha, beta, PREC_NAMES[prec]))
    check(lib.mk_apply_noalloc(op.handle, T.ptr(xd), T.ptr(y_na), alpha, beta, PREC_NAMES[prec]))
    torch.cuda.synchronize()
    return T.to_numpy(y_ref), T.to_numpy(y_na)


@pytest.mark.parametrize("prec", PRECS)
def test_noalloc_matches_apply_two_calls_shrinking(mk, prec):
    print("noalloc", prec, flush=True)
    for D, seed in ((1030, 1), (301, 2)):       # second call smaller than the first
        op, x, y_or = jordan_case(mk, D, seed)
        y_ref, y_na = apply_both(mk, op, x, prec)
        e_same, e_or = rel(y_na, y_ref), rel(y_na, y_or)
        print(f"  D={D}: vs mk_apply {e_same:.1e}, vs oracle {e_or:.1e}", flush=True)
        assert e_same <= 1e-15, (D, prec, e_same)
        assert e_or < TOL[prec], (D, prec, e_or)
        op.free()


def test_noalloc_alpha_beta_and_other_operators(mk):
    rng = np.random.default_rng(5)
    op, x, y_or = jordan_case(mk, 64, 3)
    y0 = rng.standard_normal(op.y_shape)
    for prec in PRECS:
        y_ref, y_na = apply_both(mk, op, x, prec, alpha=0.5, beta=-2.0, y0=y0)
        assert rel(y_na, y_ref) <= 1e-15
        assert rel(y_na, 0.5 * y_or - 2.0 * y0) < TOL[prec]
    op.free()
    # a generic two-site operator (its intermediate x R also comes from the array)
    # and a one-site Jordan operator
    T = mk.tensors
    D, w = 70, 5
    W = rng.standard_normal((w, 2, 2, w))
    GL, GR = rng.standard_normal((D, w, D)), rng.standard_normal((D, w, D))
    gen = mk.prepare_operator(mk.MPO_AC2_Hamiltonian(T.from_numpy(GL), T.from_numpy(W), T.from_numpy(W),
                                                     T.from_numpy(GR)), precisions=PRECS)
    x2 = rng.standard_normal((D, 2, D, 2))
    Wj = mk.models.xxz_bulk(0.7)
    ac = mk.prepare_operator(mk.JordanMPO_AC_Hamiltonian(T.from_numpy(GL), T.from_numpy(Wj), T.from_numpy(GR)),
                             precisions=PRECS)
    x1 = rng.standard_normal((D, 2, D))
    for prec in PRECS:
        y_ref, y_na = apply_both(mk, gen, x2, prec)
        assert rel(y_na, y_ref) <= 1e-15
        assert rel(y_na, O.ac2_apply_generic(GL, W, W, GR, x2)) < TOL[prec]
        y_ref, y_na = apply_both(mk, ac, x1, prec)
        assert rel(y_na, y_ref) <= 1e-15
        assert rel(y_na, O.jordan_ac_apply(O.jordan_ac_prepare(GL, Wj, GR), x1)) < TOL[prec]
    gen.free()
    ac.free()


def test_noalloc_rejects_oversized_scratch(mk):
    """D = 4096 BF16X3 needs 2 x 135 MB of operand views > 160 MiB: MK_BAD_SHAPE,
    raised on the host before any launch; the FP64 apply of the same operator
    needs no scratch and runs."""
    import torch
    from mpskit_amd._lib import BF16X3, FP64, load
    lib = load()
    T = mk.tensors
    D = 4096
    W = mk.models.xxz_bulk(1.0)
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    GL = T.empty((D, 5, D))
    GL.copy_(torch.randn((D, 5, D), dtype=torch.float64, device="cuda", generator=g))
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, T.from_numpy(W), T.from_numpy(W), GL),
                             precisions=("fp64", "bf16x3"))
    x = T.empty(op.x_shape)
    x.copy_(torch.randn(op.x_shape, dtype=torch.float64, device="cuda", generator=g))
    y1, y2 = T.zeros(op.y_shape), T.zeros(op.y_shape)
    op.ctx.sync_stream()
    assert lib.mk_apply_noalloc(op.handle, T.ptr(x), T.ptr(y1), 1.0, 0.0, BF16X3) == 1
    assert b"fixed scratch" in lib.mk_last_error()
    assert lib.mk_apply_noalloc(op.handle, T.ptr(x), T.ptr(y1), 1.0, 0.0, FP64) == 0
    assert lib.mk_apply(op.handle, T.ptr(x), T.ptr(y2), 1.0, 0.0, FP64) == 0
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    op.free()
