"""Prepared transfer matrices (``mk_prepare_transfer_left`` /
``mk_prepare_transfer_right``) against the oracle's transfers
(transfer.jl:18-55, :125-130) in FP64, FP32 and BF16x3: MPS and MPO
transfers, bond dimensions that are not multiples of 8 (padded mixed
operands), rectangular maps, alpha / beta, and the operator's own copies of
its tensors (changing the caller's tensors after the prepare changes
nothing).  The FP64 apply equals the unprepared entry point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

TOL = {"fp64": 1e-13, "fp32": 1e-5, "bf16x3": 3e-4}
PRECS = ("fp64", "fp32", "bf16x3")


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("Dm,Dn,Dc,Da", [(37, 37, 37, 37), (300, 301, 299, 302), (1024, 1024, 1024, 1024)])
def test_prepared_mps_transfers(mk, Dm, Dn, Dc, Da):
    T = mk.tensors
    rng = np.random.default_rng(Dm + Dn)
    d = 2
    A, Ab = rng.standard_normal((Dn, d, Dc)) / np.sqrt(Dn), rng.standard_normal((Dm, d, Da)) / np.sqrt(Dm)
    v = rng.standard_normal((Dm, Dn))
    print("mps transfer left", Dm, Dn, Dc, Da, flush=True)
    op = mk.prepare_transfer_left(T.from_numpy(A), T.from_numpy(Ab), precisions=PRECS)
    ref = O.transfer_left_mps(v, A, Ab)
    vd = T.from_numpy(v)
    y64 = T.to_numpy(op(vd))
    assert np.array_equal(y64, T.to_numpy(mk.transfer_left(vd, T.from_numpy(A), T.from_numpy(Ab))))
    for p in PRECS:
        e = rel(T.to_numpy(op(vd, precision=p)), ref)
        assert e < TOL[p], ("left", p, e)
    op.free()
    # right: A (Da', d, Di), Ab (Dc', d, Do), v (Di, Do)
    Ar, Abr = rng.standard_normal((Da, d, Dn)) / np.sqrt(Dn), rng.standard_normal((Dc, d, Dm)) / np.sqrt(Dm)
    w = rng.standard_normal((Dn, Dm))
    op = mk.prepare_transfer_right(T.from_numpy(Ar), T.from_numpy(Abr), precisions=PRECS)
    ref = O.transfer_right_mps(w, Ar, Abr)
    wd = T.from_numpy(w)
    assert np.array_equal(T.to_numpy(op(wd)), T.to_numpy(mk.transfer_right(wd, T.from_numpy(Ar), T.from_numpy(Abr))))
    for p in PRECS:
        e = rel(T.to_numpy(op(wd, precision=p)), ref)
        assert e < TOL[p], ("right", p, e)
    op.free()


@pytest.mark.parametrize("D,model", [(45, "xxz"), (301, "j1j2"), (1024, "j1j2")])
def test_prepared_mpo_transfers(mk, D, model):
    T = mk.tensors
    rng = np.random.default_rng(D)
    W = {"xxz": mk.models.xxz_bulk(0.7), "j1j2": mk.models.j1j2_bulk(1.0, 0.24)}[model]
    w, d = W.shape[0], W.shape[1]
    A, Ab = rng.standard_normal((D, d, D)) / np.sqrt(D), rng.standard_normal((D, d, D)) / np.sqrt(D)
    x = rng.standard_normal((D, w, D))
    print("mpo transfers", model, D, flush=True)
    Wd, Ad, Abd = T.from_numpy(W), T.from_numpy(A), T.from_numpy(Ab)
    opl = mk.prepare_transfer_left(Ad, Abd, O=Wd, precisions=PRECS)
    opr = mk.prepare_transfer_right(Ad, Abd, O=Wd, precisions=PRECS)
    refl, refr = O.transfer_left_mpo(x, W, A, Ab), O.transfer_right_mpo(x, W, A, Ab)
    xd = T.from_numpy(x)
    assert np.array_equal(T.to_numpy(opl(xd)), T.to_numpy(mk.transfer_left(xd, Wd, Ad, Abd)))
    assert np.array_equal(T.to_numpy(opr(xd)), T.to_numpy(mk.transfer_right(xd, Wd, Ad, Abd)))
    # the operators hold their own copies
    Ad.zero_()
    Abd.zero_()
    Wd.zero_()
    for p in PRECS:
        el = rel(T.to_numpy(opl(xd, precision=p)), refl)
        er = rel(T.to_numpy(opr(xd, precision=p)), refr)
        assert el < TOL[p] and er < TOL[p], (p, el, er)
        # y = alpha T x + beta y
        y0 = rng.standard_normal(refl.shape)
        out = T.from_numpy(y0)
        opl(xd, precision=p, out=out, alpha=-0.5, beta=2.0)
        e = rel(T.to_numpy(out), -0.5 * refl + 2.0 * y0)
        assert e < TOL[p], ("alpha/beta", p, e)
    opl.free()
    opr.free()


def test_prepared_transfer_rejects_bad_shapes(mk):
    T = mk.tensors
    rng = np.random.default_rng(3)
    A = T.from_numpy(rng.standard_normal((5, 2, 6)))
    with pytest.raises(mk.SpaceMismatch):
        mk.prepare_transfer_left(A, T.from_numpy(rng.standard_normal((5, 3, 6))))
    op = mk.prepare_transfer_left(A, A, precisions=("fp64",))
    with pytest.raises(mk.MKError):
        op(T.from_numpy(rng.standard_normal((5, 5))), precision="bf16x3")
    with pytest.raises(mk.SpaceMismatch):
        op(T.from_numpy(rng.standard_normal((6, 5))))
    op.free()
