"""Block-stored environments (one tensor per Jordan level, as BlockTensorMap
holds them, src/environments/abstract_envs.jl:32-58) and block-stored Jordan MPO
tensors (nonzero blocks, identities as scalars, jordanmpotensor.jl:44-69) at the
C ABI: every per-level entry point equals its dense counterpart bit for bit,
and the dense results match the oracle (fp64 1e-12)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

LEVELS = {"xxz": [1, 3, 1], "j1j2": [1, 3, 3, 1], "heis3": [1, 3, 3, 3, 1]}


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def bulk(mk, name):
    M = mk.models
    return {"xxz": M.xxz_bulk(0.7), "j1j2": M.j1j2_bulk(1.0, 0.3), "heis3": M.heisenberg_range_bulk([1.0, 0.4, 0.2])}[name]


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", list(LEVELS))
def test_levels_prepares_equal_dense(mk, name):
    import torch
    T, Bk = mk.tensors, mk.blocks
    rng = np.random.default_rng(51)
    W = bulk(mk, name)
    w, lv = W.shape[0], LEVELS[name]
    assert sum(lv) == w
    D = 67
    GL, GR = rng.standard_normal((D, w, D)) / D, rng.standard_normal((D, w, D)) / D
    GLd, GRd, Wd = T.from_numpy(GL), T.from_numpy(GR), T.from_numpy(W)
    GLl, GRl = Bk.split_levels(GLd, lv), Bk.split_levels(GRd, lv)
    precs = ("fp64", "bf16x3")
    print("levels", name, flush=True)
    # two-site Jordan
    x = T.from_numpy(rng.standard_normal((D, 2, D, 2)))
    dense = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GLd, Wd, Wd, GRd), precisions=precs)
    lvl = Bk.prepare_ac2_jordan_levels(GLl, W, W, GRl, precisions=precs)
    for p in precs:
        assert torch.equal(lvl(x, precision=p), dense(x, precision=p)), p
    y_or = O.jordan_ac2_apply(O.jordan_ac2_prepare(GL, W, W, GR), T.to_numpy(x))
    assert rel(T.to_numpy(lvl(x)), y_or) < 1e-12
    # one-site Jordan
    x1 = T.from_numpy(rng.standard_normal((D, 2, D)))
    dense1 = mk.prepare_operator(mk.JordanMPO_AC_Hamiltonian(GLd, Wd, GRd), precisions=precs)
    lvl1 = Bk.prepare_ac_jordan_levels(GLl, W, GRl, precisions=precs)
    for p in precs:
        assert torch.equal(lvl1(x1, precision=p), dense1(x1, precision=p)), p
    # zero-site (C_hamiltonian densifies every level in the reference)
    x0 = T.from_numpy(rng.standard_normal((D, D)))
    dense0 = mk.prepare_operator(mk.MPO_C_Hamiltonian(GLd, GRd), precisions=precs)
    lvl0 = Bk.prepare_c_levels(GLl, GRl, precisions=precs)
    for p in precs:
        assert torch.equal(lvl0(x0, precision=p), dense0(x0, precision=p)), p
    assert rel(T.to_numpy(lvl0(x0)), O.c_apply(GL, GR, T.to_numpy(x0))) < 1e-12
    for op in (dense, lvl, dense1, lvl1, dense0, lvl0):
        op.free()


@pytest.mark.parametrize("name", list(LEVELS))
def test_levels_transfers_equal_dense(mk, name):
    import torch
    T, Bk = mk.tensors, mk.blocks
    rng = np.random.default_rng(52)
    W = bulk(mk, name)
    w, lv = W.shape[0], LEVELS[name]
    Dm, Dn, Dc, Da = 40, 43, 41, 42
    A, Ab = rng.standard_normal((Dn, 2, Dc)), rng.standard_normal((Dm, 2, Da))
    x = rng.standard_normal((Dm, w, Dn))
    Ad, Abd, Wd = T.from_numpy(A), T.from_numpy(Ab), T.from_numpy(W)
    ys = Bk.transfer_left_levels(Bk.split_levels(T.from_numpy(x), lv), W, Ad, Abd, lv)
    y = torch.cat(ys, dim=1)
    yd = mk.transfer_left(T.from_numpy(x), Wd, Ad, Abd)
    assert torch.equal(y, yd)
    assert rel(T.to_numpy(y), O.transfer_left_mpo(x, W, A, Ab)) < 1e-12
    # right: A (Da x d x Di), Abar (Dc x d x Do), v (Di x w x Do)
    Ar, Abr = rng.standard_normal((Da, 2, Dn)), rng.standard_normal((Dc, 2, Dm))
    v = rng.standard_normal((Dn, w, Dm))
    Ard, Abrd = T.from_numpy(Ar), T.from_numpy(Abr)
    ys = Bk.transfer_right_levels(Bk.split_levels(T.from_numpy(v), lv), W, Ard, Abrd, lv)
    y = torch.cat(ys, dim=1)
    yd = mk.transfer_right(T.from_numpy(v), Wd, Ard, Abrd)
    assert torch.equal(y, yd)
    assert rel(T.to_numpy(y), O.transfer_right_mpo(v, W, Ar, Abr)) < 1e-12


def test_levels_errors(mk):
    """Level widths that do not match the MPO bond raise before any work."""
    T, Bk = mk.tensors, mk.blocks
    rng = np.random.default_rng(53)
    W = mk.models.xxz_bulk(1.0)
    D = 8
    GLd = T.from_numpy(rng.standard_normal((D, 5, D)))
    with pytest.raises(mk.SpaceMismatch):
        Bk.prepare_ac2_jordan_levels(Bk.split_levels(GLd, [1, 4]), W, W, Bk.split_levels(GLd, [2, 2, 1])[:2])
