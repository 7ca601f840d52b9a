"""GPU parity of the HIP derivative / transfer kernels against the float64 oracle.

Tolerances (relative Frobenius error of y):
  fp64   <= 1e-12   (the reference arithmetic)
  fp32   <= 1e-5    (FP32 MFMA products, fp32 accumulation)
  bf16x3 <= 3e-4    (hi*hi + hi*lo + lo*hi bf16 products, fp32 accumulation)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pathlib import Path  # noqa: E402

from oracle import brute as B  # noqa: E402
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


@pytest.fixture(scope="module")
def fixtures():
    data = np.load(Path(__file__).with_name("golden") / "apply_vectors.npz")
    cases = {}
    for k in data.files:
        c, a = k.split("/")
        cases.setdefault(c, {})[a] = data[k]
    return cases


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev(mk, a):
    return mk.tensors.from_numpy(a)


def host(mk, t):
    import torch
    torch.cuda.synchronize()
    return mk.tensors.to_numpy(t)


def build_op(mk, kind, g):
    d = {k: dev(mk, v) for k, v in g.items() if k not in ("y",)}
    if kind == "ac2gen":
        H = mk.MPO_AC2_Hamiltonian(d["GL"], d["W1"], d["W2"], d["GR"])
    elif kind == "acgen":
        H = mk.MPO_AC_Hamiltonian(d["GL"], d["W"], d["GR"])
    elif kind == "c":
        H = mk.MPO_C_Hamiltonian(d["GL"], d["GR"])
    elif kind in ("ac2jor", "ac2chain"):
        H = mk.JordanMPO_AC2_Hamiltonian(d["GL"], d["W1"], d["W2"], d["GR"])
    elif kind in ("acjor", "acchain"):
        H = mk.JordanMPO_AC_Hamiltonian(d["GL"], d["W"], d["GR"])
    else:
        return None, d
    return mk.prepare_operator(H, precisions=PRECS), d


def test_fixture_vectors_all_precisions(mk, fixtures):
    worst = {p: 0.0 for p in PRECS}
    failures = []
    for case, g in sorted(fixtures.items()):
        kind = case.split("_")[0]
        print("case", case, flush=True)
        op, d = build_op(mk, kind, g)
        if op is None:
            if kind == "tlmps":
                y = mk.transfer_left(d["v"], d["A"], d["Ab"])
            elif kind == "trmps":
                y = mk.transfer_right(d["v"], d["A"], d["Ab"])
            elif kind == "tlmpo":
                y = mk.transfer_left(d["x"], d["O"], d["A"], d["Ab"])
            elif kind == "trmpo":
                y = mk.transfer_right(d["v"], d["O"], d["A"], d["Ab"])
            e = rel(host(mk, y), g["y"])
            if not e < TOL["fp64"]:
                failures.append((case, "fp64", e))
            worst["fp64"] = max(worst["fp64"], e)
            continue
        for p in PRECS:
            e = rel(host(mk, op(d["x"], precision=p)), g["y"])
            worst[p] = max(worst[p], e)
            if not e < TOL[p]:
                failures.append((case, p, e))
    print("worst relative errors", worst)
    assert not failures, failures


@pytest.mark.parametrize("D", [37, 130, 300, 1030])
def test_jordan_ac2_odd_sizes_vs_oracle(mk, D):
    """Tile-edge bounds (D not a multiple of the 64/128 tiles) and the 128x128 path (D=1030)."""
    rng = np.random.default_rng(D)
    W = __import__("mpskit_amd").models.xxz_bulk(1.0)
    GL, GR = rng.standard_normal((D, 5, D)) / D, rng.standard_normal((D, 5, D)) / D
    x = rng.standard_normal((D, 2, D, 2))
    y_ref = O.jordan_ac2_apply(O.jordan_ac2_prepare(GL, W, W, GR), x)
    print("jordan ac2 D", D, flush=True)
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, W), dev(mk, GR)),
                             precisions=PRECS)
    xd = dev(mk, x)
    for p in PRECS:
        e = rel(host(mk, op(xd, precision=p)), y_ref)
        print(f"  {p}: {e:.2e}", flush=True)
        assert e < TOL[p], (p, e)


@pytest.mark.parametrize("model,D", [("heis3", 37), ("heis3", 130), ("expising", 45), ("expising", 300)])
def test_jordan_ac2_aa_term_vs_oracle(mk, model, D):
    """Connected A blocks (range 3, exponential decay): the HIP Jordan operator's AA
    term over _connected_channels (hamiltonian_derivatives.jl:243-270, :352-362) plus
    its CA / AB folds, against the oracle in every precision."""
    rng = np.random.default_rng(D + 7)
    M = __import__("mpskit_amd").models
    W = M.heisenberg_range_bulk([1.0, 0.45, 0.3]) if model == "heis3" else M.exp_ising_bulk(0.8, 0.6, 0.9)
    w = W.shape[0]
    GL, GR = rng.standard_normal((D, w, D)) / D, rng.standard_normal((D, w, D)) / D
    x = rng.standard_normal((D, 2, D, 2))
    T = O.jordan_ac2_prepare(GL, W, W, GR)
    assert T["AA"] is not None
    y_ref = O.jordan_ac2_apply(T, x)
    print("jordan ac2 AA", model, "D", D, flush=True)
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, W), dev(mk, GR)),
                             precisions=PRECS)
    xd = dev(mk, x)
    for p in PRECS:
        e = rel(host(mk, op(xd, precision=p)), y_ref)
        print(f"  {p}: {e:.2e}", flush=True)
        assert e < TOL[p], (p, e)


@pytest.mark.parametrize("D,w", [(45, 3), (129, 5), (70, 8)])
def test_generic_ops_odd_sizes_vs_oracle(mk, D, w):
    rng = np.random.default_rng(D + w)
    W = rng.standard_normal((w, 2, 2, w))
    Da, Df, Di, Dc = D, D + 3, D + 1, D - 2
    GL, GR = rng.standard_normal((Da, w, Df)), rng.standard_normal((Di, w, Dc))
    print("generic D", D, "w", w, flush=True)
    ops = {
        "ac2": (mk.MPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, W), dev(mk, GR)),
                rng.standard_normal((Df, 2, Di, 2)), lambda x: O.ac2_apply_generic(GL, W, W, GR, x)),
        "ac": (mk.MPO_AC_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, GR)),
               rng.standard_normal((Df, 2, Di)), lambda x: O.ac_apply_generic(GL, W, GR, x)),
        "c": (mk.MPO_C_Hamiltonian(dev(mk, GL), dev(mk, GR)), rng.standard_normal((Df, Di)),
              lambda x: O.c_apply(GL, GR, x)),
    }
    for name, (H, x, f) in ops.items():
        op = mk.prepare_operator(H, precisions=PRECS)
        y_ref = f(x)
        for p in PRECS:
            e = rel(host(mk, op(dev(mk, x), precision=p)), y_ref)
            assert e < TOL[p], (name, p, e)


def test_alpha_beta_and_unprepared_call(mk):
    rng = np.random.default_rng(9)
    D = 33
    W = __import__("mpskit_amd").models.tfim_bulk(0.9)
    GL, GR = rng.standard_normal((D, 3, D)), rng.standard_normal((D, 3, D))
    x, y0 = rng.standard_normal((D, 2, D, 2)), rng.standard_normal((D, 2, D, 2))
    H = mk.JordanMPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, W), dev(mk, W), dev(mk, GR))
    hx = O.jordan_ac2_apply(O.jordan_ac2_prepare(GL, W, W, GR), x)
    assert rel(host(mk, H * dev(mk, x)), hx) < 1e-12          # raw call prepares lazily
    op = mk.prepare_operator(H)
    out = dev(mk, y0)
    op(dev(mk, x), out=out, alpha=2.0, beta=0.5)
    assert rel(host(mk, out), 2.0 * hx + 0.5 * y0) < 1e-12


def test_transfers_odd_sizes_and_regularize(mk):
    rng = np.random.default_rng(21)
    Dm, Dn, Dc, Da, w = 67, 70, 65, 66, 5
    A, Ab = rng.standard_normal((Dn, 2, Dc)), rng.standard_normal((Dm, 2, Da))
    O5 = rng.standard_normal((w, 2, 2, 4))
    v = rng.standard_normal((Dm, Dn))
    assert rel(host(mk, mk.transfer_left(dev(mk, v), dev(mk, A), dev(mk, Ab))), O.transfer_left_mps(v, A, Ab)) < 1e-12
    vr = rng.standard_normal((Dc, Da))
    assert rel(host(mk, mk.transfer_right(dev(mk, vr), dev(mk, A), dev(mk, Ab))),
               O.transfer_right_mps(vr, A, Ab)) < 1e-12
    x = rng.standard_normal((Dm, w, Dn))
    assert rel(host(mk, mk.transfer_left(dev(mk, x), dev(mk, O5), dev(mk, A), dev(mk, Ab))),
               O.transfer_left_mpo(x, O5, A, Ab)) < 1e-12
    v3 = rng.standard_normal((Dc, 4, Da))
    assert rel(host(mk, mk.transfer_right(dev(mk, v3), dev(mk, O5), dev(mk, A), dev(mk, Ab))),
               O.transfer_right_mpo(v3, O5, A, Ab)) < 1e-12
    # regularize (bond tensor and environment-shaped)
    D = 40
    lvec, rvec = rng.standard_normal((D, D)), rng.standard_normal((D, D))
    for shape in ((D, D), (D, 3, D)):
        vv = rng.standard_normal(shape)
        t = dev(mk, vv)
        mk.regularize_(t, dev(mk, lvec), dev(mk, rvec))
        assert rel(host(mk, t), O.regularize(vv, lvec, rvec)) < 1e-12
    tm = mk.TransferMatrix(dev(mk, A), None, dev(mk, Ab))
    assert rel(host(mk, dev(mk, v) * tm), O.transfer_left_mps(v, A, Ab)) < 1e-12


def test_shape_errors_raise(mk):
    rng = np.random.default_rng(1)
    D = 8
    W = __import__("mpskit_amd").models.xxz_bulk(1.0)
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(dev(mk, rng.standard_normal((D, 5, D))), dev(mk, W),
                                                          dev(mk, W), dev(mk, rng.standard_normal((D, 5, D)))))
    with pytest.raises(mk.SpaceMismatch):
        op(dev(mk, rng.standard_normal((D, 2, D + 1, 2))))
    with pytest.raises(mk.MKError):
        op(dev(mk, rng.standard_normal((D, 2, D, 2))), precision="bf16x3")   # not prepared
    Wbad = W.copy()
    Wbad[0, :, :, 0] = 2.0 * np.eye(2)        # not a Jordan-form MPO
    with pytest.raises(mk.MKError):
        mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(dev(mk, rng.standard_normal((D, 5, D))), dev(mk, Wbad),
                                                         dev(mk, W), dev(mk, rng.standard_normal((D, 5, D)))))


@pytest.mark.parametrize("prec", PRECS)
def test_eigsolve_lowest_eigenpair(mk, prec):
    """Lanczos on a Hermitian two-site operator of an XXZ chain; FP64 residuals and
    FP64 refinement bring every precision to the float64 eigenvalue."""
    rng = np.random.default_rng(31)
    Lc = 10
    Ws = __import__("mpskit_amd").models.xxz(Lc, 0.8)
    psi = B.random_mps(Lc, 2, 12, rng)
    AL, AR = B.left_canonical(psi), B.right_canonical(psi)
    GL = np.ones((1, 1, 1))
    for j in range(4):
        GL = O.transfer_left_mpo(GL, Ws[j], AL[j], AL[j])
    GR = np.ones((1, 1, 1))
    for j in range(Lc - 1, 5, -1):
        GR = O.transfer_right_mpo(GR, Ws[j], AR[j], AR[j])
    T2 = O.jordan_ac2_prepare(GL, Ws[4], Ws[5], GR)
    shape = (GL.shape[2], 2, GR.shape[0], 2)
    n = int(np.prod(shape))
    Hd = np.column_stack([O.jordan_ac2_apply(T2, np.eye(n)[:, i].reshape(shape, order="F")).reshape(-1, order="F")
                          for i in range(n)])
    assert np.abs(Hd - Hd.T).max() < 1e-12
    e_ref = np.linalg.eigvalsh(Hd)[0]
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(dev(mk, GL), dev(mk, Ws[4]), dev(mk, Ws[5]), dev(mk, GR)),
                             precisions=PRECS)
    x0 = dev(mk, rng.standard_normal(shape))
    print("eigsolve", prec, "n", n, flush=True)
    val, vec, info = op.eigsolve(x0, krylovdim=30, maxiter=50, tol=1e-11, precision=prec)
    print(f"  E={val:.15f} ref={e_ref:.15f} info={info}", flush=True)
    assert info["converged"]
    assert abs(val - e_ref) < 1e-10
    v = host(mk, vec).reshape(-1, order="F")
    r = Hd @ v - e_ref * v
    assert np.linalg.norm(r) < 1e-9
