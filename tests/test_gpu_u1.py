"""U(1) charge-block path: block applies equal the dense apply on
charge-conserving vectors (fp64 1e-12, mixed within the apply bounds), on the
dense charge-sorted layout and on compact block storage (with alpha / beta and
eigensolves), and U(1) DMRG2 of the XXZ chain reaches the exact Sz = 0
ground-state energy (1e-10 per site), in float64 and in the mixed path, with
either storage."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402

TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 3e-4}


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def test_block_apply_matches_dense(mk):
    import torch
    L, nup = 12, 6
    psi = mk.U1FiniteMPS.random(L, nup, lambda k, q: 20, seed=3)
    H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, 0.8))
    envs = mk.FiniteEnvironments(psi, H)
    pos = 0
    # walk the center to the middle with exact two-site splits
    from mpskit_amd.u1 import u1_gauge2
    for p in range(L // 2 - 1):
        u1_gauge2(psi, p, "right", psi.AC2(p), 10 ** 6)
    pos = L // 2 - 1
    GL, GR = envs.leftenv(pos, psi), envs.rightenv(pos + 1, psi)
    precs = ("fp64", "fp32", "bf16x3")
    dense = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=precs)
    blk = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=precs)
    mk.set_charge_blocks(blk, psi.bonds[pos], psi.bonds[pos + 2])
    assert blk.flops < 0.5 * dense.flops
    x = psi.AC2(pos)
    y64 = dense(x)
    for prec in precs:
        yb = blk(x, precision=prec)
        assert rel(yb, y64) < TOL[prec], prec
    # through a batch (grouped block problems, captured Krylov steps)
    Bt = mk.BatchedDerivative([blk])
    yb = Bt.unpack(Bt.apply(Bt.pack([x]), "fp64"))[0]
    assert rel(yb, y64) < 1e-12
    # alpha / beta
    z = torch.randn_like(x)
    out = z.clone()
    blk(x, precision="fp64", out=out, alpha=0.5, beta=-2.0)
    assert rel(out, 0.5 * y64 - 2.0 * z) < 1e-12
    Bt.free()
    # compact storage: the vectors hold only the charge blocks
    cmp = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=precs)
    mk.set_charge_blocks(cmp, psi.bonds[pos], psi.bonds[pos + 2])
    from mpskit_amd.u1 import compact_charge_blocks
    lay = compact_charge_blocks(cmp, x.shape)
    assert lay.n < 0.5 * x.numel() and cmp.x_shape == (lay.n,)
    # x is charge conserving: gather then scatter returns it
    assert rel(lay.scatter(lay.gather(x)), x) == 0.0
    xc = lay.gather(x)
    for prec in precs:
        yc = lay.scatter(cmp(xc, precision=prec))
        assert rel(yc, y64) < TOL[prec], ("compact", prec)
    zc = torch.randn_like(xc)
    outc = zc.clone()
    cmp(xc, precision="bf16x3", out=outc, alpha=0.5, beta=-2.0)
    assert rel(lay.scatter(outc), 0.5 * y64 - 2.0 * lay.scatter(zc)) < TOL["bf16x3"]
    e64 = cmp.eigsolve(xc, krylovdim=20, maxiter=100, tol=1e-11)[0]
    ed = dense.eigsolve(x, krylovdim=20, maxiter=100, tol=1e-11)[0]
    eb = cmp.eigsolve(xc, krylovdim=20, maxiter=100, tol=1e-11, precision="bf16x3")[0]
    # the dense problem's lowest state may carry another charge: it is not below
    assert abs(e64 - eb) < 1e-10 * abs(e64) and ed <= e64 + 1e-10 * abs(e64)
    cmp.free()


def test_u1_dmrg_xxz_vs_exact(mk):
    L, nup, delta = 10, 5, 1.0
    e_ed = np.linalg.eigvalsh(B.dense_xxz(L, delta).toarray())[0]
    psi = mk.U1FiniteMPS.random(L, nup, lambda k, q: 16, seed=1)
    H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, delta))
    hist = mk.find_groundstate_u1(psi, H, trunc_dim=64, sweeps=4, tol=1e-12, verbosity=1)
    assert abs(hist[-1]["energy"] - e_ed) / L < 1e-10, (hist[-1]["energy"], e_ed)


def test_u1_dmrg_mixed_matches_fp64(mk):
    L, nup = 16, 8
    res = {}
    for prec in ("fp64", "bf16x3"):
        for storage in ("compact", "dense"):
            psi = mk.U1FiniteMPS.random(L, nup, lambda k, q: 16, seed=2)
            H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, 1.0))
            res[prec, storage] = mk.find_groundstate_u1(psi, H, trunc_dim=96, sweeps=4, precision=prec, tol=1e-11,
                                                        storage=storage)[-1]["energy"]
    vals = list(res.values())
    assert max(vals) - min(vals) < 1e-10 * L, res
