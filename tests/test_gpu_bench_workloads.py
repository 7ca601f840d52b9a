"""Sweep workloads at their benchmark sizes.

* U(1) XXZ DMRG2, L = 64, D = 2048 kept states over the 21 charge sectors
  [1, 3, 8, 18, 37, 68, 112, 165, 218, 258, 272, 258, 218, 165, 112, 68, 37,
  18, 8, 3, 1] around each bond's mean charge (interior sectors carry zero
  padding states up to a multiple of 8; the 2048 counts real states):
  charge-block apply == dense apply at the middle bond (fp64 1e-12, fp32 1e-5,
  bf16x3 3e-4), and one sweep in float64 and one in the mixed path from the
  same state agree on E/L to 1e-10.
* VUMPS on the infinite J1-J2 chain (J2/J1 = 0.24, unit cell 2, w = 8) at
  D = 1024, 26 iterations in float64 and in the mixed path from the same
  start: energies per site agree to 1e-10.
* 20 TFIM chains (g = 0.1 .. 2.0, L = 48, D_i = 256 .. 1024): batched applies at
  the middle bond equal per-instance applies (1e-12), and lockstep DMRG2 (maximal
  bond dimension D_i, discarded weight cut 1e-14) reaches each chain's
  free-fermion energy to 1e-10 per site.
"""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402

SECTORS = [1, 3, 8, 18, 37, 68, 112, 165, 218, 258, 272, 258, 218, 165, 112, 68, 37, 18, 8, 3, 1]
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


def sector_dims(k, q):
    i = q - (k // 2) + len(SECTORS) // 2
    return SECTORS[i] if 0 <= i < len(SECTORS) else 0


def test_u1_sector_profile_matches_the_bench():
    assert sum(SECTORS) == 2048 and len(SECTORS) == 21


def test_u1_L64_D2048(mk):
    import torch
    from mpskit_amd.u1 import u1_gauge2
    L, nup = 64, 32
    H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, 1.0))
    psi = mk.U1FiniteMPS.random(L, nup, sector_dims, seed=0)
    mid = L // 2 - 1
    print("u1 L=64 bonds", psi.bond_dims()[mid - 1:mid + 3], flush=True)
    # splits capped at 2048 states walk the center to the middle bond (exact: the
    # rank of every two-site center is at most its middle bond's 2048 states)
    for p in range(mid):
        u1_gauge2(psi, p, "right", psi.AC2(p), 2048)
    envs = mk.FiniteEnvironments(psi, H)
    GL, GR = envs.leftenv(mid, psi), envs.rightenv(mid + 1, psi)
    lb, rb = psi.bonds[mid], psi.bonds[mid + 2]
    print("middle bonds", lb.D, rb.D, "sectors", len(lb.q), len(rb.q), flush=True)
    # the right bond keeps the 21 sectors of the start; the split of the left
    # half keeps only the sectors its rank allows
    assert len(rb.q) == 21 and len(lb.q) >= 15
    precs = ("fp64", "fp32", "bf16x3")
    dense = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[mid], H[mid + 1], GR), precisions=precs)
    blk = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[mid], H[mid + 1], GR), precisions=precs)
    mk.set_charge_blocks(blk, lb, rb)
    x = psi.AC2(mid)
    x = x / torch.linalg.vector_norm(x)
    y64 = dense(x)
    for prec in precs:
        e = rel(blk(x, precision=prec), y64)
        print(f"  block vs dense {prec}: {e:.2e}", flush=True)
        assert e < TOL[prec], (prec, e)
    dense.free()
    blk.free()
    del envs, GL, GR
    # back to site 0, then one loose float64 sweep as a common start
    for p in range(mid - 1, -1, -1):
        u1_gauge2(psi, p, "left", psi.AC2(p), 2048)
    print("loose sweep", flush=True)
    mk.find_groundstate_u1(psi, H, trunc_dim=2048, sweeps=1, precision="fp64", tol=1e-5, maxiter=3, verbosity=1)
    res = {}
    for prec in ("fp64", "bf16x3"):
        p2 = copy.deepcopy(psi)
        hist = mk.find_groundstate_u1(p2, H, trunc_dim=2048, sweeps=1, precision=prec, tol=1e-10, verbosity=1)
        res[prec] = hist[-1]["energy"] / L
        del p2
    print("E/L", res, "diff", abs(res["fp64"] - res["bf16x3"]), flush=True)
    assert abs(res["fp64"] - res["bf16x3"]) < 1e-10


def test_vumps_j1j2_D1024_mixed_matches_fp64(mk):
    """26 VUMPS iterations in float64 and in the mixed path (BF16x3 eigensolver
    applies and environment transfers, each refined to the same float64
    tolerance) from one random start.  Near the critical coupling the Galerkin
    error levels off around 1.2e-5 after ~20 iterations while the energy keeps
    converging by ~1e-10 per iteration; the two runs follow each other and
    their energies per site agree to 1e-10 from iteration ~24 on (measured:
    7.8e-11 at 24)."""
    import time
    import torch
    from mpskit_amd import vumps as V
    H = V.InfiniteHamiltonian([mk.models.j1j2_bulk(1.0, 0.24)] * 2)
    res = {}
    for prec in ("fp64", "bf16x3"):
        psi = V.InfiniteMPS.random(2, 2, 1024, seed=7)
        print("vumps j1j2 D=1024", prec, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, eps, hist = V.find_groundstate_vumps(psi, H, V.VUMPS(tol=1e-12, maxiter=26, precision=prec,
                                                                   verbosity=2))
        torch.cuda.synchronize()
        res[prec] = (hist[-1]["energy"], eps, len(hist), time.perf_counter() - t0)
    print(res, "diff", abs(res["fp64"][0] - res["bf16x3"][0]), flush=True)
    assert res["fp64"][2] == res["bf16x3"][2] == 26
    assert res["fp64"][1] <= 2e-5 and res["bf16x3"][1] <= 2e-5
    assert abs(res["fp64"][0] - res["bf16x3"][0]) < 1e-10
    assert -0.4021276 < res["fp64"][0] < -0.4021275


def test_batched_tfim_20_instances(mk):
    import torch
    L, n = 48, 20
    gs = [0.1 * (i + 1) for i in range(n)]
    Ds = [int(round((256 + (768 * i) // 19) / 16.0)) * 16 for i in range(n)]
    assert Ds[0] == 256 and Ds[-1] == 1024
    Hs = [mk.FiniteMPOHamiltonian(mk.models.tfim(L, g)) for g in gs]
    # batched applies and one batched eigensolve at the middle bond of random
    # states at the full D_i, against per-instance applies and solves
    psis = [mk.FiniteMPS.random(L, 2, D, seed=100 + i) for i, D in enumerate(Ds)]
    pos = L // 2 - 1
    ops, xs = [], []
    for psi, H in zip(psis, Hs):
        envs = mk.FiniteEnvironments(psi, H)
        psi.move_center(pos)
        GL, GR = envs.leftenv(pos, psi), envs.rightenv(pos + 1, psi)
        ops.append(mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR),
                                       precisions=("fp64", "bf16x3")))
        x = psi.AC2(pos)
        xs.append(x / torch.linalg.vector_norm(x))
        del envs
    print("batched apply, D_i", [op.x_shape[0] for op in ops], flush=True)
    Bt = mk.BatchedDerivative(ops)
    flat = Bt.pack(xs)
    for prec in ("fp64", "bf16x3"):
        ys = Bt.unpack(Bt.apply(flat, prec))
        for i, (op, x, y) in enumerate(zip(ops, xs, ys)):
            e = rel(y, op(x, precision=prec))
            assert e < (1e-12 if prec == "fp64" else 1e-6), (prec, i, e)
    print("batched eigensolve", flush=True)
    vals, _, infos = Bt.eigsolve(xs, krylovdim=20, maxiter=100, tol=1e-10, precision="bf16x3")
    for i in range(n):
        v1, _, inf1 = ops[i].eigsolve(xs[i], krylovdim=20, maxiter=100, tol=1e-10)
        print(f"  instance {i}: batched {vals[i]:.12f} ({infos[i]['normres']:.1e}) single {v1:.12f} "
              f"({inf1['normres']:.1e})", flush=True)
        assert abs(vals[i] - v1) < 1e-10 * max(1.0, abs(v1))
        assert infos[i]["normres"] < 1e-8
    Bt.free()
    for op in ops:
        op.free()
    del ops, xs, flat, psis
    # lockstep DMRG2 to the closed-form energies: bond caps D_i, discarded weight
    # cut 1e-12, started from D = 16 random states (the bonds grow as needed)
    alg = mk.DMRG2(maxiter=5, trunc_dim=Ds, trunc_err=1e-12, precision="bf16x3")
    psis = [mk.FiniteMPS.random(L, 2, 16, seed=100 + i) for i in range(n)]
    solver = mk.Lanczos(tol=1e-11, krylovdim=20, maxiter=50)
    _, _, hist = mk.find_groundstate_batched(psis, Hs, alg, solver, use_graph=True, verbosity=1)
    errs = [abs(E / L - B.tfim_open_energy(L, g) / L) for E, g in zip(hist[-1]["energies"], gs)]
    print("sweeps", len(hist), "bond dims", [max(p.bond_dims()) for p in psis],
          "errors per site", ["%.1e" % e for e in errs], flush=True)
    assert max(errs) < 1e-10
