"""Batched applies and lockstep eigensolves (grouped persistent GEMM, hipGraph
replay) against the per-operator path.

Operators are two-site Jordan derivatives at a bond of DMRG states of TFIM /
XXZ chains of different lengths, so their sizes are uneven and they are
Hermitian.  Tolerances: fp64 applies 1e-12 relative to the per-operator apply;
mixed applies within the per-precision bound of test_gpu_apply relative to
fp64; eigenvalues 1e-10 relative (float64 residual refinement in every mode).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 3e-4}
PRECS = ("fp64", "fp32", "bf16x3")


@pytest.fixture(scope="module")
def mk():
    import torch
    assert torch.cuda.is_available()
    import mpskit_amd
    mpskit_amd.load()
    return mpskit_amd


def bond_operator(mk, Ws, L, D, seed, pos=None):
    H = mk.FiniteMPOHamiltonian(Ws)
    psi = mk.FiniteMPS.random(L, 2, D, seed=seed)
    envs = mk.FiniteEnvironments(psi, H)
    pos = L // 2 - 1 if pos is None else pos
    psi.move_center(pos)
    GL, GR = envs.leftenv(pos, psi), envs.rightenv(pos + 1, psi)
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, H[pos], H[pos + 1], GR), precisions=PRECS)
    return op, psi.AC2(pos)


@pytest.fixture(scope="module")
def instances(mk):
    specs = [(mk.models.tfim(10, 0.5), 10, 32, 1), (mk.models.tfim(14, 1.0), 14, 96, 2),
             (mk.models.xxz(12, 0.7), 12, 64, 3), (mk.models.tfim(16, 1.5), 16, 200, 4),
             (mk.models.xxz(8, 1.0), 8, 16, 5)]
    return [bond_operator(mk, *s) for s in specs]


def rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def test_batch_apply_matches_single(mk, instances):
    import torch
    ops = [op for op, _ in instances]
    xs = [x / torch.linalg.vector_norm(x) for _, x in instances]
    B = mk.BatchedDerivative(ops)
    assert B.grouped
    flat = B.pack(xs)
    assert torch.equal(B.unpack(flat)[2], xs[2])
    y64 = [op(x) for op, x in zip(ops, xs)]
    for prec in PRECS:
        print("batch apply", prec, flush=True)
        ys = B.unpack(B.apply(flat, prec))
        for i, (y, yr) in enumerate(zip(ys, y64)):
            e = rel(y, yr)
            assert e < TOL[prec], (prec, i, e)
    # padding between instances stays zero
    y = B.apply(flat, "fp64")
    mask = torch.ones_like(y, dtype=torch.bool)
    for o, op in zip(B.offsets, ops):
        mask[o:o + int(np.prod(op.x_shape))] = False
    assert float(y[mask].abs().max()) == 0.0
    B.free()


def test_batch_eigsolve_matches_single(mk, instances):
    import torch
    ops = [op for op, _ in instances]
    xs = [x.clone() for _, x in instances]
    ref = [op.eigsolve(x, krylovdim=20, maxiter=200, tol=1e-12)[0] for op, x in zip(ops, xs)]
    B = mk.BatchedDerivative(ops)
    for prec in PRECS:
        print("batch eigsolve", prec, flush=True)
        ev, vecs, infos = B.eigsolve(xs, krylovdim=20, maxiter=200, tol=1e-12, precision=prec, use_graph=False)
        for i in range(len(ops)):
            assert infos[i]["converged"], (prec, i, infos[i])
            assert abs(ev[i] - ref[i]) <= 1e-10 * abs(ref[i]), (prec, i, ev[i], ref[i])
            # the returned vector is a normalised eigenvector (float64 residual)
            v = vecs[i]
            r = ops[i](v) - ev[i] * v
            assert float(torch.linalg.vector_norm(r)) < 1e-9, (prec, i)
    B.free()


def test_batch_graph_replay_matches_eager(mk, instances):
    """Captured Krylov steps replayed across calls, with eager work in between,
    give the eager results."""
    import torch
    ops = [op for op, _ in instances]
    xs = [x.clone() for _, x in instances]
    B = mk.BatchedDerivative(ops)
    for prec in ("fp64", "bf16x3"):
        e_eager, _, _ = B.eigsolve(xs, krylovdim=12, maxiter=100, tol=1e-11, precision=prec, use_graph=False)
        for rep in range(3):
            print("graph replay", prec, rep, flush=True)
            e_g, v_g, inf = B.eigsolve(xs, krylovdim=12, maxiter=100, tol=1e-11, precision=prec, use_graph=True)
            for i in range(len(ops)):
                assert inf[i]["converged"]
                assert abs(e_g[i] - e_eager[i]) <= 1e-10 * abs(e_eager[i]), (prec, rep, i)
            _ = [op(x) for op, x in zip(ops, xs)]   # unrelated eager work between replays
            torch.cuda.synchronize()
    B.free()


def test_batch_fallback_non_grouped(mk, instances):
    """A batch with a J1-J2 operator (CA / AB terms) is applied operator by operator."""
    import torch
    op_j, x_j = bond_operator(mk, mk.models.j1j2(10, 1.0, 0.24), 10, 32, 7)
    ops = [instances[0][0], op_j]
    xs = [instances[0][1], x_j]
    B = mk.BatchedDerivative(ops)
    assert not B.grouped
    ys = B.unpack(B.apply(B.pack(xs), "fp64"))
    for op, x, y in zip(ops, xs, ys):
        assert rel(y, op(x)) < 1e-13
    ev, _, infos = B.eigsolve(xs, krylovdim=16, maxiter=100, tol=1e-11)
    ref = op_j.eigsolve(x_j, krylovdim=16, maxiter=100, tol=1e-11)[0]
    assert abs(ev[1] - ref) <= 1e-10 * abs(ref)
    B.free()


def test_autotuner_decides_and_caches(mk, instances):
    d0, h0 = mk.autotune_stats()
    ops = [op for op, _ in instances]
    B1 = mk.BatchedDerivative(ops)
    B1.apply(B1.pack([x for _, x in instances]), "bf16x3")
    B2 = mk.BatchedDerivative(ops)
    B2.apply(B2.pack([x for _, x in instances]), "bf16x3")
    d1, h1 = mk.autotune_stats()
    assert d1 >= 1 and h1 > h0
    B1.free()
    B2.free()


def test_batched_dmrg_matches_individual_sweeps(mk):
    """Lockstep DMRG2 of three TFIM couplings (grouped launches, captured Krylov
    steps) reaches the closed-form energies of the open chain (1e-10 per site)."""
    from oracle import brute as Bf
    L, D = 12, 64
    gs = (0.5, 1.0, 1.5)
    psis = [mk.FiniteMPS.random(L, 2, D, seed=i) for i in range(3)]
    Hs = [mk.FiniteMPOHamiltonian(mk.models.tfim(L, g)) for g in gs]
    alg = mk.DMRG2(maxiter=5, trunc_dim=D)
    _, _, hist = mk.find_groundstate_batched(psis, Hs, alg, mk.Lanczos(tol=1e-12, krylovdim=20, maxiter=100))
    for g, E in zip(gs, hist[-1]["energies"]):
        assert abs(E - Bf.tfim_open_energy(L, g)) / L < 1e-10, (g, E)

