"""The drop-in seam on the GPU: one-argument prepare of constructed Jordan
operators, and the concurrency contract.

* ``mk_prepare_ac2_jordan_terms`` / ``mk_prepare_ac_jordan_terms`` take the
  fields of a constructed ``JordanMPO_AC2_Hamiltonian`` / ``JordanMPO_AC_Hamiltonian``
  (the oracle's ``jordan_ac2_construct`` / ``jordan_ac_construct``, the
  reference's constructor, hamiltonian_derivatives.jl:57-107, :272-369) and
  fold them as ``prepare_operator!!`` does (:109-157, :371-468).  Their applies
  match the oracle (fp64 1e-12, fp32 1e-5, bf16x3 3e-4) on every bond of the
  test chains, boundaries included, including the CA / AB folds (J1-J2), the AA
  term (range 3) and terms the reference keeps apart when their fold target is
  missing.
* Two host threads, each with its own context and stream, interleave prepares,
  applies (fp64 and bf16x3), transfers and eigensolves; every result equals a
  serial run bit for bit (the reference calls these operations from concurrent
  tasks, vumps.jl:142, infinite_envs.jl:95-109, and tests its shared allocator
  for it, test/misc/allocator.jl:105-129).
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import brute as B  # noqa: E402
from oracle import oracle as O  # noqa: E402

TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 3e-4}
PRECS = ("fp64", "fp32", "bf16x3")
L = 7


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


def chain(mk, name):
    M = mk.models
    return {"tfim": M.tfim(L, 0.7), "xxz": M.xxz(L, 0.6), "j1j2": M.j1j2(L, 1.0, 0.3),
            "heis3": M.heisenberg_range(L, [1.0, 0.45, 0.3]), "expising": M.exp_ising(L, 0.8, 0.6, 0.9)}[name]


def envs(Ws, psi):
    AL, AR = B.left_canonical(psi), B.right_canonical(psi)
    GL = [np.ones((1, 1, 1))]
    for j in range(len(Ws)):
        GL.append(O.transfer_left_mpo(GL[-1], Ws[j], AL[j], AL[j]))
    GR = [None] * (len(Ws) + 1)
    GR[-1] = np.ones((1, 1, 1))
    for j in range(len(Ws) - 1, -1, -1):
        GR[j] = O.transfer_right_mpo(GR[j + 1], Ws[j], AR[j], AR[j])
    return GL, GR


def terms_op(mk, T0, dims, two_site=True):
    dev = mk.tensors.from_numpy
    t = {k: (None if v is None else dev(v)) for k, v in T0.items() if k not in ("AA", "A")}
    if two_site:
        AA = None if T0["AA"] is None else tuple(dev(a) for a in T0["AA"])
        H = mk.JordanMPO_AC2_Terms(dims, AA=AA, **t)
    else:
        A = None if T0["A"] is None else tuple(dev(a) for a in T0["A"])
        H = mk.JordanMPO_AC_Terms(dims, A=A, **t)
    return mk.prepare_operator(H, precisions=PRECS)


@pytest.mark.parametrize("name", ["tfim", "xxz", "j1j2", "heis3", "expising"])
def test_prepare_from_constructed_terms(mk, name):
    rng = np.random.default_rng(41)
    Ws = chain(mk, name)
    psi = B.random_mps(L, 2, 6, rng)
    GL, GR = envs(Ws, psi)
    print("terms", name, flush=True)
    for j in range(L - 1):
        T0 = O.jordan_ac2_construct(GL[j], Ws[j], Ws[j + 1], GR[j + 2])
        Dl, Dr = GL[j].shape[2], GR[j + 2].shape[0]
        x = rng.standard_normal((Dl, 2, Dr, 2))
        y_ref = O.jordan_ac2_apply(O.jordan_ac2_fold(T0), x)
        op = terms_op(mk, T0, (Dl, 2, Dr, 2))
        xd = mk.tensors.from_numpy(x)
        for p in PRECS:
            e = rel(mk.tensors.to_numpy(op(xd, precision=p)), y_ref)
            assert e < TOL[p], (name, j, p, e)
        op.free()
    for j in range(L):
        T0 = O.jordan_ac_construct(GL[j], Ws[j], GR[j + 1])
        Dl, Dr = GL[j].shape[2], GR[j + 1].shape[0]
        x = rng.standard_normal((Dl, 2, Dr))
        y_ref = O.jordan_ac_apply(O.jordan_ac_fold(T0), x)
        op = terms_op(mk, T0, (Dl, 2, Dr), two_site=False)
        xd = mk.tensors.from_numpy(x)
        for p in PRECS:
            e = rel(mk.tensors.to_numpy(op(xd, precision=p)), y_ref)
            assert e < TOL[p], (name, "ac", j, p, e)
        op.free()


def test_terms_without_their_fold_target(mk):
    """Terms that the reference keeps apart when the term they fold into is
    missing (II / ID without IC or CA, DE / EE without BE or AB, D without C
    or B, I without C, E without B): the unfolded oracle apply runs them on
    their own; the library adds them into an operand of their own."""
    rng = np.random.default_rng(43)
    W = mk.models.tfim_bulk(0.9)
    D = 37
    GL, GR = rng.standard_normal((D, 3, D)), rng.standard_normal((D, 3, D))
    T0 = O.jordan_ac2_construct(GL, W, W, GR)
    x = rng.standard_normal((D, 2, D, 2))
    for drop in (("IC",), ("BE",), ("IC", "BE"), ("IC", "CB")):
        Tm = dict(T0, **{k: None for k in drop})
        y_ref = O.jordan_ac2_apply(O.jordan_ac2_fold(Tm), x)
        assert rel(y_ref, O.jordan_ac2_apply(Tm, x)) < 1e-13
        op = terms_op(mk, Tm, (D, 2, D, 2))
        for p in PRECS:
            e = rel(mk.tensors.to_numpy(op(mk.tensors.from_numpy(x), precision=p)), y_ref)
            assert e < TOL[p], (drop, p, e)
        op.free()
    T1 = O.jordan_ac_construct(GL, W, GR)
    x1 = rng.standard_normal((D, 2, D))
    for drop in (("C",), ("B",), ("C", "B")):
        Tm = dict(T1, **{k: None for k in drop})
        y_ref = O.jordan_ac_apply(O.jordan_ac_fold(Tm), x1)
        op = terms_op(mk, Tm, (D, 2, D), two_site=False)
        for p in PRECS:
            e = rel(mk.tensors.to_numpy(op(mk.tensors.from_numpy(x1), precision=p)), y_ref)
            assert e < TOL[p], (drop, p, e)
        op.free()


def _workload(mk, seed, out, barrier=None):
    """prepare -> applies -> transfers -> eigensolve, repeated; results to `out`."""
    import torch
    T = mk.tensors
    rng = np.random.default_rng(seed)
    W = mk.models.xxz_bulk(0.8)
    for rep in range(3):
        D = 96 + 32 * rep + seed
        GL, GR = rng.standard_normal((D, 5, D)) / D, rng.standard_normal((D, 5, D)) / D
        GL = GL + GL.transpose(2, 1, 0)
        GR = GR + GR.transpose(2, 1, 0)
        A = rng.standard_normal((D, 2, D)) / D
        x = rng.standard_normal((D, 2, D, 2))
        if barrier is not None:
            barrier.wait()
        op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(T.from_numpy(GL), T.from_numpy(W), T.from_numpy(W),
                                                              T.from_numpy(GR)), precisions=("fp64", "bf16x3"))
        xd = T.from_numpy(x)
        y64 = op(xd, precision="fp64")
        ybf = op(xd, precision="bf16x3")
        tl = mk.transfer_left(T.from_numpy(GL), T.from_numpy(W), T.from_numpy(A), T.from_numpy(A))
        tr = mk.transfer_right(T.from_numpy(GR), T.from_numpy(W), T.from_numpy(A), T.from_numpy(A))
        val, vec, _ = op.eigsolve(xd, krylovdim=12, maxiter=3, tol=1e-14, precision="bf16x3")
        torch.cuda.current_stream().synchronize()
        out.append([T.to_numpy(t) for t in (y64, ybf, tl, tr, vec)] + [val])
        op.free()


def test_two_threads_two_contexts_match_serial(mk):
    import torch
    serial = {0: [], 1: []}
    for s in (0, 1):
        _workload(mk, s, serial[s])
    par = {0: [], 1: []}
    errors = []
    barrier = threading.Barrier(2)

    def run(s):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                _workload(mk, s, par[s], barrier)
        except Exception as e:            # surfaced below
            errors.append(e)
            barrier.abort()

    ths = [threading.Thread(target=run, args=(s,)) for s in (0, 1)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ths), "a worker thread did not finish"
    assert not errors, errors
    for s in (0, 1):
        assert len(par[s]) == len(serial[s]) == 3
        for rs, rp in zip(serial[s], par[s]):
            for a, b in zip(rs[:-1], rp[:-1]):
                assert np.array_equal(a, b)
            assert rs[-1] == rp[-1]
