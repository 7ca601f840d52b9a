"""Generate the seeded golden vectors in tests/golden/apply_vectors.npz.

Inputs are i.i.d. Gaussian (numpy PCG64, fixed seeds) or built from random
mixed-canonical MPS of the TFIM (w=3), XXZ (w=5) and J1-J2 (w=8) chains;
outputs come from the float64 oracle (oracle/oracle.py), whose every routine is
itself checked against brute-force dense Hamiltonians in tests/test_oracle.py.
The reference ships no vectors for these applies (its tests only compare
relative quantities), so these fixtures are this project's own.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "mpskit.jl_amd")]

from oracle import oracle as O  # noqa: E402
from oracle import brute as B  # noqa: E402
from mpskit_amd import models as M  # noqa: E402

BULK = {"tfim": M.tfim_bulk(0.7), "xxz": M.xxz_bulk(0.6), "j1j2": M.j1j2_bulk(1.0, 0.24),
        # connected A blocks: the two-site Jordan operator carries its AA term
        "heis3": M.heisenberg_range_bulk([1.0, 0.45, 0.3]), "expising": M.exp_ising_bulk(0.8, 0.6, 0.9)}
CHAIN = {"tfim": lambda L: M.tfim(L, 0.7), "xxz": lambda L: M.xxz(L, 0.6), "j1j2": lambda L: M.j1j2(L, 1.0, 0.24),
         "heis3": lambda L: M.heisenberg_range(L, [1.0, 0.45, 0.3]), "expising": lambda L: M.exp_ising(L, 0.8, 0.6, 0.9)}


def unit(rng, *shape):
    a = rng.standard_normal(shape)
    return a / np.linalg.norm(a)


def cases():
    out = {}
    for mi, (name, W) in enumerate(BULK.items()):
        w = W.shape[0]
        for D in (8, 16):
            rng = np.random.default_rng(1000 * mi + D)
            # i.i.d. environments (the reference AC2 benchmark's setup), bra != ket dims
            Da, Df, Di, Dc = D, D + 1, D + 2, D - 1
            GL, GR = unit(rng, Da, w, Df), unit(rng, Di, w, Dc)
            x2 = unit(rng, Df, 2, Di, 2)
            x1 = unit(rng, Df, 2, Di)
            x0 = unit(rng, Df, Di)
            k = f"{name}_D{D}"
            out[f"ac2gen_{k}"] = dict(GL=GL, W1=W, W2=W, GR=GR, x=x2, y=O.ac2_apply_generic(GL, W, W, GR, x2))
            out[f"acgen_{k}"] = dict(GL=GL, W=W, GR=GR, x=x1, y=O.ac_apply_generic(GL, W, GR, x1))
            out[f"c_{k}"] = dict(GL=GL, GR=GR, x=x0, y=O.c_apply(GL, GR, x0))
            # Jordan operators on i.i.d. square environments (GL "not started" and
            # GR "finished" levels are ignored, as in the reference)
            GLs, GRs = unit(rng, D, w, D), unit(rng, D, w, D)
            xs2, xs1 = unit(rng, D, 2, D, 2), unit(rng, D, 2, D)
            T2 = O.jordan_ac2_prepare(GLs, W, W, GRs)
            T1 = O.jordan_ac_prepare(GLs, W, GRs)
            out[f"ac2jor_{k}"] = dict(GL=GLs, W1=W, W2=W, GR=GRs, x=xs2, y=O.jordan_ac2_apply(T2, xs2))
            out[f"acjor_{k}"] = dict(GL=GLs, W=W, GR=GRs, x=xs1, y=O.jordan_ac_apply(T1, xs1))
            # transfers with distinct bra / ket tensors
            A, Ab = unit(rng, Df, 2, Dc), unit(rng, Da, 2, Di)
            v = unit(rng, Da, Df)
            out[f"tlmps_{k}"] = dict(v=v, A=A, Ab=Ab, y=O.transfer_left_mps(v, A, Ab))
            vr = unit(rng, Dc, Di)
            out[f"trmps_{k}"] = dict(v=vr, A=A, Ab=Ab, y=O.transfer_right_mps(vr, A, Ab))
            xl = unit(rng, Da, w, Df)
            out[f"tlmpo_{k}"] = dict(x=xl, O=W, A=A, Ab=Ab, y=O.transfer_left_mpo(xl, W, A, Ab))
            vr3 = unit(rng, Dc, w, Di)
            out[f"trmpo_{k}"] = dict(v=vr3, O=W, A=A, Ab=Ab, y=O.transfer_right_mpo(vr3, W, A, Ab))
        # a boundary pair: first two sites of a finite chain (w_l = 1)
        Ws = CHAIN[name](6)
        rng = np.random.default_rng(77 + mi)
        psi = B.random_mps(6, 2, 8, rng)
        AL, AR = B.left_canonical(psi), B.right_canonical(psi)
        GLs = [np.ones((1, 1, 1))]
        for j in range(6):
            GLs.append(O.transfer_left_mpo(GLs[-1], Ws[j], AL[j], AL[j]))
        GRs = [None] * 7
        GRs[6] = np.ones((1, 1, 1))
        for j in range(5, -1, -1):
            GRs[j] = O.transfer_right_mpo(GRs[j + 1], Ws[j], AR[j], AR[j])
        for j in (0, 2, 4):
            x = unit(rng, GLs[j].shape[2], 2, GRs[j + 2].shape[0], 2)
            T2 = O.jordan_ac2_prepare(GLs[j], Ws[j], Ws[j + 1], GRs[j + 2])
            out[f"ac2chain_{name}_site{j}"] = dict(GL=GLs[j], W1=Ws[j], W2=Ws[j + 1], GR=GRs[j + 2], x=x,
                                                   y=O.jordan_ac2_apply(T2, x))
            xa = unit(rng, GLs[j].shape[2], 2, GRs[j + 1].shape[0])
            T1 = O.jordan_ac_prepare(GLs[j], Ws[j], GRs[j + 1])
            out[f"acchain_{name}_site{j}"] = dict(GL=GLs[j], W=Ws[j], GR=GRs[j + 1], x=xa, y=O.jordan_ac_apply(T1, xa))
    return out


def main():
    flat = {}
    for case, arrs in cases().items():
        for k, v in arrs.items():
            flat[f"{case}/{k}"] = np.asarray(v, dtype=np.float64)
    path = Path(__file__).with_name("apply_vectors.npz")
    np.savez_compressed(path, **flat)
    print(path, len(flat), "arrays", path.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
