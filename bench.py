"""Benchmark: isolated two-site derivative applies at D=2048 on one MI355X.

Workload (the reference's AC2 contraction benchmark, AC2_benchmarks.jl:24-72,
at the isolated-apply configuration): nearest-neighbour Jordan dAC2 of the
Heisenberg XXZ chain (w=5, d=2), dense float64 GL/GR of shape D x 5 x D and
x of shape D x 2 x D x 2 with i.i.d. Gaussian entries scaled to unit norm
(seeds 1, 2, 4); W has random Jordan C/B blocks at the (1,2)/(2,3) levels and
identity corners (seed 3).  A step is one apply y = H(x) with all inputs
resident in HBM.

Prints one JSON line: value = applies/s of the mixed (split-BF16x3) path, with
the FP64 and FP32 GPU paths, the speed-up over FP64, the relative error of
each against the FP64 result, the GEMM roofline measured with HIP events, and
the float64 numpy/OpenBLAS restatement of the same apply timed on host cores.

The same line carries the DMRG2 sweep time at D=2048: one sweep of the
Heisenberg chain (L=64, 43 of its 63 bonds at 2048, truncating from 4096;
random normalised start, seed 0) in FP64 and in the mixed path, with the
energies of both (--sweep-L sets the chain length, --no-sweep skips it).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--D 2048] [--no-cpu] [--no-sweep]

Other workloads print their own JSON line:
    --workload batched   20 TFIM chains (g = 0.1 .. 2.0, L = 48, D_i = 256 .. 1024)
                         swept in lockstep with grouped launches; learned
                         Davidson preconditioner against plain Lanczos
    --workload u1        U(1) XXZ DMRG2, L = 64, D = 2048 over 21 charge sectors
                         (charge-block grouped GEMMs, captured Krylov steps)
    --workload vumps     VUMPS on the J1-J2 chain, cell 2, D = 1024
    --workload boundary  leading_boundary (VUMPS) of the 2D Ising transfer MPO, D = 1024
    --workload tdvp      imaginary-time TDVP2 on the Heisenberg chain, L = 24, D = 512
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "mpskit.jl_amd")]

PEAK_TF = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2516.6 / 3.0}
KERNEL = {"fp64": "gemm_f64", "fp32": "gemm_f32", "bf16x3": "gemm_bf16x3"}


def jordan_w(rng, d=2, nmid=3):
    """Jordan MPO tensor with random C (start) and B (end) blocks, identity corners."""
    w = nmid + 2
    W = np.zeros((w, d, d, w))
    W[0, :, :, 0] = np.eye(d)
    W[-1, :, :, -1] = np.eye(d)
    W[0, :, :, 1:-1] = rng.standard_normal((d, d, nmid))
    W[1:-1, :, :, -1] = rng.standard_normal((nmid, d, d))
    return W


def gpu_unit(T, shape, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    t = T.empty(shape)
    t.copy_(torch.randn(tuple(shape), dtype=torch.float64, device="cuda", generator=g))
    t.div_(torch.linalg.vector_norm(t))
    return t


def time_applies(op, x, y, prec, steps, warmup):
    import torch
    for _ in range(warmup):
        op(x, precision=prec, out=y)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        op(x, precision=prec, out=y)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def vendor_gemm_reference(D, d=2, reps=20):
    """torch.matmul (hipBLASLt / rocBLAS) on the GEMM shape of one apply: the
    two-segment product y = x IC + BE x is one M = N = dD, K = 2 dD GEMM.
    bf16 is timed once and reported x3 (the three products of the split
    scheme); float64 once.  Context for the roofline fractions, not a path of
    the library."""
    import torch
    n = d * D
    out = {}
    for name, dt in (("bf16", torch.bfloat16), ("fp64", torch.float64)):
        A = torch.randn(n, 2 * n, dtype=dt, device="cuda")
        B = torch.randn(2 * n, n, dtype=dt, device="cuda")
        for _ in range(3):
            C = A @ B
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            C = A @ B
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = dict(ms=ms, tflops=2.0 * n * n * 2 * n / (ms * 1e-3) / 1e12)
        del A, B, C
    out["bf16_x3_ms"] = 3 * out["bf16"]["ms"]
    out["shape"] = f"M=N={n}, K={2 * n}"
    return out


def apply_table(mk, T, steps=5):
    """applies/s of the other operators and sizes of the isolated-apply workload
    beside the headline: the Jordan dAC2
    at D = 512, 1024, 4096, and at D = 2048 the generic MPO dAC2 (w = 5, the
    same W taken as a dense MPO tensor) and the Jordan one-site dAC, in FP64,
    FP32 and BF16x3 (inputs as in the headline workload, other seeds)."""
    import torch
    W = T.from_numpy(jordan_w(np.random.default_rng(3)))
    cases = [("jordan_dAC2", D) for D in (512, 1024, 4096)] + [("generic_dAC2", 2048), ("jordan_dAC", 2048)]
    out = {}
    for kind, D in cases:
        GL, GR = gpu_unit(T, (D, 5, D), 11), gpu_unit(T, (D, 5, D), 12)
        if kind == "jordan_dAC2":
            H = mk.JordanMPO_AC2_Hamiltonian(GL, W, W, GR)
            shape = (D, 2, D, 2)
        elif kind == "generic_dAC2":
            H = mk.MPO_AC2_Hamiltonian(GL, W, W, GR)
            shape = (D, 2, D, 2)
        else:
            H = mk.JordanMPO_AC_Hamiltonian(GL, W, GR)
            shape = (D, 2, D)
        op = mk.prepare_operator(H, precisions=("fp64", "fp32", "bf16x3"))
        del GL, GR, H
        x = gpu_unit(T, shape, 13)
        y = T.empty(shape)
        row = {}
        y64 = None
        for prec in ("fp64", "fp32", "bf16x3"):
            ms = time_applies(op, x, y, prec, steps if D < 4096 else 3, 1)
            if prec == "fp64":
                y64 = y.clone()
            row[prec] = {"applies_per_s": 1e3 / ms, "tflops": op.flops / (ms * 1e-3) / 1e12}
            if y64 is not None and prec != "fp64":
                row[prec]["rel_err_vs_fp64"] = float(torch.linalg.vector_norm(y - y64) / torch.linalg.vector_norm(y64))
        row["bf16x3_speedup_vs_fp64"] = row["bf16x3"]["applies_per_s"] / row["fp64"]["applies_per_s"]
        out[f"{kind} D={D}"] = row
        print(f"# {kind} D={D}: " + ", ".join(f"{p} {row[p]['applies_per_s']:.1f}/s" for p in ("fp64", "fp32", "bf16x3")),
              file=sys.stderr, flush=True)
        op.free()
        del x, y, y64
        torch.cuda.empty_cache()
    return out


def qp_transfer_table(mk, T, D=1024, De=2, reps=10):
    """Times of the quasiparticle and excitation-leg MPO transfers (transfer.jl:133-148)
    beside the plain MPO transfer (transfer.jl:125-130), D = 1024, w = 5, d = 2, an
    excitation leg of De = 2, FP64.  Each variant is the plain transfer's two GEMMs with
    one bond or MPO leg De times wider, so its algorithmic flops are De times the plain
    transfer's 2 D^3 d (wl + wr); the rate shows what the reshapes, the leg swap of the
    excited right transfer and the widened MPO step cost on top."""
    import torch
    rng = np.random.default_rng(21)
    w, d = 5, 2
    W = T.from_numpy(jordan_w(rng))
    x = gpu_unit(T, (D, w, D), 31)
    x4 = gpu_unit(T, (D, w, De, D), 32)
    A, Ab = gpu_unit(T, (D, d, D), 33), gpu_unit(T, (D, d, D), 34)
    B = gpu_unit(T, (D, d, De, D), 35)
    base = 2.0 * D ** 3 * d * (2 * w)
    cases = {
        "transfer_left(x, O, A, Ab)": (lambda: mk.transfer_left(x, W, A, Ab), base),
        "transfer_right(v, O, A, Ab)": (lambda: mk.transfer_right(x, W, A, Ab), base),
        "transfer_left(x, O, B, Ab), excited A": (lambda: mk.transfer_left(x, W, B, Ab), De * base),
        "transfer_right(v, O, B, Ab), excited A": (lambda: mk.transfer_right(x, W, B, Ab), De * base),
        "transfer_left(x4, O, A, Ab), excitation leg": (lambda: mk.transfer_left(x4, W, A, Ab), De * base),
        "transfer_right(v4, O, A, Ab), excitation leg": (lambda: mk.transfer_right(x4, W, A, Ab), De * base),
    }
    out = {"config": f"D={D}, w={w}, d={d}, De={De}, float64"}
    for name, (fn, flops) in cases.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = dict(ms=ms, tflops=flops / (ms * 1e-3) / 1e12)
        print(f"# {name}: {ms:.3f} ms, {out[name]['tflops']:.1f} TFLOP/s", file=sys.stderr, flush=True)
    del x, x4, A, Ab, B
    torch.cuda.empty_cache()
    return out


def batched_apply(mk, T, reps=10):
    """The grouped (batched) dAC2 path: one mk_batch_apply applies every
    instance of a batch in one grouped persistent GEMM launch plus one memory-
    bound prologue.  Two batches of XXZ-like Jordan operators (random C/B
    blocks, unit-norm random GL/GR, as in the isolated apply): 4 instances at
    D = 2048 and 20 instances at the parameter-sweep sizes D_i = 256..1024.
    Returns instance applies/s per precision and the BF16x3 / FP64 ratio."""
    import torch
    out = {}
    cases = {"4 x D=2048": [2048] * 4,
             "20 x D=256..1024": [256 + (768 * i // 19) // 16 * 16 for i in range(20)]}
    for name, Ds in cases.items():
        ops = []
        for i, D in enumerate(Ds):
            W = T.from_numpy(jordan_w(np.random.default_rng(100 + i)))
            GL = gpu_unit(T, (D, 5, D), 200 + i)
            GR = gpu_unit(T, (D, 5, D), 300 + i)
            ops.append(mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, W, W, GR),
                                           precisions=("fp64", "bf16x3")))
            del GL, GR
        B = mk.BatchedDerivative(ops)
        xs = [gpu_unit(T, (D, 2, D, 2), 400 + i) for i, D in enumerate(Ds)]
        flat = B.pack(xs)
        res = {}
        y64 = None
        for prec in ("fp64", "bf16x3"):
            y = B.apply(flat, prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                y = B.apply(flat, prec)
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / reps
            if prec == "fp64":
                y64 = y
            err = float(torch.linalg.vector_norm(y - y64) / torch.linalg.vector_norm(y64))
            res[prec] = dict(ms_per_batch=sec * 1e3, instance_applies_per_s=len(Ds) / sec, rel_err_vs_fp64=err)
        res["speedup_vs_fp64"] = res["fp64"]["ms_per_batch"] / res["bf16x3"]["ms_per_batch"]
        res["grouped"] = B.grouped
        out[name] = res
        print(f"# batched {name}: fp64 {res['fp64']['ms_per_batch']:.3f} ms, bf16x3 "
              f"{res['bf16x3']['ms_per_batch']:.3f} ms per batch ({res['speedup_vs_fp64']:.2f}x)",
              file=sys.stderr, flush=True)
        B.free()
        for op in ops:
            op.free()
        del xs, flat, y, y64
    return out


def cpu_baseline(D, budget_s=20.0):
    """float64 oracle restatement (numpy/OpenBLAS) of the same apply on host cores."""
    from threadpoolctl import threadpool_info, threadpool_limits
    from oracle import oracle as O
    # the host share of one GPU is 16 threads (OMP_NUM_THREADS on the box);
    # all of them run the BLAS products
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), avail))
    rng = np.random.default_rng(5)
    d = 2
    W = jordan_w(rng)
    GL, GR = rng.standard_normal((D, 5, D)), rng.standard_normal((D, 5, D))
    x = rng.standard_normal((D, d, D, d))
    with threadpool_limits(cores):
        T2 = O.jordan_ac2_prepare(GL, W, W, GR)
        ICm, BEm, CB = O.jordan_ac2_nn_matrices(T2, D, d, D, d)
        del T2, GL, GR
        O.jordan_ac2_apply_nn(ICm, BEm, CB, x)          # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            O.jordan_ac2_apply_nn(ICm, BEm, CB, x)
            n += 1
            el = time.perf_counter() - t0
            if el > budget_s or n >= 50 or (n >= 2 and el * (n + 1) / n > budget_s):
                break
        threads = max((i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"),
                      default=cores)
    return dict(value=n / el, unit="applies/s", cores=int(threads), kind="port",
                sample=f"{n} applies of the prepared nearest-neighbour Jordan dAC2 at D={D} "
                       f"(numpy/OpenBLAS float64: x*IC + BE*x + CB term), {el:.1f} s",
                host=host_info(avail))


def host_info(avail):
    """What the host offers: logical CPUs, the ones this process may use, the model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": avail, "model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def dmrg_sweep(mk, prec, L=24, D=2048):
    """One DMRG2 sweep of the Heisenberg chain from a random state; (seconds, E/L, applies)."""
    H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, 1.0))
    psi = mk.FiniteMPS.random(L, 2, D, seed=0)
    alg = mk.DMRG2(tol=1e-12, maxiter=1, trunc_dim=D, precision=prec, exact_sweeps=True, verbosity=0,
                   alg_eigsolve=mk.Lanczos(tol=1e-10, krylovdim=20, maxiter=50))
    _, _, _, hist = mk.find_groundstate(psi, H, alg)
    return hist[-1]["time"], hist[-1]["energy"] / L, hist[-1]["numops"]


def traffic_from_profile(kernel, D):
    """HBM bytes per launch of `kernel` at this D from the committed counter summary, if present."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        rec = json.loads(f.read_text()).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("D") != D:
        return None
    return rec.get("bytes_per_launch")


def batched_workload(args):
    """20 TFIM chains swept in lockstep (grouped launches), then the learned
    Davidson preconditioner against plain Lanczos.

    Two lockstep DMRG2 sweeps use a loose local tolerance (1e-6, as the
    reference's adaptive tolerances do early on) in the mixed path; training
    pairs come from Davidson solves of one rotating instance at every third
    bond.  After the sweeps the local problems at four bonds along the chain are
    solved to a float64 residual of 1e-10: all 20 instances by one batched
    Lanczos (one grouped launch per Krylov step), and five instances spread over
    g by per-instance Lanczos, Davidson with the diagonal correction and Davidson
    with the trained network (applies and seconds of each).
    """
    import torch
    import mpskit_amd as mk
    L, n = args.L, 20
    gs = [0.1 * (i + 1) for i in range(n)]
    Ds = [int(round((256 + (768 * i) // 19) / 16.0)) * 16 for i in range(n)]
    psis = [mk.FiniteMPS.random(L, 2, D, seed=100 + i) for i, D in enumerate(Ds)]
    Hs = [mk.FiniteMPOHamiltonian(mk.models.tfim(L, g)) for g in gs]
    net = mk.MLPPreconditioner(seed=1, capacity=400_000)
    count = [0]

    def hook(sweep, pos, direction, ops, xs):
        count[0] += 1
        print(f"# sweep {sweep} bond {pos} {direction}", file=sys.stderr, flush=True)
        if count[0] % 3 == 0:
            i = (count[0] // 3) % n
            mk.davidson(ops[i], xs[i], maxdim=20, maxiter=200, tol=1e-8, mlp=net, mode="none", collect=True)

    Dmax = max(Ds)
    alg = mk.DMRG2(maxiter=2, trunc_dim=Dmax, precision="bf16x3")
    loose = mk.Lanczos(tol=1e-6, krylovdim=16, maxiter=4)
    _, envs, hist = mk.find_groundstate_batched(psis, Hs, alg, loose, use_graph=True, verbosity=1, on_bond=hook)
    # parameters are selected on held-out local problems of three instances
    def local(i, pos):
        psi, H = psis[i], Hs[i]
        psi.move_center(pos)
        op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(envs[i].leftenv(pos, psi), H[pos], H[pos + 1],
                                                               envs[i].rightenv(pos + 1, psi)),
                                 precisions=("fp64", "bf16x3"))
        return op, psi.AC2(pos).clone()

    held = [local(i, L // 2 - 1) for i in (3, 9, 15)]
    rows = net.dataset_size
    sel = net.select(held, rounds=4, steps=200, lr=1e-4)
    print(f"# training pairs {rows}, selection {sel}", file=sys.stderr, flush=True)
    for op, _ in held:
        op.free()
    st = {k: 0.0 for k in ("lanczos_s", "diag_s", "mlp_s", "batched_s")}
    st.update({k: 0 for k in ("lanczos_ops", "diag_ops", "mlp_ops", "batched_ops", "solves")})
    subset = (1, 5, 9, 13, 17)
    for pos in (5, 17, 29, 41):
        print(f"# compare solvers at bond {pos}", file=sys.stderr, flush=True)
        probs = [local(i, pos) for i in range(n)]
        Bt = mk.BatchedDerivative([op for op, _ in probs])
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, _, infos = Bt.eigsolve([x for _, x in probs], krylovdim=20, maxiter=100, tol=1e-10, precision="bf16x3")
        torch.cuda.synchronize()
        st["batched_s"] += time.perf_counter() - t
        st["batched_ops"] += max(i["numops"] for i in infos)
        Bt.free()
        for i in subset:
            op, x = probs[i]
            for mode in ("lanczos", "diag", "mlp"):
                torch.cuda.synchronize()
                t = time.perf_counter()
                if mode == "lanczos":
                    _, _, info = op.eigsolve(x, krylovdim=20, maxiter=100, tol=1e-10)
                else:
                    _, _, info = mk.davidson(op, x, maxdim=20, maxiter=300, tol=1e-10, mlp=net, mode=mode)
                torch.cuda.synchronize()
                st[mode + "_s"] += time.perf_counter() - t
                st[mode + "_ops"] += info["numops"]
            st["solves"] += 1
        for op, _ in probs:
            op.free()
    out = {
        "metric": "batched DMRG2 sweep time, 20 TFIM chains in lockstep (grouped launches); learned Davidson "
                  "vs Lanczos applies to a 1e-10 residual",
        "value": hist[-1]["time"], "unit": "s/sweep", "n_gpus": 1, "steps": 2, "warmup": 0,
        "ms_per_step": hist[-1]["time"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16x3", "data": "synthetic",
        "config": {"workload": f"20 TFIM chains g=0.1..2.0, L={L}, D_i=256..1024 (multiples of 16), "
                               "lockstep DMRG2, local tol 1e-6", "D_i": Ds},
        "sweeps": [dict(iter=h["iter"], time=h["time"], solve_time=h["solve_time"], applies_slowest=h["numops"])
                   for h in hist],
        "energies_per_site": [E / L for E in hist[-1]["energies"]],
        "batched_solve_1e-10": {"bonds": 4, "instances": n, "seconds": st["batched_s"],
                                "applies_slowest": st["batched_ops"]},
        "learned": {"training_pairs": rows, "selection_applies": sel, "local_solves": st["solves"],
                    "instances": list(subset), "bonds": [5, 17, 29, 41],
                    "applies": {k: st[k + "_ops"] for k in ("lanczos", "diag", "mlp")},
                    "seconds": {k: st[k + "_s"] for k in ("lanczos", "diag", "mlp")}, "tol": 1e-10},
    }
    print(json.dumps(out), flush=True)


# 2048 states over 21 charge sectors, a discretised Gaussian in the charge (sigma ~ 3 sectors)
U1_CENTER = [1, 3, 8, 18, 37, 68, 112, 165, 218, 258, 272, 258, 218, 165, 112, 68, 37, 18, 8, 3, 1]


def u1_workload(args):
    """U(1) XXZ DMRG2 (L = 64, Sz = 0, D = 2048 over 21 charge sectors with a
    discretised Gaussian profile; interior sectors carry zero padding states up
    to a multiple of 8): sweep time and applies in float64 and in the mixed
    path, energies of both.  The local problems run on compact block storage
    (--u1-storage dense: the dense charge-sorted layout)."""
    import torch
    import mpskit_amd as mk
    L, nup = args.L if args.L != 48 else 64, None
    nup = L // 2
    prof = U1_CENTER

    def dims(k, q):
        # centre the profile on the bond's mean charge k/2
        i = q - (k // 2) + len(prof) // 2
        return prof[i] if 0 <= i < len(prof) else 0

    res = {}
    for prec in ("fp64", "bf16x3"):
        psi = mk.U1FiniteMPS.random(L, nup, dims, seed=0)
        H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, 1.0))
        print(f"# u1 {prec} bonds {psi.bond_dims()}", file=sys.stderr, flush=True)
        hist = mk.find_groundstate_u1(psi, H, trunc_dim=2048, sweeps=args.u1_sweeps, precision=prec, tol=1e-10,
                                      verbosity=1, storage=args.u1_storage)
        res[prec] = dict(sweeps=[dict(time=h["time"], solve_time=h["solve_time"], applies=h["numops"],
                                      energy_per_site=h["energy"] / L) for h in hist],
                         mid_bond=hist[-1]["mid_bond"])
    last = {p: r["sweeps"][-1] for p, r in res.items()}
    out = {
        "metric": "U(1) XXZ DMRG2 sweep time at D=2048 over 21 charge sectors (charge-block grouped GEMMs)",
        "value": last["bf16x3"]["time"], "unit": "s/sweep", "n_gpus": 1, "steps": args.u1_sweeps, "warmup": 0,
        "ms_per_step": last["bf16x3"]["time"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16x3", "data": "synthetic",
        "config": {"workload": f"U(1) XXZ DMRG2, L={L}, Sz=0, D=2048 over 21 sectors {prof}",
                   "storage": args.u1_storage},
        "fp64": res["fp64"], "bf16x3": res["bf16x3"],
        "energy_diff_per_site": abs(last["fp64"]["energy_per_site"] - last["bf16x3"]["energy_per_site"]),
    }
    print(json.dumps(out), flush=True)


def sweep_workload(args):
    """Where one DMRG2 sweep at D=2048 spends its time: the sweep of the default
    line at chain length --sweep-L (64: 43 of the 63 bonds at 2048, truncating
    from 4096), run once per precision with host timers (device synchronised) around
    the local eigensolves, the truncated splits, the QR of the local error
    estimate, the environment updates and the operator preparation."""
    import collections
    import torch
    import mpskit_amd as mk
    alg_mod = sys.modules["mpskit_amd.algorithms"]
    env_mod = sys.modules["mpskit_amd.environments"]
    der = sys.modules["mpskit_amd.derivatives"]
    acc, cnt = collections.defaultdict(float), collections.defaultdict(int)

    def timed(obj, name, key):
        f = getattr(obj, name)

        def g(*a, **k):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = f(*a, **k)
            torch.cuda.synchronize()
            acc[key] += time.perf_counter() - t
            cnt[key] += 1
            return r
        setattr(obj, name, g)

    timed(der.PreparedDerivative, "eigsolve", "local eigensolves")
    timed(alg_mod, "truncated_split", "truncated splits")
    timed(alg_mod, "qr_left", "QR for the local error estimate")
    timed(env_mod, "transfer_left", "environment updates")
    timed(env_mod, "transfer_right", "environment updates")
    timed(alg_mod, "prepare_operator", "operator preparation")
    out = {"metric": f"DMRG2 sweep time breakdown, Heisenberg chain L={args.sweep_L}, bonds up to 2048", "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
           "config": {"workload": "one DMRG2 sweep from a random normalised state (seed 0), Lanczos tol 1e-10"}}
    for prec in ((args.only,) if args.only else ("fp64", "bf16x3")):
        acc.clear()
        cnt.clear()
        sec, e, nops = dmrg_sweep(mk, prec, L=args.sweep_L)
        out[prec] = {"seconds": sec, "energy_per_site": e, "applies": nops,
                     "parts": {k: {"seconds": acc[k], "calls": cnt[k]} for k in sorted(acc, key=lambda k: -acc[k])}}
        print(f"# {prec}: sweep {sec:.2f} s", file=sys.stderr, flush=True)
    last = args.only or "bf16x3"
    out["value"] = out[last]["seconds"]
    out["dtype"] = last
    print(json.dumps(out), flush=True)


def idmrg_workload(args):
    """Single-site IDMRG on the J1-J2 chain of the VUMPS workload (J2/J1 = 0.24,
    unit cell 2, D = --vumps-D), --vumps-iters iterations from one random start,
    float64 and BF16x3 applies: seconds per iteration (iterations 2 on) and the
    energy per site of the regauged result."""
    import torch
    import mpskit_amd as mk
    from mpskit_amd import idmrg as I, vumps as V
    D, iters = args.vumps_D, args.vumps_iters
    H = V.InfiniteHamiltonian([mk.models.j1j2_bulk(1.0, 0.24)] * 2)
    res = {}
    for prec in ("fp64", "bf16x3"):
        psi = V.InfiniteMPS.random(2, 2, D, seed=7)
        torch.cuda.synchronize()
        print(f"# idmrg {prec} D={D}", file=sys.stderr, flush=True)
        psi, envs, eps, hist = I.find_groundstate_idmrg(psi, H, I.IDMRG(tol=0.0, maxiter=iters, precision=prec,
                                                                        verbosity=2))
        ts = [h["time"] for h in hist[1:]] or [hist[0]["time"]]
        res[prec] = dict(seconds_per_iteration=sum(ts) / len(ts), energy_per_site=hist[-1]["energy"],
                         boundary_change=eps, applies_per_iteration=sum(h["numops"] for h in hist) / len(hist),
                         iterations=len(hist))
    out = {"metric": "IDMRG time per iteration, infinite J1-J2 chain (J2/J1=0.24, cell 2, w=8)",
           "value": res["bf16x3"]["seconds_per_iteration"], "unit": "s/iteration", "n_gpus": 1,
           "steps": iters, "warmup": 1, "ms_per_step": res["bf16x3"]["seconds_per_iteration"] * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
           "data": "synthetic",
           "config": {"workload": f"single-site IDMRG J1-J2 J2/J1=0.24 unit cell 2 D={D}, random start seed 7, "
                                  f"{iters} iterations"},
           "fp64": res["fp64"], "bf16x3": res["bf16x3"],
           "speedup": res["fp64"]["seconds_per_iteration"] / res["bf16x3"]["seconds_per_iteration"]}
    print(json.dumps(out), flush=True)


def idmrg2_workload(args):
    """Two-site IDMRG on the J1-J2 chain of the VUMPS workload (J2/J1 = 0.24,
    unit cell 2) growing from D = 16 to --idmrg2-D kept states, --vumps-iters
    iterations, float64 and BF16x3 applies: seconds per iteration (iterations 2
    on), bond dimension reached and the energy per site of the regauged result."""
    import torch
    import mpskit_amd as mk
    from mpskit_amd import idmrg as I, vumps as V
    D, iters = args.idmrg2_D, args.vumps_iters
    H = V.InfiniteHamiltonian([mk.models.j1j2_bulk(1.0, 0.24)] * 2)
    res = {}
    for prec in ("fp64", "bf16x3"):
        psi = V.InfiniteMPS.random(2, 2, 16, seed=7)
        torch.cuda.synchronize()
        print(f"# idmrg2 {prec} D={D}", file=sys.stderr, flush=True)
        psi, envs, eps, hist = I.find_groundstate_idmrg2(
            psi, H, I.IDMRG2(trunc_dim=D, tol=0.0, maxiter=iters, precision=prec, verbosity=2))
        ts = [h["time"] for h in hist[1:]] or [hist[0]["time"]]
        res[prec] = dict(seconds_per_iteration=sum(ts) / len(ts), energy_per_site=hist[-1]["energy"],
                         boundary_change=eps, bond_dims=[h["D"] for h in hist],
                         applies_per_iteration=sum(h["numops"] for h in hist) / len(hist), iterations=len(hist))
    out = {"metric": "IDMRG2 time per iteration, infinite J1-J2 chain (J2/J1=0.24, cell 2, w=8)",
           "value": res["bf16x3"]["seconds_per_iteration"], "unit": "s/iteration", "n_gpus": 1,
           "steps": iters, "warmup": 1, "ms_per_step": res["bf16x3"]["seconds_per_iteration"] * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
           "data": "synthetic",
           "config": {"workload": f"two-site IDMRG J1-J2 J2/J1=0.24 unit cell 2, D 16 -> {D}, random start seed 7, "
                                  f"{iters} iterations"},
           "fp64": res["fp64"], "bf16x3": res["bf16x3"],
           "speedup": res["fp64"]["seconds_per_iteration"] / res["bf16x3"]["seconds_per_iteration"]}
    print(json.dumps(out), flush=True)


def vumps_workload(args):
    """VUMPS on the infinite J1-J2 chain (J2/J1 = 0.24, unit cell 2, w = 8) at bond
    dimension D: seconds per iteration in float64 and in the mixed path (FP64
    environment normalisation, gauge and energies in both), over a fixed number
    of iterations from the same random state, with the time of each part of the
    iterations (host timers, device synchronised)."""
    import collections
    import torch
    import mpskit_amd as mk
    from mpskit_amd import vumps as V
    der = sys.modules["mpskit_amd.derivatives"]
    D, iters = args.vumps_D, args.vumps_iters
    # host timers (device synchronised) around the parts of an iteration
    acc, cnt = collections.defaultdict(float), collections.defaultdict(int)

    def timed(obj, name, key_of):
        f = getattr(obj, name)

        def g(*a, **k):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = f(*a, **k)
            torch.cuda.synchronize()
            key = key_of(a)
            acc[key] += time.perf_counter() - t
            cnt[key] += 1
            return r
        setattr(obj, name, g)

    timed(der.PreparedDerivative, "eigsolve", lambda a: "AC eigensolves" if len(a[0].x_shape) == 3 else "C eigensolves")
    timed(V, "uniform_rightorth", lambda a: "gauge (right orthogonalisation of the cell)")
    # inside the gauge step (also counted in it): its LQ factorisations and fixed-point Arnoldi
    timed(V, "lq_right", lambda a: "  gauge: positive LQ factorisations")
    timed(V, "arnoldi_dominant", lambda a: "  gauge: fixed-point Arnoldi")
    timed(V.InfiniteEnvironments, "recalculate", lambda a: "environments")
    timed(V, "regauge", lambda a: "regauge")
    timed(V, "calc_galerkin", lambda a: "Galerkin error")
    timed(V, "prepare_operator", lambda a: "operator preparation")
    res = {}
    for prec in ("fp64", "bf16x3"):
        acc.clear()
        cnt.clear()
        H = V.InfiniteHamiltonian([mk.models.j1j2_bulk(1.0, 0.24)] * 2)
        psi = V.InfiniteMPS.random(2, 2, D, seed=7)
        print(f"# vumps {prec} D={D}", file=sys.stderr, flush=True)
        alg = V.VUMPS(tol=1e-12, maxiter=iters, precision=prec, verbosity=2)
        _, _, eps, hist = V.find_groundstate_vumps(psi, H, alg)
        per = [h["time"] for h in hist[1:]] or [hist[0]["time"]]
        last = [h["time"] for h in hist[-5:]]
        res[prec] = dict(seconds_per_iteration=sum(per) / len(per), energy_per_site=hist[-1]["energy"],
                         galerkin=hist[-1]["eps"], iterations=len(hist), total_seconds=sum(h["time"] for h in hist),
                         seconds_per_iteration_last5=sum(last) / len(last),
                         energies=[h["energy"] for h in hist], errors=[h["eps"] for h in hist],
                         parts={k: {"seconds": acc[k], "calls": cnt[k]} for k in sorted(acc, key=lambda k: -acc[k])})
    out = {
        "metric": "VUMPS time per iteration, infinite J1-J2 chain (J2/J1=0.24, cell 2, w=8)",
        "value": res["bf16x3"]["seconds_per_iteration"], "unit": "s/iteration", "n_gpus": 1, "steps": iters,
        "warmup": 1, "ms_per_step": res["bf16x3"]["seconds_per_iteration"] * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3", "data": "synthetic",
        "config": {"workload": f"VUMPS J1-J2 J2/J1=0.24 unit cell 2 D={D}, random start seed 7, {iters} iterations"},
        "fp64": res["fp64"], "bf16x3": res["bf16x3"],
        "speedup": res["fp64"]["seconds_per_iteration"] / res["bf16x3"]["seconds_per_iteration"],
        "energy_diff_per_site": abs(res["fp64"]["energy_per_site"] - res["bf16x3"]["energy_per_site"]),
    }
    print(json.dumps(out), flush=True)


def tdvp_workload(args):
    """Imaginary-time TDVP2 on the Heisenberg chain (L = 24, bonds up to
    --tdvp-D, truncation at --tdvp-D), dt = 0.05, --vumps-iters steps from one
    random normalised state (seed 0), float64: seconds per time step (steps 2 on),
    applies per step and the energy per site after the steps."""
    import torch
    import mpskit_amd as mk
    from mpskit_amd import tdvp as Td
    L, D, steps, dt = 24, args.tdvp_D, args.vumps_iters, 0.05
    H = mk.FiniteMPOHamiltonian(mk.models.xxz(L, 1.0))
    psi = mk.FiniteMPS.random(L, 2, D, seed=0)
    envs = mk.environments(psi, H)
    times, ops = [], []
    print(f"# tdvp2 L={L} D={D}", file=sys.stderr, flush=True)
    for k in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        psi, envs, st = Td.timestep(psi, H, dt, Td.TDVP2(trunc_dim=D), envs, normalize=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        ops.append(st["numops"])
        print(f"# step {k + 1}: {times[-1]:.3f} s, {st['numops']} applies", file=sys.stderr, flush=True)
    E = mk.expectation_value(psi, H, envs) / L
    per = times[1:] or times
    out = {
        "metric": "imaginary-time TDVP2 time per step, Heisenberg chain L=24",
        "value": sum(per) / len(per), "unit": "s/step", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": sum(per) / len(per) * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"TDVP2 imaginary time, Heisenberg L={L}, D<={D}, dt={dt}, random start seed 0, "
                               f"{steps} steps"},
        "applies_per_step": sum(ops[1:] or ops) / len(ops[1:] or ops), "energy_per_site": E,
        "bond_dims": psi.bond_dims(),
    }
    print(json.dumps(out), flush=True)


def boundary_workload(args):
    """leading_boundary of the square-lattice Ising transfer MPO (beta = 0.42,
    disordered side of beta_c = 0.4407, unit cell 1, w = d = 2) with VUMPS at bond
    dimension --boundary-D: --vumps-iters iterations from one random state in
    float64 and in the mixed path (AC / C eigensolves in BF16x3 refined in FP64;
    environments, gauge and the Galerkin error in FP64).  Reports seconds per
    iteration (iterations 2 on) and ln Z per site against Onsager's value."""
    import torch
    import mpskit_amd as mk
    from mpskit_amd import boundary as Bd, vumps as V
    D, iters, beta = args.boundary_D, args.vumps_iters, 0.42
    O = Bd.InfiniteMPO([mk.models.classical_ising(beta)])
    res = {}
    for prec in ("fp64", "bf16x3"):
        psi = V.InfiniteMPS.random(1, 2, D, seed=7)
        torch.cuda.synchronize()
        print(f"# boundary {prec} D={D}", file=sys.stderr, flush=True)
        _, _, eps, hist = Bd.leading_boundary(psi, O, Bd.BoundaryVUMPS(tol=1e-12, maxiter=iters, precision=prec,
                                                                       verbosity=2))
        per = [h["time"] for h in hist[1:]] or [hist[0]["time"]]
        res[prec] = dict(seconds_per_iteration=sum(per) / len(per), log_z_per_site=float(np.log(hist[-1]["value"])),
                         galerkin=eps, iterations=len(hist), mixed_applies=hist[-1]["mixed_ops"],
                         fp64_applies=hist[-1]["fp64_ops"])
    out = {
        "metric": "leading_boundary VUMPS time per iteration, 2D Ising transfer MPO (beta=0.42, w=2)",
        "value": res["bf16x3"]["seconds_per_iteration"], "unit": "s/iteration", "n_gpus": 1, "steps": iters,
        "warmup": 1, "ms_per_step": res["bf16x3"]["seconds_per_iteration"] * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3", "data": "synthetic",
        "config": {"workload": f"VUMPS leading boundary, classical Ising beta=0.42, unit cell 1, D={D}, "
                               f"random start seed 7, {iters} iterations"},
        "fp64": res["fp64"], "bf16x3": res["bf16x3"],
        "speedup": res["fp64"]["seconds_per_iteration"] / res["bf16x3"]["seconds_per_iteration"],
        "log_z_diff": abs(res["fp64"]["log_z_per_site"] - res["bf16x3"]["log_z_per_site"]),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--D", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32", "bf16x3"], default=None,
                    help="time one precision only (for profiler runs); prints no JSON line")
    ap.add_argument("--no-sweep", action="store_true", help="skip the DMRG2 sweep at D=2048 (about 3 minutes)")
    ap.add_argument("--no-table", action="store_true", help="skip the table of the other isolated applies")
    ap.add_argument("--workload", choices=["apply", "batched", "vumps", "u1", "sweep", "idmrg", "idmrg2", "boundary", "tdvp"], default="apply")
    ap.add_argument("--vumps-D", type=int, default=1024)
    ap.add_argument("--sweep-L", type=int, default=64, help="chain length of the DMRG2 sweeps (default line and sweep workload)")
    ap.add_argument("--u1-sweeps", type=int, default=2)
    ap.add_argument("--u1-storage", choices=["compact", "dense"], default="compact",
                    help="charge-block storage of the U(1) workload's local problems")
    ap.add_argument("--vumps-iters", type=int, default=6)
    ap.add_argument("--idmrg2-D", type=int, default=256, help="kept states of the IDMRG2 workload")
    ap.add_argument("--boundary-D", type=int, default=1024, help="bond dimension of the boundary workload")
    ap.add_argument("--tdvp-D", type=int, default=512, help="maximal bond dimension of the TDVP2 workload")
    ap.add_argument("--L", type=int, default=48, help="chain length of the batched workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB device-to-device five times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on a single "
              "MI355X; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    import mpskit_amd as mk
    from mpskit_amd import tensors as T

    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print("# copied 5 x 1 GiB", file=sys.stderr)
        return

    if args.workload == "batched":
        batched_workload(args)
        return
    if args.workload == "vumps":
        vumps_workload(args)
        return
    if args.workload == "u1":
        u1_workload(args)
        return
    if args.workload == "sweep":
        sweep_workload(args)
        return
    if args.workload == "idmrg":
        idmrg_workload(args)
        return
    if args.workload == "idmrg2":
        idmrg2_workload(args)
        return
    if args.workload == "boundary":
        boundary_workload(args)
        return
    if args.workload == "tdvp":
        tdvp_workload(args)
        return

    D, d = args.D, 2
    W = T.from_numpy(jordan_w(np.random.default_rng(3)))
    GL = gpu_unit(T, (D, 5, D), 1)
    GR = gpu_unit(T, (D, 5, D), 2)
    x = gpu_unit(T, (D, d, D, d), 4)
    op = mk.prepare_operator(mk.JordanMPO_AC2_Hamiltonian(GL, W, W, GR), precisions=("fp64", "fp32", "bf16x3"))
    del GL, GR
    y = T.empty(x.shape)
    lib = mk.load()

    res = {}
    y64 = None
    for prec in ((args.only,) if args.only else ("fp64", "fp32", "bf16x3")):
        lib.mk_prof_reset()
        lib.mk_prof_enable(1)
        ms = time_applies(op, x, y, prec, args.steps, args.warmup)
        lib.mk_prof_enable(0)
        tot, cnt = ctypes.c_double(), ctypes.c_int64()
        lib.mk_prof_query(KERNEL[prec].encode(), ctypes.byref(tot), ctypes.byref(cnt))
        gemm_ms = tot.value / max(cnt.value, 1)
        if prec == "fp64":
            y64 = y.clone()
            err = 0.0
        elif y64 is None:
            err = float("nan")
        else:
            err = float(torch.linalg.vector_norm(y - y64) / torch.linalg.vector_norm(y64))
        res[prec] = dict(ms=ms, gemm_ms=gemm_ms, launches=cnt.value, err=err)
        print(f"# {prec}: {ms:.3f} ms/apply, gemm {gemm_ms:.3f} ms, {op.flops / ms / 1e9:.1f} TF/s, "
              f"rel err vs fp64 {err:.2e}", file=sys.stderr, flush=True)

    if args.only:
        return
    F = op.flops
    head = res["bf16x3"]
    ach = F / (head["gemm_ms"] * 1e-3) / 1e12
    out = {
        "metric": "dAC2 applies/s at D=2048 (mixed split-BF16x3 path; FP64 GPU path alongside)",
        "value": 1e3 / head["ms"],
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3",
        "data": "synthetic",
        "config": {"workload": f"isolated Jordan dAC2, XXZ w=5 d=2, D={D}, float64 in/out",
                   "D": D, "w": 5, "d": 2, "flops_per_apply": F},
        "fp64_applies_per_s": 1e3 / res["fp64"]["ms"],
        "fp32_applies_per_s": 1e3 / res["fp32"]["ms"],
        "speedup_vs_fp64": res["fp64"]["ms"] / head["ms"],
        "rel_err_vs_fp64": {"bf16x3": res["bf16x3"]["err"], "fp32": res["fp32"]["err"]},
        "tflops": {p: F / (r["ms"] * 1e-3) / 1e12 for p, r in res.items()},
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_TF["bf16x3"], "unit": "TFLOP/s",
                     "frac": ach / PEAK_TF["bf16x3"],
                     "traffic": traffic_from_profile("gemm_mixed_big2_kernel<BF16X3>", D),
                     "kernel": "gemm_mixed_big2_kernel<BF16X3> (256x256 tiles, 8 waves)",
                     "peak_note": "BF16 dense 2516.6 TF/s over 3 products per float64 product"},
        "roofline_fp64": {"bound": "mfma", "achieved": F / (res["fp64"]["gemm_ms"] * 1e-3) / 1e12,
                          "peak": PEAK_TF["fp64"], "unit": "TFLOP/s",
                          "frac": F / (res["fp64"]["gemm_ms"] * 1e-3) / 1e12 / PEAK_TF["fp64"],
                          "traffic": traffic_from_profile("gemm_f64_buf_kernel<128,128>", D),
                          "kernel": "gemm_f64_buf_kernel<128,128> (branch-free loads, rotated K step)"},
    }
    print("# vendor GEMM reference ...", file=sys.stderr, flush=True)
    out["vendor_gemm_reference"] = vendor_gemm_reference(D)
    print("# batched applies ...", file=sys.stderr, flush=True)
    out["batched_apply"] = batched_apply(mk, T)
    if not args.no_table:
        print("# apply table ...", file=sys.stderr, flush=True)
        out["apply_table"] = apply_table(mk, T)
        print("# quasiparticle transfers ...", file=sys.stderr, flush=True)
        out["qp_transfers"] = qp_transfer_table(mk, T)
    if not args.no_sweep:
        sw = {}
        for prec in ("fp64", "bf16x3"):
            print(f"# DMRG2 sweep {prec} (L={args.sweep_L}) ...", file=sys.stderr, flush=True)
            sec, e, nops = dmrg_sweep(mk, prec, L=args.sweep_L)
            sw[prec] = dict(seconds=sec, energy_per_site=e, applies=nops)
            print(f"# DMRG2 sweep {prec}: {sec:.2f} s", file=sys.stderr, flush=True)
        out["dmrg_sweep"] = {"workload": f"DMRG2, Heisenberg chain L={args.sweep_L}, bonds up to 2048 (truncating from "
                                         f"4096), one sweep from a random normalised state (seed 0), Lanczos tol 1e-10",
                             "fp64": sw["fp64"], "bf16x3": sw["bf16x3"],
                             "speedup": sw["fp64"]["seconds"] / sw["bf16x3"]["seconds"],
                             "energy_diff_per_site": abs(sw["fp64"]["energy_per_site"] -
                                                         sw["bf16x3"]["energy_per_site"])}
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(D)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
