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

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--D 2048] [--no-cpu]
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


def cpu_baseline(D, budget_s=20.0):
    """float64 oracle restatement (numpy/OpenBLAS) of the same apply on host cores."""
    from threadpoolctl import threadpool_info, threadpool_limits
    from oracle import oracle as O
    cores = min(16, os.cpu_count() or 1)
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
                       f"(numpy/OpenBLAS float64: x*IC + BE*x + CB term), {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--D", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32", "bf16x3"], default=None,
                    help="time one precision only (for profiler runs); prints no JSON line")
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
                     "frac": ach / PEAK_TF["bf16x3"], "traffic": None,
                     "kernel": "gemm_mixed_kernel<128,128,BF16X3>",
                     "peak_note": "BF16 dense 2516.6 TF/s over 3 products per float64 product"},
        "roofline_fp64": {"bound": "mfma", "achieved": F / (res["fp64"]["gemm_ms"] * 1e-3) / 1e12,
                          "peak": PEAK_TF["fp64"], "unit": "TFLOP/s",
                          "frac": F / (res["fp64"]["gemm_ms"] * 1e-3) / 1e12 / PEAK_TF["fp64"],
                          "kernel": "gemm_f64_kernel<128,128>"},
    }
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(D)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
