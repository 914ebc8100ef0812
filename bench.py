#!/usr/bin/env python3
"""Benchmark: batched RGF selected solve at 64 blocks of 256 on one MI355X.

One step = ``RGF.selected_solve`` (G^R diagonal, G^<, G^> blocks) for all
512 energies of the synthetic production workload (N = 64 blocks of
256 x 256, complex128 storage).  A, Sigma^< and Sigma^> of all energies are
resident in HBM before timing (102 GB + 2 x 34 GB); the energies are solved
in chunks of ``--chunk`` whose outputs (G^<, G^> with the BTD pattern of A,
G^R block-diagonal) are written into one reused set of output stacks, since
the outputs of all 512 energies (~240 GB) do not fit next to the inputs.

Printed: one JSON line with energies/s (whole job), ms/step, per-call
latency percentiles, the roofline of the dominant kernel (the batched MFMA
GEMM; achieved = algorithmic 8 m n k flops per launch / HIP-event launch
time over the timed region) and the CPU baseline (the NumPy oracle, a
restatement of the reference algorithm, timed on a bounded energy sample on
the host cores).
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "RGF energies/s at 64 blocks×256, mixed vs c128 GPU; p50/p99 solve latency"
PEAK_TFLOPS = {"c128": 78.6, "c64": 157.3}  # MI355X dense FP64 / FP32 matrix peaks
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(w, budget_s: float):
    """Time the NumPy oracle (reference algorithm restated) on a few energies."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    import oracle.rgf as orgf
    from oracle.dsdbcoo import DSDBCOO as ODSDBCOO
    from quatrex_amd.datastructures import btd_pattern

    bs = w.block_sizes
    rows, cols, aoff = btd_pattern(bs, offdiag=True)
    drows, dcols, doff = btd_pattern(bs, offdiag=False)

    def containers(energies):
        diag, upper, lower = w.a_blocks(energies)
        ne = len(energies)
        a = np.empty((ne, len(rows)), complex)
        sl = np.empty((ne, len(drows)), complex)
        sg = np.empty((ne, len(drows)), complex)
        for i in range(len(bs)):
            b = bs[i]
            o = aoff[(i, i)]
            a[:, o:o + b * b] = diag[i].reshape(ne, -1)
            o = doff[(i, i)]
            sl[:, o:o + b * b] = w.sl_diag[i].reshape(1, -1)
            sg[:, o:o + b * b] = w.sg_diag[i].reshape(1, -1)
            if i + 1 < len(bs):
                o = aoff[(i, i + 1)]
                a[:, o:o + b * bs[i + 1]] = upper[i].reshape(ne, -1)
                o = aoff[(i + 1, i)]
                a[:, o:o + b * bs[i + 1]] = lower[i].reshape(ne, -1)
        mk = lambda d, r, c: ODSDBCOO(d, r, c, bs, (ne,))
        return mk(a, rows, cols), mk(sl, drows, dcols), mk(sg, drows, dcols)

    cores = host_cores()
    with threadpool_limits(limits=cores):
        solver = orgf.RGF(max_batch_size=1)
        a, sl, sg = containers(w.energies[:1])
        t0 = time.perf_counter()
        solver.selected_solve(a, sl, sg, return_retarded=True)
        t1 = time.perf_counter() - t0
        n = int(max(1, min(8, budget_s // max(t1, 1e-3))))
        a, sl, sg = containers(w.energies[1:1 + n])
        t0 = time.perf_counter()
        solver.selected_solve(a, sl, sg, return_retarded=True)
        dt = time.perf_counter() - t0
    return {
        "value": n / dt,
        "unit": "energies/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"NumPy oracle selected_solve (return_retarded) on {n} energies of the same "
                   f"workload (N={len(bs)}, b={bs[0]}, complex128, OpenBLAS/LAPACK, "
                   f"{cores} BLAS threads); first energy timed separately as warm-up "
                   f"({t1:.2f} s)"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="c128", choices=["c128", "c64"])
    ap.add_argument("--energies", type=int, default=512)
    ap.add_argument("--blocks", type=int, default=64)
    ap.add_argument("--block-size", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU work for the baseline sample (0 disables)")
    ap.add_argument("--no-timer", action="store_true",
                    help="do not bracket kernels with HIP events (no roofline)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Energies shard with no communication; "
              f"run one process per GPU on disjoint energy ranges.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from quatrex_amd import _lib
    from quatrex_amd.datastructures import DSDBCOO, empty_btd
    from quatrex_amd.greens_function_solver import RGF
    from quatrex_amd.workloads import make_workload, solve_flops

    dev = torch.device("cuda:0")
    E, N, b, chunk = args.energies, args.blocks, args.block_size, args.chunk
    if E % chunk:
        raise SystemExit("--energies must be a multiple of --chunk")
    log(f"workload: {E} energies, {N} blocks of {b}, precision {args.precision}, chunk {chunk}")
    w = make_workload(seed=args.seed, n_energies=E, n_blocks=N, block_size=b)
    t0 = time.perf_counter()
    a, sl, sg = w.assemble(dev)
    torch.cuda.synchronize()
    log(f"inputs resident: A {a.data.numel() * 16 / 1e9:.1f} GB, sigma "
        f"{2 * sl.data.numel() * 16 / 1e9:.1f} GB ({time.perf_counter() - t0:.1f} s)")
    xl = DSDBCOO.zeros_like(a.stack_view(0, chunk))
    xg = DSDBCOO.zeros_like(a.stack_view(0, chunk))
    xr, _ = empty_btd(w.block_sizes, (chunk,), offdiag=False, device=dev)
    rgf = RGF(max_batch_size=chunk, precision=args.precision)
    views = [(a.stack_view(c, c + chunk), sl.stack_view(c, c + chunk), sg.stack_view(c, c + chunk))
             for c in range(0, E, chunk)]

    def step(lat):
        for va, vl, vg in views:
            t = time.perf_counter()
            rgf.selected_solve(va, vl, vg, out=(xl, xg, xr), return_retarded=True)
            lat.append(time.perf_counter() - t)

    for i in range(args.warmup):
        t = time.perf_counter()
        step([])
        log(f"warmup step {i} done ({time.perf_counter() - t:.2f} s)")
    torch.cuda.synchronize()
    lat = []
    timer = None if args.no_timer else _lib.KernelTimer(dev)
    if timer:
        timer.__enter__()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(lat)
        log(f"timed step {i} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if timer:
        timer.__exit__(None, None, None)
    ms_per_step = 1e3 * elapsed / args.steps
    value = E * args.steps / elapsed
    lat_ms = np.array(lat) * 1e3

    roofline = None
    kernels = None
    if timer:
        g, inv = timer.stats["gemm"], timer.stats["inv"]
        # Sub-batches run on concurrent streams, so launches overlap: the GEMM
        # throughput is its flops over the wall time during which at least one
        # GEMM launch was in flight (union of the HIP-event intervals).
        achieved = g["flops"] / (g["busy_ms"] / 1e3) / 1e12
        peak = PEAK_TFLOPS[args.precision]
        per_launch = g["flops"] / max(g["launches"], 1)
        avg_ms = g["ms"] / max(g["launches"], 1)
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                    "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": None,
                    "kernel": "zgemm_kernel (batched complex GEMM, 64x64 tiles)",
                    "flops_per_launch": per_launch, "avg_launch_ms": round(avg_ms, 4),
                    "per_launch_tflops": round(per_launch / (avg_ms / 1e3) / 1e12, 3),
                    "busy_ms": round(g["busy_ms"], 2)}
        kernels = {
            "gemm": {"launches": g["launches"], "sum_ms": round(g["ms"], 2),
                     "busy_ms": round(g["busy_ms"], 2), "tflops": round(achieved, 3)},
            "inverse": {"launches": inv["launches"], "sum_ms": round(inv["ms"], 2),
                        "busy_ms": round(inv["busy_ms"], 2),
                        "tflops": round(inv["flops"] / max(inv["busy_ms"], 1e-9) / 1e9, 3)},
            "wall_ms": round(elapsed * 1e3, 2),
        }
    flops = solve_flops(N, b)
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "energies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "c128" else "f32",
        "data": "synthetic",
        "config": {"workload": f"RGF selected_solve, {E} energies, {N} blocks of {b}, "
                               f"complex128 storage, {args.precision} arithmetic",
                   "energies": E, "n_blocks": N, "block_size": b, "precision": args.precision,
                   "energies_per_call": chunk},
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2),
                       "p99": round(float(np.percentile(lat_ms, 99)), 2),
                       "per": f"selected_solve call of {chunk} energies"},
        "algorithmic_tflops": round(value * flops / 1e12, 3),
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": None,
    }
    del a, sl, sg, xl, xg, xr, views
    torch.cuda.empty_cache()
    if args.cpu_budget > 0:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
