#!/usr/bin/env python3
"""Benchmark: batched RGF selected solve at 64 blocks of 256 on one MI355X.

One step = ``RGF.selected_solve`` (G^R diagonal, G^<, G^> blocks) for all
512 energies of the synthetic production workload (N = 64 blocks of
256 x 256, complex128 storage).  A, Sigma^< and Sigma^> of all energies are
resident in HBM before timing (102 GB + 2 x 34 GB); the energies are solved
in chunks of ``--chunk`` (default 128) whose outputs (G^<, G^> with the BTD
pattern of A, G^R block-diagonal) are written into one reused set of output
stacks, since the outputs of all 512 energies (~240 GB) do not fit next to
the inputs.

The primary arithmetic (``--precision``, default complex128 with the 4M
GEMM split) gives ``value``; every mode of ``--compare`` (default:
c128-3m = complex128 with the 3M split, mixed = complex64 products and
inverses + one FP64 Newton-Schulz step per block inverse, oz2-12 =
block products on int8 matrix cores by modular arithmetic with 12 moduli
and FP64 inverses, bf16x7 = block products on bf16 matrix cores from 7
digit slices and FP64 inverses) is timed the
same way and reported beside it with its speed-up and its error against the
complex128 GPU result (per block, electron density diag G^<, transmission
T(E) = Tr[Gamma_L G^R_{0,N-1} Gamma_R G^A_{N-1,0}]) on the first
``--accuracy-energies`` energies.  The CPU baseline leg times the NumPy
oracle on a bounded energy sample and compares every GPU mode with it on
those energies.

Printed: one JSON line with energies/s (whole job), ms/step, per-call
latency percentiles, the roofline of the step (matrix-core flops of every
GEMM and inverse launch, counted per launch with HIP events over the timed
region, / wall time; the dominant GEMM kernel's per-launch figures beside
it) and the CPU baseline.
"""

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "RGF energies/s at 64 blocks×256, mixed vs c128 GPU; p50/p99 solve latency"
# MI355X dense matrix peaks of the GEMM arithmetic each mode runs on
PEAK_TFLOPS = {"c128": 78.6, "c128-3m": 78.6, "c64": 157.3, "mixed": 157.3}
# modular-arithmetic products on int8 cores: rated against the FP64 matrix
# peak of the arithmetic they replace (complex flops 8 m n k per product)
PEAK_TFLOPS.update({f"oz2-{n}": 78.6 for n in range(8, 21)})
PEAK_TFLOPS.update({f"bf16x{s}": 78.6 for s in range(2, 8)})
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


def cpu_baseline(w, budget_s: float, max_energies: int):
    """Time the NumPy oracle (reference algorithm restated) on a few energies.

    Returns (oracle outputs on the timed energies, baseline record).
    """
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
        n = int(max(1, min(max_energies - 1, budget_s // max(t1, 1e-3))))
        a, sl, sg = containers(w.energies[1:1 + n])
        corner = np.zeros((n, bs[0], bs[-1]), complex)
        t0 = time.perf_counter()
        xl, xg, xr = solver.selected_solve(a, sl, sg, return_retarded=True, corner=corner)
        dt = time.perf_counter() - t0
    ref = {"energies": (1, 1 + n), "density": np.stack(
               [np.diagonal(xl.blocks[i, i], axis1=-2, axis2=-1) for i in range(len(bs))], 1),
           "diag_l": [xl.blocks[i, i] for i in range(len(bs))],
           "corner": corner}
    return ref, {
        "value": n / dt,
        "unit": "energies/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"NumPy oracle selected_solve (return_retarded) on {n} energies of the same "
                   f"workload (N={len(bs)}, b={bs[0]}, complex128, OpenBLAS/LAPACK, "
                   f"{cores} BLAS threads); first energy timed separately as warm-up "
                   f"({t1:.2f} s)"),
    }


TRAFFIC_FILE = ROOT / "profiles" / "zgemm_256_hbm_traffic.json"
# the dominant kernel: complex128 A B with both operands untransposed (70 % of
# the RGF's GEMM launches), register-direct 64 x 64 workgroup tiles
GEMM_KERNEL_KEY = "zgemm_reg_kernel<double,N,N,4M,64x64>"


STEP_COUNTERS_FILE = ROOT / "profiles" / "rgf_step_64x256_counters.json"


def measured_step_traffic(precision: str, required_step_flops: float):
    """HBM bytes per step from the committed whole-step counter run (rocprofv3
    FETCH_SIZE / WRITE_SIZE summed over every dispatch of one step, calibrated
    on a copy) when it was taken at this precision and step size; else None."""
    try:
        rec = json.loads(STEP_COUNTERS_FILE.read_text())
    except (OSError, ValueError):
        return None
    m = rec.get("modes", {}).get(precision)
    if not m or abs(rec.get("required_flops_per_step", -1.0) - required_step_flops) > 1.0:
        return None
    return {"bytes_per_step": m["traffic_bytes_per_step"],
            "source": f"rocprofv3 FETCH_SIZE + WRITE_SIZE over one step: "
                      f"profiles/{STEP_COUNTERS_FILE.name}"}


def measured_traffic(flops_per_launch: float):
    """HBM bytes per launch of the dominant GEMM from the committed counter
    measurement (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated on a copy), when
    it was taken at this launch shape; else None."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    for name, k in rec.get("kernels", {}).items():
        if name.startswith(GEMM_KERNEL_KEY) and \
                abs(k.get("flops_per_launch", -1.0) - flops_per_launch) < 1.0:
            return k["traffic_bytes"]
    return None


def timed_leg(rgf, views, out, steps, warmup, dev, timer_on, peak, dominant_flops,
              required_step_flops, fused=False):
    """Warm-up + timed steps of one arithmetic mode; returns its record.
    dominant_flops: flops of one launch of the dominant GEMM shape (one
    sub-batch of b x b x b products), the shape the traffic was measured at;
    required_step_flops: flops one step requires (all energies); fused: the
    calls run as one fused small-block kernel launch each."""
    import numpy as np
    import torch

    from quatrex_amd import _lib

    def step(lat):
        for va, vl, vg in views:
            t = time.perf_counter()
            rgf.selected_solve(va, vl, vg, out=out, return_retarded=True)
            lat.append(time.perf_counter() - t)

    for i in range(warmup):
        t = time.perf_counter()
        step([])
        log(f"{rgf.precision}: warmup step {i} done ({time.perf_counter() - t:.2f} s)")
    torch.cuda.synchronize()
    lat = []
    timer = _lib.KernelTimer(dev) if timer_on else None
    if timer:
        timer.__enter__()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks = [t0]
    for i in range(steps):
        step(lat)
        log(f"{rgf.precision}: timed step {i} done")
        marks.append(time.perf_counter())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if timer:
        timer.__exit__(None, None, None)
    lat_ms = np.array(lat) * 1e3
    # a selected_solve call returns after checking the device error flags of
    # its work (a stream synchronization), so the marks bound each step
    step_ms = np.diff(np.array(marks)) * 1e3
    rec = {"elapsed": elapsed,
           "step_ms": {"median": round(float(np.median(step_ms)), 2),
                       "min": round(float(step_ms.min()), 2),
                       "max": round(float(step_ms.max()), 2)},
           "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2),
                          "p99": round(float(np.percentile(lat_ms, 99)), 2)}}
    if timer:
        g, inv = timer.stats["gemm"], timer.stats["inv"]
        # The roofline covers the whole step: the matrix-core flops of every
        # GEMM and inverse launch of the timed steps, as launched (8 m n k per
        # complex product, 8 n^3 per inverse), over the wall time of the timed
        # steps.  The dominant kernel's per-launch figures (comparable with
        # rocprofv3's average duration) are reported beside it: its launches
        # from the 4 sub-batch streams overlap, so one launch's rate is its
        # share of the chip, not the chip's rate.
        executed = g["flops"] + inv["flops"]
        achieved = executed / elapsed / 1e12
        per_launch = g["flops"] / max(g["launches"], 1)
        avg_ms = g["ms"] / max(g["launches"], 1)
        pl_tflops = per_launch / (avg_ms / 1e3) / 1e12
        busy_union = g["flops"] / (g["busy_ms"] / 1e3) / 1e12
        step_traffic = measured_step_traffic(rgf.precision, required_step_flops)
        rec["roofline"] = {
            "bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": step_traffic["bytes_per_step"] if step_traffic else None,
            "method": ("whole step: flops of every GEMM and inverse launch of the timed steps "
                       "(8 m n k per complex product, 8 n^3 per inverse, as launched) / wall "
                       "time of the timed steps; traffic = HBM bytes per step"),
            "executed_flops_per_step": executed / steps,
            "required_flops_per_step": required_step_flops,
            "traffic_source": step_traffic["source"] if step_traffic else None,
            "dominant_kernel": {
                "name": ("rgf_small_kernel (fused small-block solve, one workgroup per energy)"
                         if fused else "zgemm_reg_kernel (batched complex GEMM, register-direct)"),
                "launches": g["launches"], "flops_per_launch": per_launch,
                "avg_launch_ms": round(avg_ms, 4),
                "per_launch_tflops": round(pl_tflops, 3),
                "per_launch_frac": round(pl_tflops / peak, 4),
                "launches_in_flight": round(g["ms"] / max(g["busy_ms"], 1e-9), 3),
                "busy_union_tflops": round(busy_union, 3),
                "traffic_per_launch": (measured_traffic(dominant_flops)
                                       if rgf.precision == "c128" else None),
                "traffic_launch_flops": dominant_flops}}
        rec["kernels"] = {
            "gemm": {"launches": g["launches"], "sum_ms": round(g["ms"], 2),
                     "busy_ms": round(g["busy_ms"], 2), "flops": g["flops"],
                     "busy_union_tflops": round(busy_union, 3)},
            "inverse": {"launches": inv["launches"], "sum_ms": round(inv["ms"], 2),
                        "busy_ms": round(inv["busy_ms"], 2), "flops": inv["flops"],
                        "busy_union_tflops": round(inv["flops"] / max(inv["busy_ms"], 1e-9) / 1e9,
                                                   3)},
            "wall_ms": round(elapsed * 1e3, 2),
        }
    return rec


def accuracy_outputs(w, rgf, a, sl, sg, n_acc):
    """G^<, G^>, G^R blocks, density and transmission on energies [0, n_acc)."""
    import numpy as np
    import torch

    from quatrex_amd.datastructures import DSDBCOO, empty_btd
    from quatrex_amd.greens_function_solver import transmission

    bs = w.block_sizes
    va, vl, vg = a.stack_view(0, n_acc), sl.stack_view(0, n_acc), sg.stack_view(0, n_acc)
    xl, xg = DSDBCOO.zeros_like(va), DSDBCOO.zeros_like(va)
    xr, _ = empty_btd(bs, (n_acc,), offdiag=False, device=a.device)
    corner = torch.zeros((n_acc, bs[0], bs[-1]), dtype=torch.complex128, device=a.device)
    rgf.selected_solve(va, vl, vg, out=(xl, xg, xr), return_retarded=True, corner=corner)
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    t = transmission(corner, np.broadcast_to(gl, (n_acc,) + gl.shape),
                     np.broadcast_to(gr, (n_acc,) + gr.shape))
    dens = torch.stack([torch.diagonal(xl.blocks[i, i], dim1=-2, dim2=-1)
                        for i in range(len(bs))], 1)
    torch.cuda.synchronize()
    return {"xl": xl, "xg": xg, "xr": xr, "corner": corner, "transmission": t, "density": dens}


def rel(x, y):
    import numpy as np
    import torch

    if isinstance(x, torch.Tensor):
        return float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def accuracy_report(w, test, ref):
    """Errors of one mode against the complex128 GPU outputs."""
    n = w.n_blocks
    worst = {"lesser": 0.0, "greater": 0.0, "retarded_diag": 0.0}
    for key, name in (("xl", "lesser"), ("xg", "greater")):
        for i in range(n):
            for j in (i - 1, i, i + 1):
                if 0 <= j < n:
                    worst[name] = max(worst[name], rel(test[key].blocks[i, j],
                                                       ref[key].blocks[i, j]))
    for i in range(n):
        worst["retarded_diag"] = max(worst["retarded_diag"],
                                     rel(test["xr"].blocks[i, i], ref["xr"].blocks[i, i]))
    t, tr = test["transmission"], ref["transmission"]
    return {"block_rel_max": {k: float(f"{v:.3e}") for k, v in worst.items()},
            "density_rel": float(f"{rel(test['density'], ref['density']):.3e}"),
            "transmission_rel_max": float(f"{float(((t - tr).abs() / tr.abs()).max()):.3e}")}


def oracle_report(w, test, oref):
    """Errors of one mode against the NumPy oracle on the CPU-sample energies."""
    import numpy as np

    from oracle.rgf import transmission as o_transmission

    e0, e1 = oref["energies"]
    n = w.n_blocks
    worst = 0.0
    for i in range(n):
        worst = max(worst, rel(test["xl"].blocks[i, i][e0:e1].cpu().numpy(), oref["diag_l"][i]))
    dens = test["density"][e0:e1].cpu().numpy()
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    t_ref = o_transmission(oref["corner"], gl, gr)
    t = test["transmission"][e0:e1].cpu().numpy()
    return {"energies": [e0, e1], "lesser_diag_block_rel_max": float(f"{worst:.3e}"),
            "density_rel": float(f"{rel(dens, oref['density']):.3e}"),
            "transmission_rel_max": float(f"{np.max(np.abs(t - t_ref) / np.abs(t_ref)):.3e}")}


def obc_workload(args):
    """Open-boundary lead solves batched over energies (one JSON line).

    Lead principal layer of b x b (H00 Hermitian, H01 ~ CN / sqrt(b)),
    M = (E + i eta) I - H over ``--energies`` energies.  Times the surface
    Green's function by Sancho-Rubio (FP64, tolerance 1e-12) and by the
    Spectral OBC on Beyn's contour method (annuli r_i = 1 / r_o for each
    radius of ``--beyn-radius``, 64 nodes, m_0 probe vectors from the largest
    in-annulus eigenvalue count), compares the two, and checks Sancho-Rubio
    against the NumPy oracle on a few energies.  The Beyn legs are repeated
    on a second lead whose onsite spread is ``--evanescent-scale`` times
    wider (fewer modes near the unit circle).
    """
    import numpy as np
    import torch

    from quatrex_amd import _lib
    from quatrex_amd.boundary_conditions import SanchoRubio, Spectral
    from quatrex_amd.nevp import Beyn, Full
    from quatrex_amd.workloads import lead_blocks

    dev = torch.device("cuda:0")
    ne, b = args.energies, args.block_size
    m_ji, m_ii, m_ij, _ = lead_blocks(seed=args.seed, n_energies=ne, block_size=b)
    a_ji, a_ii, a_ij = (torch.as_tensor(x, device=dev) for x in (m_ji, m_ii, m_ij))
    torch.cuda.synchronize()
    log(f"obc workload: {ne} energies, lead layer {b}, resident")

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t0) / args.steps

    sr = SanchoRubio(max_iterations=200, convergence_tol=1e-12)
    timer = _lib.KernelTimer(dev)
    with timer:
        x_sr, t_sr = timed(lambda: sr(a_ii, a_ij, a_ji))
    g = timer.stats["gemm"]
    sr_iterations = sr.iterations
    log(f"sancho-rubio: {t_sr * 1e3:.1f} ms per batch, {sr_iterations} iterations")
    from quatrex_amd.ops import inv, mm

    res_sr = torch.linalg.norm(x_sr - inv(a_ii - mm(a_ji, x_sr, a_ij)), dim=(-2, -1))
    res_sr = (res_sr / torch.linalg.norm(x_sr, dim=(-2, -1))).max().item()

    def spectral_leg(nevp, n, label, lead, x_ref):
        spec = Spectral(nevp)
        l_ii, l_ij, l_ji = lead
        x_sp, t_sp = timed(lambda: spec(l_ii[:n], l_ij[:n], l_ji[:n]))
        diff = torch.linalg.norm(x_sp - x_ref[:n], dim=(-2, -1)) / torch.linalg.norm(
            x_ref[:n], dim=(-2, -1))
        log(f"spectral/{label} on {n} energies: {t_sp * 1e3:.1f} ms per batch, median "
            f"diff vs SR {diff.median().item():.2e}")
        return {"nevp": label, "energies": n, "energies_per_s": round(n / t_sp, 3),
                "rel_diff_vs_sancho_rubio": {"median": float(f"{diff.median().item():.3e}"),
                                             "max": float(f"{diff.max().item():.3e}")}}

    # Beyn's two-moment projection resolves at most b eigenvalues inside the
    # annulus r_i < |w| < r_o.  For each outer radius the probe count m_0 is
    # the largest in-annulus count over a sample of energies (companion
    # eigenvalues, Full) plus a margin, capped at b; the Spectral OBC on Beyn
    # is then compared with Sancho-Rubio on a subset of energies.
    radii = [float(r) for r in str(args.beyn_radius).split(",")]
    n_beyn = min(ne, args.beyn_energies)

    def beyn_legs(lead, x_ref, check=None):
        l_ii, l_ij, l_ji = lead
        n_cnt = min(ne, 8)
        w_full, _ = Full()((l_ji[:n_cnt], l_ii[:n_cnt], l_ij[:n_cnt]))
        mag = w_full.abs()
        legs, annulus = [], []
        for r_o in radii:
            inside = ((mag > 1.0 / r_o) & (mag < r_o)).sum(-1)
            cnt = int(inside.max().item())
            m0 = min(cnt + 8, b)
            annulus.append({"r_o": r_o, "r_i": round(1.0 / r_o, 6), "in_annulus_max": cnt,
                            "in_annulus_min": int(inside.min().item()), "of": 2 * b,
                            "m_0": m0, "resolvable": cnt <= b})
            legs.append(spectral_leg(Beyn(r_o=r_o, r_i=1.0 / r_o, m_0=m0, num_quad_points=64),
                                     n_beyn, f"beyn(r_o={r_o}, m_0={m0}, 64 nodes)", lead,
                                     x_ref))
            if check is not None:
                # the same Spectral OBC on Beyn in the NumPy oracle, 2 energies;
                # both Beyn objects fresh, so both draw the same probe block
                import oracle.nevp as onevp
                import oracle.obc as oobc

                t0 = time.perf_counter()
                x_o = oobc.Spectral(onevp.Beyn(r_o=r_o, r_i=1.0 / r_o, m_0=m0,
                                               num_quad_points=64))(*(x[:2] for x in check))
                got = Spectral(Beyn(r_o=r_o, r_i=1.0 / r_o, m_0=m0, num_quad_points=64))(
                    *(x[:2] for x in lead)).cpu().numpy()
                legs[-1]["oracle"] = {
                    "energies": 2, "cpu_s": round(time.perf_counter() - t0, 1),
                    "rel_diff": float(f"{np.linalg.norm(got - x_o) / np.linalg.norm(x_o):.3e}")}
                log(f"  oracle Spectral/Beyn on 2 energies: rel diff "
                    f"{legs[-1]['oracle']['rel_diff']:.2e}")
        return legs, annulus

    lead = (a_ii, a_ij, a_ji)
    legs, annulus = beyn_legs(lead, x_sr)
    # the Spectral OBC on the full companion eigenproblem shows the mode
    # machinery at this size
    legs.append(spectral_leg(Full(), min(ne, 8), "full", lead, x_sr))
    # A lead whose onsite spread is --evanescent-scale times wider has most
    # of its modes far from the unit circle, few enough inside the annuli
    # for the projection to resolve them: Beyn against Sancho-Rubio there.
    evanescent = None
    if args.evanescent_scale > 0:
        e_ji, e_ii, e_ij, _ = lead_blocks(seed=args.seed, n_energies=n_beyn, block_size=b,
                                          onsite_scale=args.evanescent_scale)
        lead_e = tuple(torch.as_tensor(x, device=dev) for x in (e_ii, e_ij, e_ji))
        x_e = sr(*lead_e)
        legs_e, annulus_e = beyn_legs(lead_e, x_e,
                                      (e_ii, e_ij, e_ji) if args.cpu_budget > 0 else None)
        evanescent = {"onsite_scale": args.evanescent_scale, "energies": n_beyn,
                      "sancho_rubio_iterations": sr.iterations, "spectral": legs_e,
                      "lead_eigenvalues_in_annulus": annulus_e}
    oracle = None
    if args.cpu_budget > 0:
        import oracle.obc as oobc

        k = 2
        t0 = time.perf_counter()
        x_o = oobc.SanchoRubio(max_iterations=200, convergence_tol=1e-12)(
            m_ii[:k], m_ij[:k], m_ji[:k])
        t_o = time.perf_counter() - t0
        oracle = {"energies": k, "sancho_rubio_rel_err": float(
            np.linalg.norm(x_sr[:k].cpu().numpy() - x_o) / np.linalg.norm(x_o)),
            "cpu_s_per_energy": round(t_o / k, 3), "cores": host_cores()}
    flops_sr = sr_iterations * 7 * 8.0 * b ** 3 * ne
    result = {
        "metric": "OBC lead solves: energies/s at 1024 energies x 256 (Sancho-Rubio, Beyn)",
        "value": round(ne / t_sr, 3),
        "unit": "energies/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(t_sr * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"lead surface Green's function, {ne} energies, layer {b}",
                   "energies": ne, "block_size": b, "sancho_rubio_tol": 1e-12,
                   "beyn": {"radii": radii, "quad_points": 64,
                            "m_0": "in-annulus count + 8, at most the layer size"}},
        "sancho_rubio": {"energies_per_s": round(ne / t_sr, 3), "iterations": sr_iterations,
                         "max_fixed_point_residual": float(f"{res_sr:.3e}"),
                         "algorithmic_tflops": round(flops_sr / t_sr / 1e12, 3),
                         "gemm_tflops": round(g["flops"] / max(g["busy_ms"], 1e-9) / 1e9, 3)},
        "spectral": legs,
        "lead_eigenvalues_in_annulus": annulus,
        "evanescent_lead": evanescent,
        "oracle": oracle,
    }
    print(json.dumps(result), flush=True)


def uneven_workload(args):
    """Devices of mixed block size / block count / energy count in one graph.

    24 devices, block size from {64, 128, 256, 416}, block count uniform in
    [16, 128], energies uniform in [64, 1024] scaled by ``--energy-scale``
    (the full mix is ~1.8e16 flops, minutes of GPU time).  Each device keeps
    H, Sigma and its results resident; lanes solve energy chunks side by
    side; each lane's schedule is captured into a HIP graph and the lanes'
    graphs are replayed concurrently, unless a lane's schedule exceeds the
    capture limit (``engine.CAPTURE_BLOCK_STEPS_MAX``), as at full energy
    scale: then the eager launches are timed.
    Tile shape and chunk per block size come from the autotuner.  complex128
    gives ``value``; the mixed mode is timed the same way and compared.
    """
    import numpy as np
    import torch

    from quatrex_amd.engine import Device, UnevenBatch, autotune, uneven_mix
    from quatrex_amd.workloads import make_workload

    dev = torch.device("cuda:0")
    mix = uneven_mix(args.seed, n_devices=args.devices, energy_scale=args.energy_scale)
    t0 = time.perf_counter()
    devices = [Device(make_workload(seed=100 + k, n_energies=e, n_blocks=n, block_size=b), dev,
                      f"dev{k}") for k, (b, n, e) in enumerate(mix)]
    torch.cuda.synchronize()
    log(f"uneven: {len(devices)} devices, {sum(d.n_energies for d in devices)} energies "
        f"resident ({time.perf_counter() - t0:.1f} s)")

    def mem(where):
        free, total = torch.cuda.mem_get_info(dev)
        log(f"memory {where}: torch {torch.cuda.memory_allocated(dev) / 2**30:.1f} GiB "
            f"allocated, {(total - free) / 2**30:.1f} of {total / 2**30:.0f} GiB in use")

    mem("with the devices resident")
    results = {}
    tables = {}
    for prec in ["c128"] + [m for m in args.compare.split(",") if m == "mixed"]:
        t0 = time.perf_counter()
        tables[prec] = autotune([d.b for d in devices], dev, precision=prec,
                                budget_bytes=args.lane_budget << 30, verbose=log)
        log(f"{prec}: autotune {time.perf_counter() - t0:.1f} s: {tables[prec]}")
        mem(f"after the {prec} autotune")
        ub = UnevenBatch(devices, n_lanes=args.lanes, precision=prec, tuning=tables[prec],
                         lane_budget_bytes=args.lane_budget << 30)
        t0 = time.perf_counter()
        ub.run_eager()
        torch.cuda.synchronize()
        t_eager = time.perf_counter() - t0
        log(f"{prec}: eager {t_eager:.2f} s")
        mem(f"after the {prec} eager run")
        graphed = ub.can_capture()
        if graphed:
            t0 = time.perf_counter()
            ub.capture()
            torch.cuda.synchronize()
            log(f"{prec}: captured ({time.perf_counter() - t0:.1f} s)")
            ub.replay()
            ub.check()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                ub.replay()
            torch.cuda.synchronize()
            t_graph = (time.perf_counter() - t0) / args.steps
            ub.check()
            log(f"{prec}: graph replay {t_graph:.2f} s")
        else:
            # schedules above the capture limit run eagerly only
            log(f"{prec}: {max(ub.lane_block_steps())} block steps per lane, above the "
                f"capture limit; timing eager launches")
            t_graph = t_eager
        results[prec] = {
            "energies_per_s": round(ub.total_energies / t_graph, 3),
            "eager_energies_per_s": round(ub.total_energies / t_eager, 3),
            "launch": "graph" if graphed else "eager",
            "lane_block_steps": ub.lane_block_steps(),
            "step_s": round(t_graph, 3),
            "graph_s": round(t_graph, 3) if graphed else None, "eager_s": round(t_eager, 3),
            "required_tflops": round(ub.total_flops / t_graph / 1e12, 3),
            "lane_load_balance": round(min(ub.lane_flops) / max(ub.lane_flops), 3),
            "items": sum(len(l.items) for l in ub.lanes),
            "density": [d.density.clone() for d in devices],
            "transmission": [d.transmission.clone() for d in devices],
        }
        ub.close()
        del ub
        gc.collect()
        torch.cuda.empty_cache()
        mem(f"after the {prec} leg")
    prim = results["c128"]
    out_modes = {}
    for prec, r in results.items():
        rec = {k: v for k, v in r.items() if k not in ("density", "transmission")}
        if prec != "c128":
            dn = sum(float(torch.linalg.norm(a - b) ** 2)
                     for a, b in zip(r["density"], prim["density"])) ** 0.5
            dd = sum(float(torch.linalg.norm(b) ** 2) for b in prim["density"]) ** 0.5
            tr = max(float(((a - b).abs() / b.abs().clamp_min(1e-300)).max())
                     for a, b in zip(r["transmission"], prim["transmission"]))
            rec["error_vs_c128"] = {"density_rel": float(f"{dn / dd:.3e}"),
                                    "transmission_rel_max": float(f"{tr:.3e}")}
            rec["speedup_vs_c128"] = round(r["energies_per_s"] / prim["energies_per_s"], 3)
        out_modes[prec] = rec
    result = {
        "metric": "uneven batched RGF: energies/s over a 24-device size mix, "
                  + ("HIP-graph replay" if prim["launch"] == "graph" else "eager launches"),
        "value": prim["energies_per_s"], "unit": "energies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1,
        "ms_per_step": round(prim["step_s"] * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "uneven device mix", "devices": [list(m) for m in mix],
                   "energy_scale": args.energy_scale, "lanes": args.lanes,
                   "total_energies": sum(d.n_energies for d in devices)},
        "modes": out_modes, "autotune": {p: {str(k): v for k, v in t.items()}
                                         for p, t in tables.items()},
    }
    print(json.dumps(result), flush=True)


def serve_workload(args):
    """Serving loop with Poisson arrivals of SCF-iteration requests.

    Devices from the uneven size mix stay resident; each request carries
    128-1024 energies (scaled by ``--energy-scale``) and updated Sigma
    blocks.  For each arrival rate of ``--rates`` the loop serves
    ``--requests`` requests and reports p50/p99 latency (arrival to
    completion) and energies/s; ``value`` is the energies/s at the highest
    rate.
    """
    import torch

    from quatrex_amd.engine import uneven_mix
    from quatrex_amd.serving import Server, poisson_requests
    from quatrex_amd.workloads import make_workload

    dev = torch.device("cuda:0")
    mix = uneven_mix(args.seed, n_devices=args.devices, energy_scale=1.0)
    t0 = time.perf_counter()
    works = {f"dev{k}": make_workload(seed=100 + k, n_energies=1, n_blocks=n, block_size=b)
             for k, (b, n, _) in enumerate(mix)}
    srv = Server(dev, n_lanes=args.lanes, hbm_budget_bytes=args.hbm_budget << 30,
                 chunk=args.chunk, lane_budget_bytes=args.lane_budget << 30)
    for k, w in works.items():
        srv.add_device(k, w)
    torch.cuda.synchronize()
    log(f"serve: {len(works)} devices resident ({srv.resident / 2**30:.1f} GiB, "
        f"{time.perf_counter() - t0:.1f} s)")
    lo, hi = max(1, round(128 * args.energy_scale)), max(2, round(1024 * args.energy_scale))
    # warm-up: one request per device size class
    warm = poisson_requests(list(works), 1e6, args.lanes, (lo, lo), seed=7, devices=works)
    srv.run(warm, realtime=False)
    rates = [float(r) for r in args.rates.split(",")]
    per_rate = []
    for k, rate in enumerate(rates):
        reqs = poisson_requests(list(works), rate, args.requests, (lo, hi), seed=args.seed + k,
                                devices=works)
        st = srv.run(reqs, realtime=True)
        log(f"serve rate {rate}/s: {st['energies_per_s']:.1f} energies/s "
            f"({st['required_tflops']:.1f} TF/s required), p50 "
            f"{st['latency_ms']['p50']:.0f} ms, p99 {st['latency_ms']['p99']:.0f} ms")
        per_rate.append(dict(rate_per_s=rate, **{k2: (round(v, 3) if isinstance(v, float) else v)
                                                 for k2, v in st.items()}))
    top = per_rate[-1]
    result = {
        "metric": "serving: energies/s and p50/p99 request latency, Poisson SCF requests",
        "value": round(top["energies_per_s"], 3), "unit": "energies/s", "n_gpus": 1,
        "steps": len(rates), "warmup": 1, "ms_per_step": round(top["makespan_s"] * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serving loop", "devices": [list(m[:2]) for m in mix],
                   "energies_per_request": [lo, hi], "lanes": args.lanes,
                   "hbm_budget_gib": args.hbm_budget, "requests_per_rate": args.requests},
        "latency_ms": top["latency_ms"], "rates": per_rate,
    }
    print(json.dumps(result), flush=True)


def kernel_probe(args):
    """Standalone launches for counter runs: a plain device copy of known
    size (the calibration for HBM byte counters), then the dominant GEMM and
    the batched inverse at the benchmark's sub-batch shape."""
    import ctypes

    import torch

    from quatrex_amd import _lib
    from quatrex_amd.ops import gemm, inv

    dev = torch.device("cuda:0")
    copy_bytes = 1 << 30
    src = torch.ones(copy_bytes // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    b, batch = args.block_size, args.probe_batch
    g = torch.Generator(device=dev).manual_seed(0)
    a = torch.randn((batch, b, b), dtype=torch.complex128, device=dev, generator=g)
    bm = torch.randn((batch, b, b), dtype=torch.complex128, device=dev, generator=g)
    c = torch.empty_like(a)
    for _ in range(5):
        gemm(a, bm, out=c)
    a = a + b * torch.eye(b, dtype=torch.complex128, device=dev)
    for _ in range(3):
        inv(a, check=False)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": copy_bytes, "copy_launches": 3,
                      "gemm": {"m": b, "n": b, "k": b, "batch": batch, "launches": 5,
                               "compulsory_bytes": 16 * 3 * b * b * batch,
                               "flops": 8.0 * b ** 3 * batch},
                      "inverse": {"n": b, "batch": batch, "launches": 3}}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="c128", choices=list(PEAK_TFLOPS))
    ap.add_argument("--compare", default="c128-3m,mixed,oz2-12,bf16x7",
                    help="comma-separated modes timed beside the primary one ('' for none)")
    ap.add_argument("--compare-steps", type=int, default=3,
                    help="timed steps of each compared mode (at most --steps; 1 warm-up)")
    ap.add_argument("--energies", type=int, default=512)
    ap.add_argument("--blocks", type=int, default=64)
    ap.add_argument("--block-size", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--accuracy-energies", type=int, default=8,
                    help="energies on which modes are compared (0 disables)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU work for the baseline sample (0 disables)")
    ap.add_argument("--no-timer", action="store_true",
                    help="do not bracket kernels with HIP events (no roofline)")
    ap.add_argument("--probe-batch", type=int, default=32,
                    help="probe: matrices per launch (the RGF sub-batch: 128 energies / 4)")
    ap.add_argument("--workload", default="rgf",
                    choices=["rgf", "obc", "uneven", "serve", "probe"],
                    help="rgf: the selected solve (default); obc: lead surface Green's "
                         "functions (use --energies 1024 --block-size 256); uneven: mixed "
                         "device sizes in one graph launch")
    ap.add_argument("--devices", type=int, default=24, help="uneven: number of devices")
    ap.add_argument("--energy-scale", type=float, default=1 / 16,
                    help="uneven: energies per device scaled from [64, 1024]")
    ap.add_argument("--lanes", type=int, default=4, help="uneven: concurrent lanes")
    ap.add_argument("--lane-budget", type=int, default=24,
                    help="uneven/serve: HBM per lane in GiB (bounds energies per chunk)")
    ap.add_argument("--rates", default="1,2,4", help="serve: arrival rates (requests/s)")
    ap.add_argument("--requests", type=int, default=16, help="serve: requests per rate")
    ap.add_argument("--hbm-budget", type=int, default=200, help="serve: HBM budget in GiB")
    ap.add_argument("--beyn-radius", default="1.5,2,3",
                    help="outer contour radii of the Beyn legs (inner = 1 / radius)")
    ap.add_argument("--beyn-energies", type=int, default=32,
                    help="obc: energies of each Spectral-on-Beyn leg")
    ap.add_argument("--evanescent-scale", type=float, default=8.0,
                    help="obc: onsite spread of the second lead of the Beyn legs (0: none)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Energies shard with no communication; "
              f"run one process per GPU on disjoint energy ranges.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "obc":
        obc_workload(args)
        return
    if args.workload == "uneven":
        uneven_workload(args)
        return
    if args.workload == "serve":
        serve_workload(args)
        return
    if args.workload == "probe":
        kernel_probe(args)
        return

    import torch

    from quatrex_amd.datastructures import DSDBCOO, empty_btd
    from quatrex_amd.greens_function_solver import RGF
    from quatrex_amd.workloads import make_workload, required_flops

    dev = torch.device("cuda:0")
    E, N, b, chunk = args.energies, args.blocks, args.block_size, args.chunk
    if E % chunk:
        raise SystemExit("--energies must be a multiple of --chunk")
    modes = [args.precision] + [m for m in args.compare.split(",") if m and m != args.precision]
    log(f"workload: {E} energies, {N} blocks of {b}, modes {modes}, chunk {chunk}")
    w = make_workload(seed=args.seed, n_energies=E, n_blocks=N, block_size=b)
    t0 = time.perf_counter()
    a, sl, sg = w.assemble(dev)
    torch.cuda.synchronize()
    log(f"inputs resident: A {a.data.numel() * 16 / 1e9:.1f} GB, sigma "
        f"{2 * sl.data.numel() * 16 / 1e9:.1f} GB ({time.perf_counter() - t0:.1f} s)")
    xl = DSDBCOO.zeros_like(a.stack_view(0, chunk))
    xg = DSDBCOO.zeros_like(a.stack_view(0, chunk))
    xr, _ = empty_btd(w.block_sizes, (chunk,), offdiag=False, device=dev)
    views = [(a.stack_view(c, c + chunk), sl.stack_view(c, c + chunk), sg.stack_view(c, c + chunk))
             for c in range(0, E, chunk)]

    legs = {}
    for m in modes:
        rgf = RGF(max_batch_size=chunk, precision=m)
        peak = PEAK_TFLOPS[m]
        # the RGF splits a call into 4 concurrent sub-batches (QX_RGF_STREAMS)
        dominant = 8.0 * b ** 3 * ((chunk + 3) // 4)
        # compared modes: a bounded number of steps, so that the default run
        # stays within minutes
        primary = m == args.precision
        steps = args.steps if primary else max(1, min(args.steps, args.compare_steps))
        warm = args.warmup if primary else min(args.warmup, 1)
        # blocks <= 32, <= 16 of them: one fused kernel launch per call
        fused = (N <= 16 and b <= 32 and m in ("c128", "c128-3m")
                 and os.environ.get("QX_RGF_FUSED", "1") != "0")
        legs[m] = timed_leg(rgf, views, (xl, xg, xr), steps, warm, dev,
                            not args.no_timer, peak, dominant, E * required_flops(N, b), fused)
        legs[m]["value"] = E * steps / legs[m]["elapsed"]
        legs[m]["steps"] = steps
        log(f"{m}: {legs[m]['value']:.1f} energies/s")
    del xl, xg, xr

    n_acc = min(args.accuracy_energies, E)
    acc = {}
    if n_acc > 0:
        for m in set(modes) | {"c128"}:
            acc[m] = accuracy_outputs(w, RGF(max_batch_size=n_acc, precision=m), a, sl, sg, n_acc)
        log("accuracy outputs done")
    del a, sl, sg, views
    torch.cuda.empty_cache()

    oref, baseline = None, None
    if args.cpu_budget > 0:
        log("cpu baseline ...")
        oref, baseline = cpu_baseline(w, args.cpu_budget, max(2, n_acc))
        if n_acc < 2:
            oref = None

    prim = legs[args.precision]
    result = {
        "metric": METRIC,
        "value": round(prim["value"], 3),
        "unit": "energies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * prim["elapsed"] / args.steps, 2),
        "step_ms": prim["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64+i8" if args.precision.startswith("oz2-") else
                  "f64+bf16" if args.precision.startswith("bf16x") else
                  {"c128": "f64", "c128-3m": "f64", "c64": "f32",
                   "mixed": "f32+f64"}[args.precision]),
        "data": "synthetic",
        "config": {"workload": f"RGF selected_solve, {E} energies, {N} blocks of {b}, "
                               f"complex128 storage, {args.precision} arithmetic",
                   "energies": E, "n_blocks": N, "block_size": b, "precision": args.precision,
                   "energies_per_call": chunk},
        "latency_ms": dict(prim["latency_ms"], per=f"selected_solve call of {chunk} energies"),
        "roofline": prim.get("roofline"),
        "kernels": prim.get("kernels"),
        "cpu_baseline": baseline,
    }
    modes_out = {}
    for m in modes:
        r = {"energies_per_s": round(legs[m]["value"], 3), "steps": legs[m]["steps"],
             "speedup_vs_primary": round(legs[m]["value"] / prim["value"], 3),
             "latency_ms": legs[m]["latency_ms"]}
        if "roofline" in legs[m]:
            r["solve_tflops"] = legs[m]["roofline"]["achieved"]
            r["solve_frac"] = legs[m]["roofline"]["frac"]
            r["gemm_busy_union_tflops"] = legs[m]["kernels"]["gemm"]["busy_union_tflops"]
            r["inverse_busy_ms"] = legs[m]["kernels"]["inverse"]["busy_ms"]
        if acc and m != "c128":
            r["error_vs_c128_gpu"] = accuracy_report(w, acc[m], acc["c128"])
        if acc and oref is not None:
            r["error_vs_oracle"] = oracle_report(w, acc[m], oref)
        modes_out[m] = r
    result["modes"] = modes_out
    if acc:
        result["accuracy_energies"] = n_acc
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
