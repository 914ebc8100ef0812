"""Uneven batched RGF: many devices of different shapes in one launch.

A *device* here is one transport problem: block-tridiagonal Hamiltonian
(N blocks of b), boundary self-energies folded into the first and last
diagonal blocks, block-diagonal Sigma^< / Sigma^> and its energy grid.  Its
Hamiltonian, self-energies and results (electron density Im diag G^<,
transmission T(E)) stay resident in HBM; the energy-dependent system matrix
and the Green's-function blocks of an energy chunk live only in the buffers
of the *lane* that solves it.

Execution model (MI355X-first):

- a lane = one library handle (own RGF workspace) + one HIP stream; lanes
  run side by side on the GPU, so small devices (b = 64) share the chip
  with large ones instead of leaving most CUs idle;
- work items (device, energy chunk) are assigned to lanes
  longest-processing-time first, by algorithmic flops, so lanes finish
  together;
- one item = system-matrix assembly kernel (A = E I - H - Sigma_B on the
  nnz pattern) -> RGF selected solve with fused density and corner-block
  outputs -> transmission;
- ``UnevenBatch.capture()`` records each lane's whole schedule (up to
  ``CAPTURE_BLOCK_STEPS_MAX`` block steps) into one HIP
  graph; ``replay()`` launches the lanes' graphs on their streams (one
  launch per lane), so the thousands of small kernel launches of small
  devices cost no host time;
- ``autotune`` picks, per block size, the GEMM kernel (the register-direct
  default or an LDS-staged 64 x 64 or 128 x 128 tile) and
  the energies per call by timing probe solves; its table is cached.

The per-device computation is the reference's RGF.selected_solve
(rgf.py:115-414) on the system matrix of electron/solver.py:390-407.
"""

from __future__ import annotations

import gc
import json
import os
import threading
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from quatrex_amd import _lib
from quatrex_amd.datastructures import DSDBCOO, btd_from_blocks, btd_pattern
from quatrex_amd.greens_function_solver import RGF, transmission
from quatrex_amd.workloads import Workload, make_workload, required_flops


def _f(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128, device=dev)


class Device:
    """Resident data and results of one device (built from a ``Workload``)."""

    def __init__(self, w: Workload, device, name: str = ""):
        dev = torch.device(device)
        self.name = name
        self.device = dev
        self.block_sizes = [int(b) for b in w.block_sizes]
        self.n_blocks = len(self.block_sizes)
        self.n_orb = int(sum(self.block_sizes))
        self.n_energies = len(w.energies)
        self.energies = torch.as_tensor(np.asarray(w.energies, dtype=np.complex128), device=dev)
        n = self.n_blocks
        # static part of A on its BTD pattern: -(H + Sigma_B)
        diag = []
        for i in range(n):
            h = np.array(w.h_diag[i], dtype=np.complex128)
            if i == 0:
                h = h + w.sigma_b_left
            if i == n - 1:
                h = h + w.sigma_b_right
            diag.append(_f(-h, dev))
        upper = [_f(-u, dev) for u in w.h_upper]
        lower = [_f(-np.conj(u.T), dev) for u in w.h_upper]
        static = btd_from_blocks(diag, upper, lower, global_stack_shape=(), device=dev)
        self.rows, self.cols = static.rows, static.cols
        self.nnz = static.nnz
        self.t_static = static.data.reshape(-1).contiguous()
        self.s_static = torch.as_tensor((self.rows == self.cols).astype(np.complex128), device=dev)
        drows, dcols, _ = btd_pattern(self.block_sizes, offdiag=False)
        self.drows, self.dcols = drows, dcols
        self.nnz_d = len(drows)
        sl = btd_from_blocks([_f(x, dev) for x in w.sl_diag], None, None,
                             global_stack_shape=(), device=dev)
        sg = btd_from_blocks([_f(x, dev) for x in w.sg_diag], None, None,
                             global_stack_shape=(), device=dev)
        self.sl = sl.data.reshape(1, -1).contiguous()
        self.sg = sg.data.reshape(1, -1).contiguous()
        self.gamma_l = _f(1j * (w.sigma_b_left - np.conj(w.sigma_b_left.T)), dev)
        self.gamma_r = _f(1j * (w.sigma_b_right - np.conj(w.sigma_b_right.T)), dev)
        self.density = torch.zeros((self.n_energies, self.n_orb), dtype=torch.float64, device=dev)
        self.transmission = torch.zeros(self.n_energies, dtype=torch.float64, device=dev)
        # pattern metadata (block slices, dense flags) shared by every
        # container of this device
        self._meta = {True: ({}, {}), False: ({}, {})}
        self._sigma_views: dict = {}
        self._ws_bytes: dict = {}

    def container(self, data: torch.Tensor, diag_only: bool) -> DSDBCOO:
        r, c = (self.drows, self.dcols) if diag_only else (self.rows, self.cols)
        m = DSDBCOO(data, r, c, self.block_sizes, (data.shape[0],))
        m._slices, m._dense = self._meta[diag_only]
        return m

    def sigma_views(self, c: int):
        """Sigma^<, Sigma^> of a chunk of c energies (stride-0 broadcast)."""
        v = self._sigma_views.get(c)
        if v is None:
            v = (self.container(self.sl.expand(c, self.nnz_d), True),
                 self.container(self.sg.expand(c, self.nnz_d), True))
            self._sigma_views[c] = v
        return v

    @property
    def b(self) -> int:
        return max(self.block_sizes)

    def flops_per_energy(self) -> float:
        return required_flops(self.n_blocks, self.b)

    def pool_bytes_per_energy(self) -> int:
        """Lane pool bytes per energy of a chunk (A, two block-diagonal
        outputs, the corner block) plus the transient Gamma and transmission
        temporaries of the chunk (six b0 x b0 blocks)."""
        b0 = self.block_sizes[0]
        return 16 * (self.nnz + 2 * self.nnz_d + self.block_sizes[-1] * b0 + 6 * b0 * b0)

    def workspace_bytes(self, n_energies: int, precision: str = "c128") -> int:
        """Handle workspace of a chunk of ``n_energies`` (qx_workspace_query)."""
        key = (n_energies, precision)
        v = self._ws_bytes.get(key)
        if v is None:
            v = _lib.workspace_bytes(_lib.handle(self.device), self.block_sizes, n_energies,
                                     precision)
            self._ws_bytes[key] = v
        return v

    def lane_bytes_per_energy(self, precision: str = "c128") -> int:
        """HBM a lane needs per energy of a chunk: its pool share and the RGF
        workspace (exact per energy: the workspace is linear in the chunk)."""
        return self.pool_bytes_per_energy() + self.workspace_bytes(1, precision)

    def resident_bytes(self) -> int:
        return 16 * (2 * self.nnz + 2 * self.nnz_d) + 24 * self.n_energies * self.n_orb


def make_device(seed: int, n_energies: int, n_blocks: int, block_size: int, device,
                name: str = "") -> Device:
    return Device(make_workload(seed=seed, n_energies=n_energies, n_blocks=n_blocks,
                                block_size=block_size), device, name)


class Lane:
    """One handle (own workspace) + one stream; buffers are views of one pool."""

    def __init__(self, device, index: int = 0):
        self.device = torch.device(device)
        self.index = index
        self.h = _lib.new_handle(self.device)
        _lib.set_tuning(self.h, streams=1, gemm_tile=0)
        self.stream = torch.cuda.Stream(self.device)  # default priority
        self.items: list = []
        self._pool = torch.empty(0, dtype=torch.complex128, device=self.device)
        self._views: dict = {}

    def reserve(self, elems: int) -> None:
        if elems > self._pool.numel():
            self._views.clear()
            self._pool = torch.empty(elems, dtype=torch.complex128, device=self.device)

    def release(self) -> None:
        """Give back the pool and the handle's workspace (the handle is
        recreated; its workspace grows again on the next call)."""
        self.stream.synchronize()
        self._views.clear()
        self._pool = torch.empty(0, dtype=torch.complex128, device=self.device)
        _lib.load().qx_destroy(self.h)
        self.h = _lib.new_handle(self.device)
        _lib.set_tuning(self.h, streams=1, gemm_tile=0)

    def close(self) -> None:
        """Free the pool and destroy the handle (with its workspace)."""
        self.stream.synchronize()
        self._views.clear()
        self._pool = torch.empty(0, dtype=torch.complex128, device=self.device)
        if self.h is not None:
            _lib.load().qx_destroy(self.h)
            self.h = None

    def views(self, d: Device, c: int):
        """(A, X^<, X^>, corner) containers for a chunk of c energies of d."""
        key = (id(d), c)
        v = self._views.get(key)
        if v is not None:
            return v
        need = c * (d.nnz + 2 * d.nnz_d + d.block_sizes[-1] * d.block_sizes[0])
        self.reserve(need)
        p, o = self._pool, 0
        a = p[o:o + c * d.nnz].view(c, d.nnz)
        o += c * d.nnz
        xl = p[o:o + c * d.nnz_d].view(c, d.nnz_d)
        o += c * d.nnz_d
        xg = p[o:o + c * d.nnz_d].view(c, d.nnz_d)
        o += c * d.nnz_d
        nl, n0 = d.block_sizes[-1], d.block_sizes[0]
        corner = p[o:o + c * nl * n0].view(c, n0, nl)
        v = (d.container(a, False), d.container(xl, True), d.container(xg, True), corner)
        self._views[key] = v
        return v

    def sync_check(self, what: str) -> None:
        with _lib.use_handle(self.h), torch.cuda.stream(self.stream):
            _lib.sync_check(self.device, what)


class Job:
    """Energies, self-energies and result buffers of one solve of a device.

    The device's own grid by default; a serving request brings its energies
    and (optionally) updated Sigma^< / Sigma^> blocks.
    """

    def __init__(self, d: Device, energies=None, sigma_lesser=None, sigma_greater=None):
        dev = d.device
        self.device = d
        if energies is None:
            self.energies, self.density, self.transmission = (d.energies, d.density,
                                                              d.transmission)
        else:
            self.energies = torch.as_tensor(np.asarray(energies, dtype=np.complex128), device=dev)
            n = self.energies.shape[0]
            self.density = torch.zeros((n, d.n_orb), dtype=torch.float64, device=dev)
            self.transmission = torch.zeros(n, dtype=torch.float64, device=dev)
        self.sl, self.sg = d.sl, d.sg
        if sigma_lesser is not None:
            self.sl = btd_from_blocks([_f(x, dev) for x in sigma_lesser], None, None,
                                      global_stack_shape=(), device=dev).data.reshape(1, -1)
        if sigma_greater is not None:
            self.sg = btd_from_blocks([_f(x, dev) for x in sigma_greater], None, None,
                                      global_stack_shape=(), device=dev).data.reshape(1, -1)
        self._views: dict = {}

    @property
    def n_energies(self) -> int:
        return int(self.energies.shape[0])

    def sigma_views(self, c: int):
        if self.sl is self.device.sl and self.sg is self.device.sg:
            return self.device.sigma_views(c)
        v = self._views.get(c)
        if v is None:
            d = self.device
            v = (d.container(self.sl.expand(c, d.nnz_d), True),
                 d.container(self.sg.expand(c, d.nnz_d), True))
            self._views[c] = v
        return v

    def extra_bytes(self) -> int:
        d = self.device
        own = 0 if self.energies is d.energies else 24 * self.n_energies * (d.n_orb + 1)
        if self.sl is not d.sl:
            own += 32 * d.nnz_d
        return own


def solve_item(lane: Lane, d: Device, e0: int, e1: int, precision: str = "c128",
               tune: dict | None = None, job: Job | None = None) -> None:
    """Assemble and solve energies [e0, e1) of d (or of ``job`` on d) on the
    lane; the current stream must be the lane's."""
    c = e1 - e0
    a, xl, xg, corner = lane.views(d, c)
    h = lane.h
    if tune:
        _lib.set_tuning(h, 1, int(tune.get("gemm_tile", 0)))
    energies = job.energies if job else d.energies
    st = ctypes_stream(lane)
    rc = _lib.load().qx_assemble_system_matrix(
        h, st, c, d.nnz, energies[e0:e1].data_ptr(), d.s_static.data_ptr(),
        d.t_static.data_ptr(), None, 0, a.data.data_ptr(), d.nnz)
    _lib.check(rc, h, "assemble_system_matrix")
    sl, sg = job.sigma_views(c) if job else d.sigma_views(c)
    density = job.density if job else d.density
    trans = job.transmission if job else d.transmission
    RGF(max_batch_size=c, precision=precision).selected_solve(
        a, sl, sg, out=(xl, xg), corner=corner, density=density[e0:e1])
    gl = d.gamma_l.expand(c, *d.gamma_l.shape).contiguous()
    gr = d.gamma_r.expand(c, *d.gamma_r.shape).contiguous()
    trans[e0:e1].copy_(transmission(corner, gl, gr))


def ctypes_stream(lane: Lane):
    import ctypes

    return ctypes.c_void_p(lane.stream.cuda_stream)


@dataclass
class Item:
    device: int
    e0: int
    e1: int
    flops: float = 0.0


def plan_items(devices, n_lanes: int):
    """Chunks of every device's energies, assigned to ``n_lanes`` lanes.

    ``devices`` holds (n_energies, flops per energy, largest chunk) per
    device.  A chunk is at most that size and at most a quarter of one
    lane's fair share of the flops, so a device that dominates the mix is
    spread over several lanes; chunks go, largest first, to the least loaded
    lane (longest-processing-time rule).  Returns the item lists per lane and
    the flops per lane.
    """
    items = []
    total = sum(n * f for n, f, _ in devices)
    share = total / (4 * n_lanes)
    for k, (n, f, chunk) in enumerate(devices):
        c = max(1, min(int(chunk), max(1, int(share // f)) if f > 0 else int(chunk)))
        for e0 in range(0, n, c):
            e1 = min(e0 + c, n)
            items.append(Item(k, e0, e1, (e1 - e0) * f))
    items.sort(key=lambda it: -it.flops)
    load = [0.0] * n_lanes
    assign = [[] for _ in range(n_lanes)]
    for it in items:
        k = int(np.argmin(load))
        assign[k].append(it)
        load[k] += it.flops
    return assign, load


# Largest per-lane schedule (block steps) recorded into one HIP graph: ~800
# block steps are ~10^5 kernel nodes per graph.  A schedule of ~2100 block
# steps per lane ended the process with a host segmentation fault while it
# was recorded, cause not established, so larger schedules are not captured.
CAPTURE_BLOCK_STEPS_MAX = 800


@dataclass
class UnevenBatch:
    """Devices of mixed shapes solved together on ``n_lanes`` lanes."""

    devices: list
    n_lanes: int = 4
    precision: str = "c128"
    tuning: dict = field(default_factory=dict)  # block size -> {"gemm_tile", "chunk"}
    default_chunk: int = 16
    lane_budget_bytes: int = 24 << 30

    def __post_init__(self):
        self.lanes = [Lane(self.devices[0].device, i) for i in range(self.n_lanes)]
        self.graph = None
        self.plan()

    def chunk_for(self, d: Device) -> int:
        c = int(self.tuning.get(d.b, {}).get("chunk", self.default_chunk))
        cap = max(1, self.lane_budget_bytes // d.lane_bytes_per_energy(self.precision))
        return max(1, min(c, cap, d.n_energies))

    def plan(self) -> None:
        """Split every device's energies into chunks and assign them to lanes
        (`plan_items`)."""
        assign, load = plan_items(
            [(d.n_energies, d.flops_per_energy(), self.chunk_for(d)) for d in self.devices],
            len(self.lanes))
        for lane, items in zip(self.lanes, assign):
            lane.items = items
        self.lane_flops = load
        self.graph = None

    def _run_lane(self, lane: Lane) -> None:
        with _lib.use_handle(lane.h), torch.cuda.stream(lane.stream):
            for it in lane.items:
                d = self.devices[it.device]
                solve_item(lane, d, it.e0, it.e1, self.precision, self.tuning.get(d.b))

    def run_eager(self) -> None:
        """Every lane from its own host thread, concurrently on the GPU."""
        cur = torch.cuda.current_stream(self.lanes[0].device)
        for lane in self.lanes:
            lane.stream.wait_stream(cur)
        errors = []

        def work(lane):
            try:
                self._run_lane(lane)
            except Exception as e:  # surfaced after the join
                errors.append(e)

        threads = [threading.Thread(target=work, args=(lane,)) for lane in self.lanes]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for lane in self.lanes:
            cur.wait_stream(lane.stream)
        if errors:
            raise errors[0]
        for lane in self.lanes:
            lane.sync_check("UnevenBatch")

    def close(self) -> None:
        """Free the lanes' pools, graphs and library handles (with their RGF
        workspaces); the batch cannot run afterwards."""
        self.graph = None
        for lane in self.lanes:
            lane.close()

    def lane_block_steps(self) -> list[int]:
        """Block steps (sum of the block counts of its items) per lane: the
        number of kernel nodes a lane's graph records grows with it."""
        return [sum(self.devices[it.device].n_blocks for it in lane.items)
                for lane in self.lanes]

    def can_capture(self) -> bool:
        return max(self.lane_block_steps()) <= CAPTURE_BLOCK_STEPS_MAX

    def capture(self) -> None:
        """Record each lane's schedule into its own HIP graph.

        Runs the schedule eagerly first (sizes every workspace and buffer so
        nothing is allocated while recording).  One graph per lane, replayed
        on the lane's stream: graph launches on different streams run
        concurrently, where the branches of a single multi-stream graph are
        not guaranteed to.  Refused (RuntimeError) when a lane's schedule
        holds more than ``CAPTURE_BLOCK_STEPS_MAX`` block steps; run such a
        schedule with ``run_eager``.
        """
        steps = max(self.lane_block_steps())
        if steps > CAPTURE_BLOCK_STEPS_MAX:
            raise RuntimeError(
                f"UnevenBatch.capture: a lane holds {steps} block steps, above the "
                f"capture limit of {CAPTURE_BLOCK_STEPS_MAX}; use run_eager()")
        self.run_eager()
        torch.cuda.synchronize(self.lanes[0].device)
        self.graph = []
        for lane in self.lanes:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=lane.stream):
                self._run_lane(lane)
            self.graph.append(g)

    def replay(self) -> None:
        """Launch every lane's graph on its stream, one launch per lane, each
        from its own host thread (a graph launch enqueues its kernel nodes
        from the host, so concurrent launches keep the lanes' queues fed)."""
        if self.graph is None:
            raise RuntimeError("capture() first")
        cur = torch.cuda.current_stream(self.lanes[0].device)
        for lane in self.lanes:
            lane.stream.wait_stream(cur)

        def launch(lane, g):
            with torch.cuda.stream(lane.stream):
                g.replay()

        threads = [threading.Thread(target=launch, args=(lane, g))
                   for lane, g in zip(self.lanes, self.graph)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for lane in self.lanes:
            cur.wait_stream(lane.stream)

    def check(self) -> None:
        torch.cuda.synchronize(self.lanes[0].device)
        for lane in self.lanes:
            lane.sync_check("UnevenBatch replay")

    @property
    def total_energies(self) -> int:
        return sum(d.n_energies for d in self.devices)

    @property
    def total_flops(self) -> float:
        return sum(d.n_energies * d.flops_per_energy() for d in self.devices)


# --------------------------------------------------------------------------
# autotuner

# GEMM kernel choices (qx_set_tuning codes): 0 = the library default
# (register-direct, tile shape by product), 64 / 128 = the LDS-staged
# 64 x 64 / 128 x 128 tiles; the first is the default
TILES = (0, 64, 128)
CHUNKS = (8, 16, 32, 64, 128, 256)
# A tile other than the default must beat it by this fraction in the probe:
# the probe runs one lane alone, the batch runs several side by side, and a
# single-lane win of a few percent has turned into a loss of ~12 % there.
TILE_MARGIN = 0.03
PROBE_REPEATS = 3  # timings per candidate (2 solves each); the fastest counts


def _cache_path() -> Path:
    base = os.environ.get("QX_AUTOTUNE_CACHE")
    if base:
        return Path(base)
    return Path.home() / ".cache" / "quatrex_amd" / "autotune.json"


def autotune(block_sizes, device, precision: str = "c128", n_blocks: int = 6,
             budget_bytes: int = 24 << 30, cache: bool = True, verbose=None) -> dict:
    """Per block size: GEMM tile shape and energies per call, by measurement.

    For every block size a probe device of ``n_blocks`` blocks is solved for
    each (tile, chunk) candidate that fits ``budget_bytes``; the fastest
    time per energy (best of ``PROBE_REPEATS`` timings) wins, a non-default
    tile only by more than ``TILE_MARGIN``.  Results are cached per GPU model and precision
    (``QX_AUTOTUNE_CACHE`` or ~/.cache/quatrex_amd/autotune.json).
    """
    dev = torch.device(device)
    gpu = torch.cuda.get_device_name(dev)
    path = _cache_path()
    table = {}
    if cache and path.exists():
        try:
            table = json.loads(path.read_text())
        except (OSError, ValueError):
            table = {}
    key = f"{gpu}|{precision}|tiles={','.join(map(str, TILES))}"
    known = table.get(key, {})
    out = {}
    lane = None
    for b in sorted(set(int(x) for x in block_sizes)):
        if str(b) in known:
            out[b] = known[str(b)]
            continue
        if lane is None:
            lane = Lane(dev)
        best = None
        for chunk in CHUNKS:
            d = make_device(seed=b, n_energies=chunk, n_blocks=n_blocks, block_size=b,
                            device=dev)
            if chunk * d.lane_bytes_per_energy(precision) > budget_bytes:
                continue
            for tile in TILES:
                tune = {"gemm_tile": tile, "chunk": chunk}
                per_e = float("inf")
                with _lib.use_handle(lane.h), torch.cuda.stream(lane.stream):
                    solve_item(lane, d, 0, chunk, precision, tune)  # warm-up
                    lane.stream.synchronize()
                    for _ in range(PROBE_REPEATS):
                        t0 = time.perf_counter()
                        for _ in range(2):
                            solve_item(lane, d, 0, chunk, precision, tune)
                        lane.stream.synchronize()
                        per_e = min(per_e, (time.perf_counter() - t0) / (2 * chunk))
                if verbose:
                    verbose(f"autotune b={b} chunk={chunk} tile={tile}: "
                            f"{per_e * 1e3:.3f} ms/energy")
                score = per_e * (1.0 if tile == TILES[0] else 1.0 + TILE_MARGIN)
                if best is None or score < best[2]:
                    best = (per_e, tune, score)
            del d
        out[b] = dict(best[1], ms_per_energy_probe=round(best[0] * 1e3, 4))
        known[str(b)] = out[b]
    if lane is not None:
        lane.close()
        gc.collect()  # the probe devices' containers hold reference cycles
    if cache:
        table[key] = known
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(table, indent=1))
        except OSError:
            pass
    return out


def uneven_mix(seed: int, n_devices: int = 24, energy_scale: float = 1.0,
               block_sizes=(64, 128, 256, 416), block_counts=(16, 128),
               energies=(64, 1024)):
    """Seeded device mix: block size from ``block_sizes``, block count uniform
    in ``block_counts``, energies uniform in ``energies`` scaled by
    ``energy_scale`` (at least 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_devices):
        b = int(rng.choice(block_sizes))
        n = int(rng.integers(block_counts[0], block_counts[1] + 1))
        e = int(rng.integers(energies[0], energies[1] + 1))
        out.append((b, n, max(1, int(round(e * energy_scale)))))
    return out
