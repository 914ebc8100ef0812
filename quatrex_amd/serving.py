"""Serving loop: continuous batching of SCF-iteration solve requests.

A request names a registered device (its Hamiltonian stays resident in
HBM across SCF iterations), carries the energies of this iteration and
optionally updated Sigma^< / Sigma^> blocks, and asks for the electron
density Im diag G^< and the transmission at every energy.

Scheduling:

- requests wait in arrival order; one is admitted as soon as a lane is idle
  and the HBM budget has room for its lane buffers (chunk x per-energy
  bytes) and its own buffers (energies, Sigma, results), so small requests
  are not held back by a large one at the head once the large one fits
  nowhere (first-fit over the queue, FIFO among those that fit);
- every lane runs its request's energy chunks on its own stream and handle,
  so requests admitted while others run overlap with them on the device;
- a request's arithmetic does not depend on what runs beside it (every
  kernel works on its own matrices, deterministic reductions), so its
  result is the same alone and in the mix; the tests check this.

Latency = completion - arrival (queueing included); the loop reports
p50/p99 latency and energies/s over the whole run.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from quatrex_amd.engine import Device, Job, Lane, solve_item


@dataclass
class Request:
    key: str
    energies: np.ndarray
    arrival: float = 0.0            # seconds after the start of the run
    sigma_lesser: list | None = None
    sigma_greater: list | None = None
    # filled by the server
    job: Job | None = None
    t_admit: float | None = None
    t_done: float | None = None
    lane: int | None = None

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival

    @property
    def density(self) -> torch.Tensor:
        return self.job.density

    @property
    def transmission(self) -> torch.Tensor:
        return self.job.transmission


def next_admission(needs, used: int, budget: int):
    """Index of the first queued request (arrival order) whose bytes
    ``needs[i]`` fit beside ``used`` in ``budget``, or None."""
    for i, need in enumerate(needs):
        if used + need <= budget:
            return i
    return None


def lane_increment(retained: tuple, need: tuple) -> int:
    """Bytes a lane retaining (pool, workspace) = ``retained`` adds when it
    serves a request needing ``need``: each buffer only grows, to the larger
    of the two, and is kept after the request."""
    return max(0, need[0] - retained[0]) + max(0, need[1] - retained[1])


@dataclass
class Server:
    device: torch.device
    n_lanes: int = 4
    hbm_budget_bytes: int = 160 << 30
    precision: str = "c128"
    chunk: int = 16
    tuning: dict = field(default_factory=dict)
    lane_budget_bytes: int = 24 << 30

    def __post_init__(self):
        self.device = torch.device(self.device)
        self.devices: dict[str, Device] = {}
        self.lanes = [Lane(self.device, i) for i in range(self.n_lanes)]
        self.resident = 0
        # (pool, workspace) bytes each lane keeps between requests
        self.retained = [(0, 0)] * self.n_lanes

    def add_device(self, key: str, workload) -> Device:
        """Register a device; its Hamiltonian and Sigma blocks stay resident."""
        d = Device(workload, self.device, key)
        self.devices[key] = d
        self.resident += d.resident_bytes()
        return d

    def _chunk(self, d: Device) -> int:
        c = int(self.tuning.get(d.b, {}).get("chunk", self.chunk))
        return max(1, min(c, self.lane_budget_bytes // d.lane_bytes_per_energy(self.precision)))

    def _lane_need(self, req: Request) -> tuple:
        """(pool, workspace) bytes a lane grows to for the request's chunks."""
        d = self.devices[req.key]
        c = min(self._chunk(d), len(req.energies))
        return c * d.pool_bytes_per_energy(), d.workspace_bytes(c, self.precision)

    def _extra(self, req: Request) -> int:
        """Bytes a request holds while it is served: its energies and
        results, and its own Sigma blocks."""
        d = self.devices[req.key]
        return 24 * len(req.energies) * (d.n_orb + 1) + \
            (64 * d.nnz_d if req.sigma_lesser is not None else 0)

    def _need(self, req: Request) -> int:
        """Bytes a request adds on a lane that retains nothing."""
        return sum(self._lane_need(req)) + self._extra(req)

    def _work(self, lane: Lane, req: Request) -> None:
        d = self.devices[req.key]
        with torch.cuda.stream(lane.stream):
            job = Job(d, req.energies, req.sigma_lesser, req.sigma_greater)
            req.job = job
            c = self._chunk(d)
            from quatrex_amd import _lib

            with _lib.use_handle(lane.h):
                for e0 in range(0, job.n_energies, c):
                    e1 = min(e0 + c, job.n_energies)
                    solve_item(lane, d, e0, e1, self.precision, self.tuning.get(d.b), job)
                lane.stream.synchronize()
                _lib.sync_check(self.device, "Server request")

    def run(self, requests: list[Request], realtime: bool = True) -> dict:
        """Serve ``requests`` (arrival times relative to the start); blocks.

        ``realtime`` releases each request at its arrival time; otherwise all
        are queued at once (arrival 0 for the latency).
        """
        for r in requests:
            if r.key not in self.devices:
                raise KeyError(f"unknown device {r.key!r}")
            if self._need(r) + self.resident > self.hbm_budget_bytes:
                raise MemoryError(f"request for {r.key!r} exceeds the HBM budget alone")
        pending = sorted(requests, key=lambda r: r.arrival)
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        cond = threading.Condition()
        queue: list[Request] = []
        idle = list(range(self.n_lanes))
        # device bytes charged: resident devices, what every lane retains,
        # and the requests in service
        state = {"extra": 0, "done": 0, "error": None}
        charged = lambda: self.resident + sum(map(sum, self.retained)) + state["extra"]  # noqa: E731

        def best_lane(req):
            """Idle lane that grows least for req, and the bytes it adds."""
            need = self._lane_need(req)
            return min(((lane_increment(self.retained[k], need), k) for k in idle),
                       default=(None, None))
        t_start = time.perf_counter()
        now = lambda: time.perf_counter() - t_start

        def lane_main(k: int, req: Request):
            try:
                self._work(self.lanes[k], req)
            except Exception as e:  # reported by run()
                state["error"] = e
            with cond:
                req.t_done = now()
                state["extra"] -= self._extra(req)
                state["done"] += 1
                idle.append(k)
                cond.notify_all()

        threads = []
        i = 0
        with cond:
            while state["done"] < len(requests):
                if state["error"] is not None:
                    break
                t = now()
                while i < len(pending) and pending[i].arrival <= t:
                    queue.append(pending[i])
                    i += 1
                # admit: first request in arrival order that fits, on the
                # idle lane that grows least, while a lane is idle; if none
                # fits, idle lanes give back what they retain
                while idle and queue:
                    adds = [best_lane(r)[0] + self._extra(r) for r in queue]
                    q = next_admission(adds, charged(), self.hbm_budget_bytes)
                    if q is None:
                        held = [k for k in idle if sum(self.retained[k]) > 0]
                        if not held:
                            break
                        k = max(held, key=lambda j: sum(self.retained[j]))
                        self.lanes[k].release()
                        self.retained[k] = (0, 0)
                        continue
                    req = queue.pop(q)
                    _, k = best_lane(req)
                    idle.remove(k)
                    need = self._lane_need(req)
                    self.retained[k] = (max(self.retained[k][0], need[0]),
                                        max(self.retained[k][1], need[1]))
                    state["extra"] += self._extra(req)
                    req.t_admit, req.lane = now(), k
                    th = threading.Thread(target=lane_main, args=(k, req))
                    threads.append(th)
                    th.start()
                wait = None
                if i < len(pending):
                    wait = max(0.0, pending[i].arrival - now())
                cond.wait(timeout=wait if wait is not None else 0.5)
        for th in threads:
            th.join()
        if state["error"] is not None:
            raise state["error"]
        makespan = max(r.t_done for r in requests)
        lat = np.array([r.latency for r in requests]) * 1e3
        n_e = sum(len(r.energies) for r in requests)
        # request costs span orders of magnitude (b = 64 to 416, 16 to 128
        # blocks), so the required flops per second measure throughput too
        flops = sum(len(r.energies) * self.devices[r.key].flops_per_energy() for r in requests)
        return {"requests": len(requests), "energies": n_e,
                "energies_per_s": n_e / makespan, "makespan_s": makespan,
                "required_tflops": flops / makespan / 1e12,
                "latency_ms": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99)),
                               "max": float(lat.max())}}


def poisson_requests(keys, rate_per_s: float, n: int, energies_range, seed: int = 0,
                     sigma_update: bool = True, devices: dict | None = None) -> list:
    """n requests with exponential inter-arrival times (rate ``rate_per_s``),
    device chosen uniformly from ``keys``, energy count uniform in
    ``energies_range``; with ``devices`` (key -> Workload) each request also
    carries updated Sigma^< / Sigma^> blocks (scaled copies of the device's,
    as an SCF iteration would produce)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate_per_s)
        key = keys[int(rng.integers(len(keys)))]
        ne = int(rng.integers(energies_range[0], energies_range[1] + 1))
        e = rng.uniform(-1.0, 1.0, ne) + 1e-3j
        sl = sg = None
        if sigma_update and devices is not None:
            w = devices[key]
            scale = 1.0 + 0.1 * rng.standard_normal()
            sl = [scale * b for b in w.sl_diag]
            sg = [scale * b for b in w.sg_diag]
        out.append(Request(key, e, t, sl, sg))
    return out
