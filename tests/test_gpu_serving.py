"""Serving loop: requests of mixed device sizes, admitted by lanes and HBM
budget, overlap on the device; each result equals the same request served
alone, and matches a direct RGF solve of that request."""

import numpy as np
import pytest
import torch

from quatrex_amd.greens_function_solver import RGF, transmission
from quatrex_amd.serving import Request, Server, poisson_requests
from quatrex_amd.workloads import make_workload

pytestmark = pytest.mark.gpu

SHAPES = {"a": (16, 5, 21), "b": (32, 3, 22), "c": (24, 4, 23)}  # b, N, seed


@pytest.fixture(scope="module")
def workloads():
    return {k: make_workload(seed=s, n_energies=2, n_blocks=n, block_size=b)
            for k, (b, n, s) in SHAPES.items()}


def _server(device, workloads, lanes, budget=8 << 30):
    srv = Server(device, n_lanes=lanes, hbm_budget_bytes=budget, chunk=4)
    for k, w in workloads.items():
        srv.add_device(k, w)
    return srv


def _clone(r):
    return Request(r.key, r.energies.copy(), r.arrival, r.sigma_lesser, r.sigma_greater)


def _direct(w, req, device):
    """Direct RGF solve of the request (its energies and Sigma) on the HIP path."""
    from dataclasses import replace

    w2 = replace(w, energies=req.energies, sl_diag=req.sigma_lesser or w.sl_diag,
                 sg_diag=req.sigma_greater or w.sg_diag)
    a, sl, sg = w2.assemble(device)
    ne, n_orb = len(req.energies), int(np.sum(w.block_sizes))
    dens = torch.zeros((ne, n_orb), dtype=torch.float64, device=device)
    corner = torch.zeros((ne, w.block_sizes[0], w.block_sizes[-1]), dtype=torch.complex128,
                         device=device)
    RGF(ne).selected_solve(a, sl, sg, density=dens, corner=corner)
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    t = transmission(corner, np.broadcast_to(gl, (ne,) + gl.shape),
                     np.broadcast_to(gr, (ne,) + gr.shape))
    return dens, t


def test_alone_equals_mix_and_direct(device, workloads):
    reqs = poisson_requests(list(workloads), rate_per_s=200.0, n=9, energies_range=(3, 11),
                            seed=1, devices=workloads)
    mix = _server(device, workloads, lanes=3)
    stats = mix.run(reqs, realtime=False)
    assert stats["requests"] == 9 and stats["latency_ms"]["p99"] >= stats["latency_ms"]["p50"]
    # requests overlapped: some request was admitted before another finished
    spans = sorted((r.t_admit, r.t_done) for r in reqs)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    alone_srv = _server(device, workloads, lanes=1)
    for r in reqs[:4]:
        alone = _clone(r)
        alone_srv.run([alone], realtime=False)
        assert torch.equal(alone.density, r.density)
        assert torch.equal(alone.transmission, r.transmission)
        dens, t = _direct(workloads[r.key], r, device)
        assert torch.allclose(r.density, dens, rtol=1e-12, atol=1e-12 * dens.abs().max().item())
        assert torch.allclose(r.transmission, t, rtol=1e-12)


def test_hbm_budget_serialises(device, workloads):
    srv = _server(device, workloads, lanes=2)
    reqs = [Request("b", np.linspace(-0.5, 0.5, 8) + 1e-3j), Request("b", np.linspace(-0.4, 0.6, 8) + 1e-3j)]
    srv.hbm_budget_bytes = srv.resident + srv._need(reqs[0]) + 1  # room for one at a time
    srv.run(reqs, realtime=False)
    first, second = sorted(reqs, key=lambda r: r.t_admit)
    assert second.t_admit >= first.t_done
    with pytest.raises(MemoryError):
        srv.hbm_budget_bytes = srv.resident + 10
        srv.run([Request("b", np.zeros(4) + 1e-3j)])
