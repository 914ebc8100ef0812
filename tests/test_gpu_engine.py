"""Uneven batched launch (devices of mixed shapes on concurrent lanes, graph
capture) against direct per-device RGF solves on the HIP path, which the RGF
parity tests pin to the oracle."""

import numpy as np
import pytest
import torch

from quatrex_amd.engine import Device, UnevenBatch, autotune
from quatrex_amd.greens_function_solver import RGF, transmission
from quatrex_amd.workloads import make_workload

pytestmark = pytest.mark.gpu

SHAPES = [(16, 5, 6, 11), (32, 3, 9, 12), (24, 4, 5, 13)]  # b, N, energies, seed


def _direct(w, device):
    """Density and transmission of one device from a plain RGF call."""
    a, sl, sg = w.assemble(device)
    ne, n_orb = len(w.energies), int(np.sum(w.block_sizes))
    dens = torch.zeros((ne, n_orb), dtype=torch.float64, device=device)
    corner = torch.zeros((ne, w.block_sizes[0], w.block_sizes[-1]), dtype=torch.complex128,
                         device=device)
    RGF(ne).selected_solve(a, sl, sg, density=dens, corner=corner)
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    t = transmission(corner, np.broadcast_to(gl, (ne,) + gl.shape),
                     np.broadcast_to(gr, (ne,) + gr.shape))
    return dens.cpu().numpy(), t.cpu().numpy()


@pytest.fixture(scope="module")
def batch(device):
    ws = [make_workload(seed=s, n_energies=e, n_blocks=n, block_size=b) for b, n, e, s in SHAPES]
    devs = [Device(w, device, f"d{k}") for k, w in enumerate(ws)]
    ub = UnevenBatch(devs, n_lanes=2, default_chunk=4)
    refs = [_direct(w, device) for w in ws]
    return ws, devs, ub, refs


def _check(devs, refs, tol):
    for d, (dens, t) in zip(devs, refs):
        assert np.allclose(d.density.cpu().numpy(), dens, rtol=tol, atol=tol * np.abs(dens).max())
        assert np.allclose(d.transmission.cpu().numpy(), t, rtol=tol)


def test_eager_matches_direct(batch):
    ws, devs, ub, refs = batch
    # the plan covers every energy of every device exactly once
    for k, d in enumerate(devs):
        spans = sorted((it.e0, it.e1) for lane in ub.lanes for it in lane.items
                       if it.device == k)
        assert spans[0][0] == 0 and spans[-1][1] == d.n_energies
        assert all(a[1] == b[0] for a, b in zip(spans, spans[1:]))
    ub.run_eager()
    _check(devs, refs, 1e-12)


def test_graph_replay_matches_eager(batch):
    ws, devs, ub, refs = batch
    ub.capture()
    for rep in range(3):
        for d in devs:
            d.density.zero_()
            d.transmission.zero_()
        ub.replay()
        ub.check()
        _check(devs, refs, 1e-12)
        # eager work between replays
        torch.ones(1000, device=devs[0].device).sum().item()


def test_autotune_table(device, tmp_path, monkeypatch):
    monkeypatch.setenv("QX_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    import quatrex_amd.engine as eng

    monkeypatch.setattr(eng, "CHUNKS", (2, 4))
    table = autotune([16, 32], device, n_blocks=3)
    assert set(table) == {16, 32}
    for v in table.values():
        assert v["gemm_tile"] in (0, 64, 128) and v["chunk"] in (2, 4)
    again = autotune([16, 32], device, n_blocks=3)
    assert again == table


def test_capture_refused_above_the_limit(device):
    # 30 items of a 32-block device on one lane: 960 block steps
    from quatrex_amd.engine import CAPTURE_BLOCK_STEPS_MAX

    w = make_workload(seed=3, n_energies=30, n_blocks=32, block_size=8)
    ub = UnevenBatch([Device(w, device, "long")], n_lanes=1, default_chunk=1)
    assert ub.lane_block_steps() == [960] and 960 > CAPTURE_BLOCK_STEPS_MAX
    assert not ub.can_capture()
    with pytest.raises(RuntimeError, match="capture limit"):
        ub.capture()
    assert ub.graph is None
    ub.close()
    assert all(lane.h is None for lane in ub.lanes)
