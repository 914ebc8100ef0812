"""Host-side scheduling logic of the batching engine and the serving loop
(CPU only: no device is touched).

- `engine.plan_items`: energy chunks of uneven devices over lanes (the
  longest-processing-time rule with a chunk cap from the lane's fair share).
- `serving.next_admission`: first-fit admission in arrival order under an HBM
  budget.
"""

import numpy as np
import pytest

from quatrex_amd.engine import plan_items
from quatrex_amd.serving import next_admission


def _covers(devices, assign):
    seen = [np.zeros(n, dtype=int) for n, _, _ in devices]
    for items in assign:
        for it in items:
            assert 0 <= it.e0 < it.e1 <= devices[it.device][0]
            seen[it.device][it.e0:it.e1] += 1
    return all((s == 1).all() for s in seen)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n_lanes", [1, 3, 4])
def test_plan_covers_every_energy_once_and_balances(seed, n_lanes):
    rng = np.random.default_rng(seed)
    devices = [(int(rng.integers(1, 200)), float(rng.uniform(1e9, 5e11)),
                int(rng.integers(1, 64))) for _ in range(int(rng.integers(1, 30)))]
    assign, load = plan_items(devices, n_lanes)
    assert len(assign) == n_lanes and len(load) == n_lanes
    assert _covers(devices, assign)
    # chunk caps and flop accounting
    for lane_items, lane_load in zip(assign, load):
        for it in lane_items:
            assert it.e1 - it.e0 <= devices[it.device][2]
            assert it.flops == pytest.approx((it.e1 - it.e0) * devices[it.device][1])
        assert lane_load == pytest.approx(sum(it.flops for it in lane_items))
    # LPT bound: the most loaded lane exceeds the mean by at most one item
    biggest = max(it.flops for items in assign for it in items)
    assert max(load) <= sum(load) / n_lanes + biggest * (1 + 1e-12)


def test_plan_spreads_a_dominant_device():
    # one device carries ~97 % of the flops: it must not sit on one lane
    devices = [(512, 3e11, 128)] + [(8, 1e9, 8)] * 10
    assign, load = plan_items(devices, 4)
    lanes_with_big = sum(any(it.device == 0 for it in items) for items in assign)
    assert lanes_with_big == 4
    assert max(load) / min(load) < 1.2


def test_plan_degenerate_inputs():
    assign, load = plan_items([(5, 1.0, 1)], 3)
    assert _covers([(5, 1.0, 1)], assign) and sum(load) == 5.0
    assign, load = plan_items([(3, 0.0, 2)], 2)  # zero-cost device
    assert _covers([(3, 0.0, 2)], assign)


def test_admission_first_fit_in_arrival_order():
    assert next_admission([10, 5, 3], used=0, budget=12) == 0
    assert next_admission([10, 5, 3], used=8, budget=12) == 2   # 10 and 5 do not fit
    assert next_admission([4, 4], used=4, budget=12) == 0       # FIFO among fitting
    assert next_admission([13], used=0, budget=12) is None
    assert next_admission([], used=0, budget=12) is None
    assert next_admission([12], used=0, budget=12) == 0         # exact fit admitted
