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


def _bound_case(monkeypatch, data, stack_shape, s0, s1, output=False):
    import torch  # noqa: F401

    from quatrex_amd import _lib
    from quatrex_amd.datastructures import DSDBCOO, btd_pattern
    from quatrex_amd.greens_function_solver import _Bound

    monkeypatch.setattr(_lib, "require_device_tensor", lambda t, what: None)
    bs = [2, 3]
    rows, cols, _ = btd_pattern(bs, offdiag=True)
    m = DSDBCOO(data, rows, cols, bs, stack_shape)
    return m, _Bound(m, s0, s1, output=output)


@pytest.mark.parametrize("s0,s1", [(0, 4), (1, 3)])
def test_bound_strides_for_contiguous_and_broadcast_stacks(monkeypatch, s0, s1):
    """Block layouts handed to the C ABI address only memory the tensor owns:
    contiguous stacks by offset and stride nnz; a stride-0 broadcast over the
    only stack axis with stack stride 0; a broadcast over the first of two
    stack axes (period K, not expressible by one stride) through a
    contiguous copy of the batch; outputs only when contiguous."""
    import torch

    nnz = 2 * 2 + 2 * 3 * 2 + 3 * 3
    E, K = 4, 3
    full = torch.arange(E * K * nnz, dtype=torch.float64).to(torch.complex128).reshape(E, K, nnz)
    m, b = _bound_case(monkeypatch, full, (E, K), s0, s1)
    assert b.layout.stack_stride == nnz and b.layout.n_stack == (s1 - s0) * K
    assert b.base == full.data_ptr() + s0 * K * nnz * 16
    one = torch.arange(nnz, dtype=torch.float64).to(torch.complex128).reshape(1, nnz)
    m, b = _bound_case(monkeypatch, one.expand(E, nnz), (E,), s0, s1)
    assert b.layout.stack_stride == 0 and b.base == one.data_ptr()
    exp = torch.arange(K * nnz, dtype=torch.float64).to(torch.complex128).reshape(1, K, nnz).expand(E, K, nnz)
    m, b = _bound_case(monkeypatch, exp, (E, K), s0, s1)
    assert m.data.stride(0) == 0  # the container keeps the broadcast
    assert b.layout.stack_stride == nnz and b.layout.n_stack == (s1 - s0) * K
    assert b._copy is not None and b.base == b._copy.data_ptr()
    assert b._copy.numel() == (s1 - s0) * K * nnz
    assert torch.equal(b._copy, exp[s0:s1].reshape(-1, nnz))
    with pytest.raises(ValueError, match="contiguous"):
        _bound_case(monkeypatch, exp, (E, K), s0, s1, output=True)


def test_bound_plan_cache_matches_fresh_layouts(monkeypatch):
    """The block plan and layout cached in the pattern cache (shared by views
    and zeros_like copies) give the offsets a fresh container computes, per
    flag combination and batch size; the base pointer follows each call."""
    import ctypes

    import numpy as np
    import torch

    from quatrex_amd import _lib
    from quatrex_amd.datastructures import DSDBCOO, btd_pattern
    from quatrex_amd.greens_function_solver import _Bound

    monkeypatch.setattr(_lib, "require_device_tensor", lambda t, what: None)
    bs = [2, 3, 2]
    rows, cols, _ = btd_pattern(bs, offdiag=True)
    nnz = len(rows)

    def offsets(b):
        n = len(bs)
        lay = b.layout
        get = lambda p, k: np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_int64)),
                                                 shape=(k,)).copy()
        return (lay.n_stack, lay.stack_stride, tuple(get(lay.diag, n)),
                tuple(get(lay.upper, n - 1)), tuple(get(lay.lower, n - 1)))

    data = torch.arange(5 * nnz, dtype=torch.float64).to(torch.complex128).reshape(5, nnz)
    m = DSDBCOO(data, rows, cols, bs, (5,))
    z = DSDBCOO.zeros_like(m)  # shares the pattern cache
    cases = [dict(output=False), dict(output=True), dict(output=True, write_lower=False),
             dict(output=True, write_offdiag=False)]
    for kw in cases:
        for s0, s1 in ((0, 5), (1, 3), (0, 5)):
            fresh = _Bound(DSDBCOO(data.clone(), rows, cols, bs, (5,)), s0, s1, **kw)
            first = _Bound(m, s0, s1, **kw)
            again = _Bound(z, s0, s1, **kw)
            assert offsets(first) == offsets(fresh) == offsets(again)
            assert first.base == data.data_ptr() + s0 * nnz * 16
            assert again.base == z.data.data_ptr() + s0 * nnz * 16
    lw = _Bound(m, 0, 5, output=True, write_lower=False)
    assert offsets(lw)[4] == (-1, -1)  # lower blocks not written
    off = _Bound(m, 0, 5, output=True, write_offdiag=False)
    assert offsets(off)[3] == (-1, -1) and offsets(off)[4] == (-1, -1)


def test_host_eig_pool_matches_sequential_eig():
    """nevp.host_eig: one LAPACK eig per matrix on a thread pool (one BLAS
    thread each) returns, in input order, the eigenpairs numpy.linalg.eig
    gives matrix by matrix (same routine; rounding may differ with the BLAS
    thread count), for a list of different sizes and for a stacked array."""
    import numpy as np

    from quatrex_amd.nevp import host_eig

    rng = np.random.default_rng(4)
    mats = [rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
            for m in (1, 5, 17, 40, 3, 64)]
    for got, a in zip(host_eig(mats), mats):
        w, v = got
        assert w.shape == (a.shape[0],) and v.shape == a.shape
        # eigenpairs: A v = w v, and the same spectrum as the direct call
        assert np.allclose(a @ v, v * w[None, :], atol=1e-10 * max(1.0, np.abs(a).max()))
        w0, _ = np.linalg.eig(a)
        assert np.allclose(np.sort_complex(w), np.sort_complex(w0), atol=1e-10)
    stack = rng.standard_normal((9, 12, 12)) + 1j * rng.standard_normal((9, 12, 12))
    pairs = host_eig(stack)
    assert len(pairs) == 9
    for (w, v), a in zip(pairs, stack):
        assert np.allclose(a @ v, v * w[None, :], atol=1e-10 * np.abs(a).max())
    assert host_eig([]) == []
