"""qx_inv_batched_noalloc (scratch in a fixed device array) against
qx_inv_batched (handle workspace), through the C ABI: same kernels, so the
results must be bitwise equal, on two calls in a row (the second smaller,
so stale contents of the shared array from the first call are present), for
the panel inverse (n > 32, which uses the scratch buffer) and the small one
(n <= 32).  Device memory in use does not change across the calls, the
singular-matrix flags are written per matrix, and a size beyond the fixed
array is rejected before anything runs."""

import ctypes

import numpy as np
import pytest
import torch

from quatrex_amd import _lib

pytestmark = pytest.mark.gpu


def _mats(batch, n, seed, device):
    g = torch.Generator(device=device).manual_seed(seed)
    a = torch.randn((batch, n, n), dtype=torch.complex128, device=device, generator=g)
    return a + n * torch.eye(n, dtype=torch.complex128, device=device)


def _ref(lib, h, a, stream):
    batch, n, _ = a.shape
    x = torch.empty_like(a)
    rc = lib.qx_inv_batched(h, stream, batch, n, a.data_ptr(), n, n * n, None, 0, 0,
                            x.data_ptr(), n, n * n, _lib.PREC_C128)
    assert rc == _lib.QX_OK, lib.qx_last_error(h)
    return x


def _noalloc(lib, h, a, x, info, stream):
    batch, n, _ = a.shape
    return lib.qx_inv_batched_noalloc(h, stream, batch, n, a.data_ptr(), n, n * n, None, 0, 0,
                                      x.data_ptr(), n, n * n,
                                      None if info is None else info.data_ptr(), _lib.PREC_C128)


@pytest.mark.parametrize("shapes", [((48, 256), (8, 96)), ((40, 32), (6, 20))],
                         ids=["panel", "small"])
def test_noalloc_matches_workspace_inverse(device, shapes):
    lib = _lib.load()
    h = _lib.new_handle(device)
    stream = _lib.stream_ptr(device)
    try:
        cases = [(_mats(b, n, 11 + k, device), k) for k, (b, n) in enumerate(shapes)]
        refs = [_ref(lib, h, a, stream) for a, _ in cases]
        outs = [torch.full_like(a, float("nan")) for a, _ in cases]
        infos = [torch.full((a.shape[0],), 7, dtype=torch.int32, device=device)
                 for a, _ in cases]
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(device)[0]
        for (a, k), x, info in zip(cases, outs, infos):
            assert _noalloc(lib, h, a, x, info, stream) == _lib.QX_OK, lib.qx_last_error(h)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info(device)[0] == free0
        for x, r, info in zip(outs, refs, infos):
            assert torch.equal(x, r)
            assert int(info.abs().sum()) == 0
        # status flags without an info array go to the scratch array
        x = torch.empty_like(cases[1][0])
        assert _noalloc(lib, h, cases[1][0], x, None, stream) == _lib.QX_OK
        torch.cuda.synchronize()
        assert torch.equal(x, refs[1])
    finally:
        lib.qx_destroy(h)


def test_noalloc_singular_flags_and_size_limit(device):
    lib = _lib.load()
    h = _lib.new_handle(device)
    stream = _lib.stream_ptr(device)
    try:
        a = _mats(4, 64, 3, device)
        a[2] = 0  # exactly singular
        x = torch.empty_like(a)
        info = torch.full((4,), 5, dtype=torch.int32, device=device)
        assert _noalloc(lib, h, a, x, info, stream) == _lib.QX_OK
        torch.cuda.synchronize()
        flags = info.cpu().numpy()
        assert flags[2] != 0 and np.all(flags[[0, 1, 3]] == 0)
        # 64 matrices of 256 need more than the fixed 64 MiB
        big = torch.empty((64, 256, 256), dtype=torch.complex128, device=device)
        out = torch.empty_like(big)
        rc = _noalloc(lib, h, big, out, None, stream)
        assert rc == _lib.QX_EINVAL
        assert b"scratch" in lib.qx_last_error(h)
    finally:
        lib.qx_destroy(h)
