"""HIP kernels through the C ABI: batched complex GEMM and batched inverse.

Checked against NumPy float64/complex128 on the CPU (the plain fp64
reference of the same op).  Tolerances: complex128 arithmetic 1e-12
relative Frobenius per matrix; complex64 arithmetic 2e-5.
"""

import ctypes

import numpy as np
import pytest
import torch

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _lib():
    from quatrex_amd import _lib

    return _lib


def _cplx(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _zgemm(dev, A, B, C, opa, opb, alpha, beta, prec, lda=None, ldb=None, ldc=None):
    L = _lib()
    lib = L.load()
    h, st = L.handle(dev), L.stream_ptr(dev)
    batch = C.shape[0]
    m, n = C.shape[1], C.shape[2]
    k = A.shape[2] if opa == L.OP_N else A.shape[1]
    al = (ctypes.c_double * 2)(alpha.real, alpha.imag)
    be = (ctypes.c_double * 2)(beta.real, beta.imag)
    rc = lib.qx_zgemm_batched(h, st, batch, m, n, k, opa, opb, al, A.data_ptr(), A.stride(1),
                              A.stride(0), B.data_ptr(), B.stride(1), B.stride(0), be,
                              C.data_ptr(), C.stride(1), C.stride(0), prec)
    L.check(rc, h, "zgemm")
    L.sync_check(dev, "zgemm")


SHAPES = [(1, 1, 1), (5, 7, 3), (16, 16, 16), (64, 64, 64), (100, 37, 129), (130, 66, 70),
          (256, 256, 256)]


@pytest.mark.parametrize("mnk", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("ops", [(0, 0), (0, 2), (2, 0), (2, 2)], ids=["NN", "NC", "CN", "CC"])
@pytest.mark.parametrize("prec", [0, 1, 3], ids=["c128", "c64", "c128_3m"])
def test_zgemm(device, mnk, ops, prec):
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    batch = 3
    opa, opb = ops
    a_shape = (batch, m, k) if opa == 0 else (batch, k, m)
    b_shape = (batch, k, n) if opb == 0 else (batch, n, k)
    a, b, c = _cplx(rng, a_shape), _cplx(rng, b_shape), _cplx(rng, (batch, m, n))
    alpha, beta = 0.7 - 0.3j, (0.0 if (m + n) % 2 else -1.2 + 0.5j)
    # padded leading dimensions: operate on a column slice of a wider buffer
    A = torch.zeros(a_shape[:2] + (a_shape[2] + 3,), dtype=torch.complex128, device=device)
    A[..., :a_shape[2]] = torch.as_tensor(a, device=device)
    B = torch.as_tensor(b, device=device)
    C = torch.as_tensor(c, device=device)
    if beta == 0.0:
        C[:] = float("nan")  # beta == 0 must not read C
    _zgemm(device, A[..., :a_shape[2]], B, C, opa, opb, alpha, beta, prec)
    opA = a if opa == 0 else np.conj(np.swapaxes(a, 1, 2))
    opB = b if opb == 0 else np.conj(np.swapaxes(b, 1, 2))
    ref = alpha * opA @ opB + (beta * c if beta != 0.0 else 0)
    tol = {0: 1e-12, 1: 2e-5, 3: 1e-12}[prec]
    got = C.cpu().numpy()
    for e in range(batch):
        assert rel_err(got[e], ref[e]) < tol
    # the padding columns of A are untouched
    assert torch.all(A[..., a_shape[2]:] == 0)


BIG = [(256, 256, 256, 64), (200, 136, 72, 130), (384, 130, 33, 100)]


@pytest.mark.parametrize("mnkb", BIG, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("ops", [(0, 0), (0, 2), (2, 0), (2, 2)], ids=["NN", "NC", "CN", "CC"])
@pytest.mark.parametrize("prec", [0, 1, 3], ids=["c128", "c64", "c128_3m"])
def test_zgemm_large_stack(device, mnkb, ops, prec):
    """Stacks with >= 256 tiles of 128 x 128 (the large-tile kernel shape)."""
    m, n, k, batch = mnkb
    rng = np.random.default_rng(m + n + k + batch)
    opa, opb = ops
    a_shape = (batch, m, k) if opa == 0 else (batch, k, m)
    b_shape = (batch, k, n) if opb == 0 else (batch, n, k)
    a, b, c = _cplx(rng, a_shape), _cplx(rng, b_shape), _cplx(rng, (batch, m, n))
    alpha, beta = 1.1 + 0.2j, -0.5 + 0.25j
    A, B = torch.as_tensor(a, device=device), torch.as_tensor(b, device=device)
    C = torch.as_tensor(c, device=device)
    _zgemm(device, A, B, C, opa, opb, alpha, beta, prec)
    got = C.cpu().numpy()
    tol = {0: 1e-12, 1: 2e-5, 3: 1e-12}[prec]
    for e in (0, batch // 2, batch - 1):
        opA = a[e] if opa == 0 else np.conj(a[e].T)
        opB = b[e] if opb == 0 else np.conj(b[e].T)
        assert rel_err(got[e], alpha * opA @ opB + beta * c[e]) < tol


REG_CASES = [(256, 256, 256, 8), (128, 128, 64, 4), (200, 136, 72, 3), (5, 7, 3, 2),
             (416, 416, 416, 2)]


@pytest.mark.parametrize("mnkb", REG_CASES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("ops", [(0, 0), (0, 2), (2, 0), (2, 2)], ids=["NN", "NC", "CN", "CC"])
@pytest.mark.parametrize("prec", [0, 3, 1], ids=["c128", "c128_3m", "c64"])
def test_zgemm_register_kernels_match_lds(device, monkeypatch, mnkb, ops, prec):
    """Every register-direct GEMM shape gives the bits of the LDS-staged
    kernel: both sum k in the same order with the same 4M / 3M split (full
    tiles, ragged tiles and k not a multiple of 4).  complex64 arithmetic
    (3M, register-direct on request: QX_ZGEMM_F32=reg): both round the
    operands to float and form the 3M sums in float."""
    monkeypatch.setenv("QX_ZGEMM_F32", "reg")
    L = _lib()
    m, n, k, batch = mnkb
    rng = np.random.default_rng(m + 2 * n + 3 * k)
    opa, opb = ops
    a_shape = (batch, m, k) if opa == 0 else (batch, k, m)
    b_shape = (batch, k, n) if opb == 0 else (batch, n, k)
    A = torch.as_tensor(_cplx(rng, a_shape), device=device)
    B = torch.as_tensor(_cplx(rng, b_shape), device=device)
    C0 = torch.as_tensor(_cplx(rng, (batch, m, n)), device=device)
    alpha, beta = 0.9 + 0.4j, 0.3 - 0.7j
    h = L.handle(device)
    outs = {}
    try:
        for code in (L.GEMM_LDS_64, L.GEMM_REG_64x64, L.GEMM_REG_64x128, L.GEMM_REG_128x64):
            L.set_tuning(h, 4, code)
            C = C0.clone()
            _zgemm(device, A, B, C, opa, opb, alpha, beta, prec)
            outs[code] = C
    finally:
        L.set_tuning(h, 4, L.GEMM_AUTO)
    for code in (L.GEMM_REG_64x64, L.GEMM_REG_64x128, L.GEMM_REG_128x64):
        assert torch.equal(outs[code], outs[L.GEMM_LDS_64]), code


def _inv(dev, A, X, prec, S=None):
    L = _lib()
    lib = L.load()
    h, st = L.handle(dev), L.stream_ptr(dev)
    n = A.shape[-1]
    sp = S.data_ptr() if S is not None else None
    rc = lib.qx_inv_batched(h, st, A.shape[0], n, A.data_ptr(), A.stride(1), A.stride(0), sp,
                            n, n * n, X.data_ptr(), X.stride(1), X.stride(0), prec)
    L.check(rc, h, "inv")
    L.sync_check(dev, "inv")


@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 31, 32, 33, 64, 100, 256, 416, 512])
@pytest.mark.parametrize("prec", [0, 1], ids=["c128", "c64"])
def test_inverse(device, n, prec):
    rng = np.random.default_rng(n)
    batch = 4
    a = _cplx(rng, (batch, n, n))
    if n > 1:
        a[:, 0, 0] = 0.0  # forces a row interchange at the first pivot
    s = 0.1 * _cplx(rng, (batch, n, n))
    A = torch.as_tensor(a, device=device)
    S = torch.as_tensor(s, device=device)
    X = torch.empty_like(A)
    _inv(device, A, X, prec, S)
    ref = np.linalg.inv(a - s)
    got = X.cpu().numpy()
    cond = max(np.linalg.cond(a[e] - s[e]) for e in range(batch))
    tol = (1e-15 if prec == 0 else 1e-7) * cond * max(n, 4)
    for e in range(batch):
        assert rel_err(got[e], ref[e]) < tol, (e, rel_err(got[e], ref[e]), tol)


@pytest.mark.parametrize("n", [1, 3, 20, 32])
@pytest.mark.parametrize("batch", [1, 5, 9])
def test_small_inverse_matches_split_kernels(device, monkeypatch, n, batch):
    """n <= 32 (complex128): one launch, one matrix per wave or per
    workgroup; same pivot rule as the panel kernels, so the same permutation
    and results to rounding (batch not a multiple of the four matrices per
    workgroup of the per-wave kernel)."""
    rng = np.random.default_rng(10 * n + batch)
    a = _cplx(rng, (batch, n, n))
    s = 0.1 * _cplx(rng, (batch, n, n))
    A, S = torch.as_tensor(a, device=device), torch.as_tensor(s, device=device)
    outs = []
    for small in ("1", "2", "0"):  # per wave, per workgroup, split panel kernels
        monkeypatch.setenv("QX_INV_SMALL", small)
        X = torch.empty_like(A)
        _inv(device, A, X, 0, S)
        outs.append(X.cpu().numpy())
    cond = max(np.linalg.cond(a[e] - s[e]) for e in range(batch))
    for e in range(batch):
        for o in outs[:2]:
            assert rel_err(o[e], outs[2][e]) < 1e-15 * cond * max(n, 4)
            assert rel_err(o[e], np.linalg.inv(a[e] - s[e])) < 1e-15 * cond * max(n, 4)


@pytest.mark.parametrize("n", [12, 32, 100, 300])
@pytest.mark.parametrize("scale", [2.0 ** -1000, 1e-150, 1e150, 2.0 ** 1000])
def test_inverse_extreme_scales(device, n, scale):
    """Pivot reciprocals are formed from parts scaled by a power of two:
    matrices whose entries are near the ends of the double range invert as
    accurately as unit-scale ones (inv(s A) = inv(A) / s), on the per-workgroup
    small path (n <= 32) and both panel kernels (n <= 256, n > 256)."""
    rng = np.random.default_rng(n)
    a = _cplx(rng, (2, n, n))
    a[:, 0, 0] = 0.0
    A = torch.as_tensor(a * scale, device=device)
    X = torch.empty_like(A)
    _inv(device, A, X, 0)
    got = X.cpu().numpy() * scale
    for e in range(2):
        ref = np.linalg.inv(a[e])
        assert np.isfinite(got[e]).all()
        assert rel_err(got[e], ref) < 1e-15 * np.linalg.cond(a[e]) * n, (e, rel_err(got[e], ref))


@pytest.mark.parametrize("n", [33, 64, 100, 256, 300, 416, 512])
@pytest.mark.parametrize("prec", [0, 1], ids=["c128", "c64"])
def test_panel_kernels_agree(device, monkeypatch, n, prec):
    """The column-group panel kernel (default; 4 or 8 columns per wave)
    against the four-threads-per-row kernel (n <= 256) and the
    one-thread-per-row kernel (all n): the same Gauss-Jordan elimination with
    partial pivoting, so equal to rounding (near-ties between pivot
    candidates may be broken differently)."""
    rng = np.random.default_rng(3 * n + prec)
    batch = 3
    a = _cplx(rng, (batch, n, n))
    a[:, 0, 0] = 0.0
    A = torch.as_tensor(a, device=device)
    variants = [("0", "4"), ("0", "8"), ("1", "8")] + ([("4", "8")] if n <= 256 else [])
    outs = {}
    for kind, cw in variants:
        monkeypatch.setenv("QX_INV_PANEL", kind)
        monkeypatch.setenv("QX_INV_CW", cw)
        X = torch.empty_like(A)
        _inv(device, A, X, prec)
        outs[(kind, cw)] = X.cpu().numpy()
    cond = max(np.linalg.cond(a[e]) for e in range(batch))
    tol = (1e-15 if prec == 0 else 1e-7) * cond * n
    for e in range(batch):
        ref = np.linalg.inv(a[e])
        for v, o in outs.items():
            assert rel_err(o[e], ref) < tol, (v, e, rel_err(o[e], ref))
            assert rel_err(o[e], outs[("1", "8")][e]) < tol, (v, e)


@pytest.mark.parametrize("n", [33, 64, 65, 100, 128, 200, 256])
@pytest.mark.parametrize("prec", [0, 1], ids=["c128", "c64"])
@pytest.mark.parametrize("in_place", [False, True])
def test_panel_width_64(device, monkeypatch, n, prec, in_place):
    """64-column panels (eight waves of eight columns) with rank-64 updates in
    two k-chunks, against the 32-column panels and NumPy: the same
    elimination blocked differently, so equal to rounding; in place and with
    a separate output (the panel count's parity picks the start buffer)."""
    rng = np.random.default_rng(7 * n + prec)
    batch = 3
    a = _cplx(rng, (batch, n, n))
    a[:, 0, 0] = 0.0  # a row interchange at the first pivot
    a[:, n - 1, n - 1] = 0.0
    outs = {}
    for nb in ("64", "32"):
        monkeypatch.setenv("QX_INV_NB", nb)
        A = torch.as_tensor(a, device=device)
        X = A if in_place else torch.empty_like(A)
        _inv(device, A, X, prec)
        outs[nb] = X.cpu().numpy()
    cond = max(np.linalg.cond(a[e]) for e in range(batch))
    tol = (1e-15 if prec == 0 else 1e-7) * cond * n
    for e in range(batch):
        ref = np.linalg.inv(a[e])
        assert rel_err(outs["64"][e], ref) < tol, (e, rel_err(outs["64"][e], ref), tol)
        assert rel_err(outs["64"][e], outs["32"][e]) < tol


@pytest.mark.parametrize("n", [40, 100, 256])
@pytest.mark.parametrize("prec", [0, 1], ids=["c128", "c64"])
def test_update_chunk_depth_bitwise(device, monkeypatch, n, prec):
    """The trailing update staged in 16-deep k-chunks (four workgroups per
    CU, the default from 256 matrices per call) or 32-deep ones (two): the
    same products summed in the same k order, so the same bits."""
    rng = np.random.default_rng(11 * n + prec)
    a = _cplx(rng, (5, n, n))
    a[:, 0, 0] = 0.0
    outs = {}
    for depth in ("16", "32"):
        monkeypatch.setenv("QX_INV_UPDATE", depth)
        A = torch.as_tensor(a, device=device)
        X = torch.empty_like(A)
        _inv(device, A, X, prec)
        outs[depth] = X.cpu().numpy()
    assert np.array_equal(outs["16"].view(np.uint64), outs["32"].view(np.uint64))
    cond = max(np.linalg.cond(a[e]) for e in range(5))
    tol = (1e-15 if prec == 0 else 1e-7) * cond * n
    for e in range(5):
        assert rel_err(outs["16"][e], np.linalg.inv(a[e])) < tol


@pytest.mark.parametrize("n,batch", [(40, 5), (64, 3), (100, 5), (256, 6), (256, 300),
                                     (300, 4), (512, 2), (130, 260)])
@pytest.mark.parametrize("prec", [0, 1], ids=["c128", "c64"])
def test_lookahead_bitwise(device, monkeypatch, n, batch, prec):
    """Look-ahead (the next panel launched beside the rest of the current
    update) against the panel / update sequence: the same kernels' arithmetic
    on the same data, so the same bits, for one and several column tiles and
    both panel row counts (n > 256).  From 256 matrices on (300, 260) both
    settings run the sequence."""
    rng = np.random.default_rng(7 * n + batch + prec)
    a = _cplx(rng, (batch, n, n))
    a[:, 0, 0] = 0.0  # the first pivot needs a row interchange
    outs = {}
    for la in ("1", "0"):
        monkeypatch.setenv("QX_INV_LOOKAHEAD", la)
        A = torch.as_tensor(a, device=device)
        X = torch.empty_like(A)
        _inv(device, A, X, prec)
        outs[la] = X.cpu().numpy()
    assert np.array_equal(outs["1"].view(np.uint64), outs["0"].view(np.uint64))
    for e in range(2):
        cond = np.linalg.cond(a[e])
        tol = (1e-15 if prec == 0 else 1e-7) * cond * n
        assert rel_err(outs["1"][e], np.linalg.inv(a[e])) < tol


@pytest.mark.parametrize("nb", ["32", "64"])
@pytest.mark.parametrize("n", [64, 300])
def test_singular_panel_raises(device, monkeypatch, nb, n):
    monkeypatch.setenv("QX_INV_NB", nb)
    a = np.zeros((2, n, n), complex)
    a[0] = np.eye(n)
    a[1, :, : n // 2] = 1.0  # rank 1
    A = torch.as_tensor(a, device=device)
    X = torch.empty_like(A)
    with pytest.raises(np.linalg.LinAlgError):
        _inv(device, A, X, 0)


def test_inverse_in_place(device):
    rng = np.random.default_rng(5)
    a = _cplx(rng, (3, 40, 40))
    A = torch.as_tensor(a, device=device)
    _inv(device, A, A, 0)
    got = A.cpu().numpy()
    for e in range(3):
        assert rel_err(got[e], np.linalg.inv(a[e])) < 1e-12
    # the one-wave path (n <= 16) in place
    b = _cplx(rng, (5, 12, 12))
    B = torch.as_tensor(b, device=device)
    _inv(device, B, B, 0)
    for e in range(5):
        assert rel_err(B[e].cpu().numpy(), np.linalg.inv(b[e])) < 1e-12


def test_singular_block_raises(device):
    a = np.zeros((2, 8, 8), complex)
    a[0] = np.eye(8)
    A = torch.as_tensor(a, device=device)
    X = torch.empty_like(A)
    with pytest.raises(np.linalg.LinAlgError):
        _inv(device, A, X, 0)
    # the error flag is cleared after it was reported
    A[1] = torch.eye(8, dtype=A.dtype, device=device)
    _inv(device, A, X, 0)
