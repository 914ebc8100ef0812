"""Ozaki-style split GEMM (bf16 matrix cores) against NumPy complex128.

Pinned by the truncation bound: s 7-bit slices leave each operand with a
remainder below 2^(-7 s) of its line's largest entry, and the digit pairs
p + q <= s + 1 keep every product term of that order, so the relative error
is a small multiple of 2^(-7 s) (a NumPy emulation of the same split gives
5.4e-4 .. 5.4e-14 for s = 2 .. 7 at 256^3) and must fall by about 2^-7 per
added slice; a wrong sign, scale or digit would leave an O(1) error.  Rows and columns of very different magnitude check the
per-line power-of-two scaling, zero lines the exponent of an empty line.
"""

import numpy as np
import pytest
import torch

from quatrex_amd.ops import gemm_ozaki
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _cplx(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _ref(a, b, opa, opb):
    A = a if opa == "N" else np.conj(np.swapaxes(a, -1, -2))
    B = b if opb == "N" else np.conj(np.swapaxes(b, -1, -2))
    return A @ B


def _bound(s):
    return 64.0 * 2.0 ** (-7 * s) + 1e-14


@pytest.mark.parametrize("mnk", [(64, 64, 64), (100, 37, 129), (256, 256, 256)],
                         ids=lambda t: "x".join(map(str, t)))
@pytest.mark.parametrize("ops", [("N", "N"), ("N", "C"), ("C", "N"), ("C", "C")],
                         ids=["NN", "NC", "CN", "CC"])
def test_error_falls_with_slices(device, mnk, ops):
    m, n, k = mnk
    rng = np.random.default_rng(m + n + k)
    batch = 3
    opa, opb = ops
    a = _cplx(rng, (batch, m, k) if opa == "N" else (batch, k, m))
    b = _cplx(rng, (batch, k, n) if opb == "N" else (batch, n, k))
    ref = _ref(a, b, opa, opb)
    errs = []
    for s in range(2, 8):
        got = gemm_ozaki(torch.as_tensor(a, device=device), torch.as_tensor(b, device=device),
                         op_a=opa, op_b=opb, slices=s).cpu().numpy()
        err = rel_err(got, ref)
        assert err < _bound(s), (s, err)
        errs.append(err)
    assert errs[-1] < 2e-13
    assert all(e1 < e0 for e0, e1 in zip(errs[:4], errs[1:5]))


def test_scaled_lines_zero_lines_alpha_beta(device):
    rng = np.random.default_rng(5)
    m = n = k = 96
    a = _cplx(rng, (2, m, k)) * (2.0 ** rng.integers(-300, 300, (2, m, 1)))
    b = _cplx(rng, (2, k, n)) * (2.0 ** rng.integers(-300, 300, (2, 1, n)))
    a[:, 5, :] = 0.0
    b[:, :, 7] = 0.0
    c = _cplx(rng, (2, m, n))
    alpha, beta = 0.5 - 1.5j, 2.0 + 0.25j
    out = torch.as_tensor(c, device=device)
    gemm_ozaki(torch.as_tensor(a, device=device), torch.as_tensor(b, device=device),
               alpha=alpha, beta=beta, out=out, slices=7)
    prod = a @ b
    got = out.cpu().numpy()
    assert np.all(got[:, 5, :] == beta * c[:, 5, :])
    # the row/column scales span 2^600: compare relative to each entry's scale
    scale = np.abs(a).max(-1, keepdims=True) * np.abs(b).max(-2, keepdims=True) * np.sqrt(k)
    scale[:, 5, :] = 1.0
    scale[:, :, 7] = 1.0
    err = np.abs(got - (alpha * prod + beta * c)) / (np.abs(alpha) * scale + np.abs(beta * c))
    assert err.max() < 1e-13
