"""Modular-arithmetic GEMM (Ozaki scheme II on int8 matrix cores) against
NumPy complex128.

Pinned by the truncation bound and by the exactness of the reconstruction:
- every line of op(A), op(B) is scaled to integers of T bits, so the
  truncation error is a small multiple of 2^-T (T = 43 for 12 moduli at
  k <= 64, 42 at k <= 256, rising by ~4 per two added moduli), and it must
  fall accordingly as moduli are added; the FP64 reconstruction adds a floor
  of ~N 2^-52 of M, which relative to random products of length k is about
  2^-52 4 sqrt(k) 2^(log2 M - 2T - log2(4 k)) < 2^-40; a wrong residue,
  weight or scale gives an O(1) error instead;
- a product at the bound the moduli must cover (all entries (1 + i)(1 - d):
  Im C' = 2 k 2^(2T) (1 - d)^2 with Re C' = 0) is recovered, which fails if
  M / 2 does not exceed the integer product;
- rows and columns of very different magnitude check the per-line power-of-two
  scaling, zero lines the exponent of an empty line, and alpha / beta the
  epilogue.
"""

import numpy as np
import pytest
import torch

from quatrex_amd.ops import gemm
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _cplx(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _ref(a, b, opa, opb):
    A = a if opa == "N" else np.conj(np.swapaxes(a, -1, -2))
    B = b if opb == "N" else np.conj(np.swapaxes(b, -1, -2))
    return A @ B


def _t_bits(n, k):
    """T of the library: the largest integer with 4 k_pad 2^(2T) < M."""
    moduli = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193,
              191, 181, 179, 173][:n]
    kp = -(-k // 64) * 64
    return int(np.floor((np.sum(np.log2(moduli)) - np.log2(4 * kp)) / 2 - 1e-9))


@pytest.mark.parametrize("mnk", [(64, 64, 64), (100, 37, 129), (256, 256, 256)],
                         ids=lambda t: "x".join(map(str, t)))
@pytest.mark.parametrize("ops", [("N", "N"), ("N", "C"), ("C", "N"), ("C", "C")],
                         ids=["NN", "NC", "CN", "CC"])
def test_error_falls_with_moduli(device, mnk, ops):
    m, n, k = mnk
    rng = np.random.default_rng(m + 2 * n + 3 * k)
    batch = 3
    opa, opb = ops
    a = _cplx(rng, (batch, m, k) if opa == "N" else (batch, k, m))
    b = _cplx(rng, (batch, k, n) if opb == "N" else (batch, n, k))
    ref = _ref(a, b, opa, opb)
    A, B = torch.as_tensor(a, device=device), torch.as_tensor(b, device=device)
    errs = []
    for nmod in (8, 10, 12, 14):
        got = gemm(A, B, op_a=opa, op_b=opb, precision=f"oz2-{nmod}").cpu().numpy()
        err = rel_err(got, ref)
        assert err < 8.0 * 2.0 ** -_t_bits(nmod, k) + 2.0 ** -40, (nmod, err)
        errs.append(err)
    assert errs[2] < 1e-12
    # truncation-dominated up to 12 moduli: ~16x per two moduli
    assert all(e1 < e0 / 4 for e0, e1 in zip(errs[:2], errs[1:3]))
    assert errs[3] <= errs[2] * 1.5


def test_product_at_the_moduli_bound(device):
    k = 256
    d = 1e-12
    a = np.full((2, 32, k), (1 + 1j) * (1 - d))
    A = torch.as_tensor(a, device=device)
    got = gemm(A, A, op_b="C", precision="oz2-12").cpu().numpy()
    # (1 + i)(1 - i) summed over k: 2 k (1 - d)^2, real
    ref = np.full((2, 32, 32), 2 * k * (1 - d) ** 2 + 0j)
    # every entry rounds the same way when scaled: the truncation adds up
    # coherently to ~2 2^-(T+1) = 2.3e-13 at T = 42
    assert rel_err(got, ref) < 1e-12
    got = gemm(A, torch.as_tensor(np.swapaxes(a, 1, 2).copy(), device=device),
               precision="oz2-12").cpu().numpy()
    # (1 + i)^2 = 2 i: imaginary part at the bound, real part zero
    assert rel_err(got, np.full((2, 32, 32), 2j * k * (1 - d) ** 2)) < 1e-12
    assert np.abs(got.real).max() < 1e-9


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
    gemm(torch.as_tensor(a, device=device), torch.as_tensor(b, device=device),
         alpha=alpha, beta=beta, out=out, precision="oz2-13")
    prod = a @ b
    got = out.cpu().numpy()
    assert np.all(got[:, 5, :] == beta * c[:, 5, :])
    # the row/column scales span 2^600: compare relative to each entry's scale
    scale = np.abs(a).max(-1, keepdims=True) * np.abs(b).max(-2, keepdims=True) * np.sqrt(k)
    scale[:, 5, :] = 1.0
    scale[:, :, 7] = 1.0
    err = np.abs(got - (alpha * prod + beta * c)) / (np.abs(alpha) * scale + np.abs(beta * c))
    assert err.max() < 1e-12
