"""Batched complex linear algebra on torch device tensors through the C ABI.

``gemm`` and ``inv`` are the two workhorses of the open-boundary solvers:
they run libquatrex_amd's MFMA GEMM and pivoted Gauss-Jordan inverse on the
current torch stream.  Inputs are (..., m, k) stacks; they are made
contiguous complex128 on the HIP device (no CPU fallback).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from quatrex_amd import _lib


def as_device(x, device=None) -> torch.Tensor:
    """complex128 HIP tensor view/copy of ``x`` (numpy arrays are uploaded)."""
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(x))
    if device is None:
        device = t.device if t.device.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
    t = t.to(device=device, dtype=torch.complex128)
    _lib.require_device_tensor(t, "operand")
    return t.contiguous()


def _flat(t: torch.Tensor):
    return t.reshape((-1,) + tuple(t.shape[-2:]))


def gemm(a, b, op_a: str = "N", op_b: str = "N", alpha=1.0, beta=0.0, out=None,
         precision: str = "c128") -> torch.Tensor:
    """out = alpha op(a) op(b) + beta out, op in {"N", "C"} (conjugate transpose).

    Leading (batch) dimensions of ``a`` and ``b`` must be equal.
    """
    a = as_device(a)
    b = as_device(b, a.device)
    ops = {"N": _lib.OP_N, "C": _lib.OP_C}
    oa, ob = ops[op_a], ops[op_b]
    m, k = (a.shape[-2], a.shape[-1]) if oa == _lib.OP_N else (a.shape[-1], a.shape[-2])
    kb, n = (b.shape[-2], b.shape[-1]) if ob == _lib.OP_N else (b.shape[-1], b.shape[-2])
    if k != kb:
        raise ValueError(f"inner dimensions differ: {k} != {kb}")
    if a.shape[:-2] != b.shape[:-2]:
        raise ValueError("batch shapes of a and b differ")
    a3, b3 = _flat(a), _flat(b)
    batch = a3.shape[0]
    if out is None:
        if beta != 0:
            raise ValueError("beta != 0 needs out")
        out = torch.empty(a.shape[:-2] + (m, n), dtype=torch.complex128, device=a.device)
    if out.dtype != torch.complex128 or not out.is_contiguous():
        raise ValueError("out must be a contiguous complex128 tensor")
    c3 = _flat(out)
    al = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
    be = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
    h, st = _lib.handle(a.device), _lib.stream_ptr(a.device)
    rc = _lib.load().qx_zgemm_batched(
        h, st, batch, m, n, k, oa, ob, al, a3.data_ptr(), a3.shape[-1], a3.stride(0),
        b3.data_ptr(), b3.shape[-1], b3.stride(0), be, c3.data_ptr(), c3.shape[-1],
        c3.stride(0), _lib.PRECISIONS[precision])
    _lib.check(rc, h, "gemm")
    return out


def gemm_ozaki(a, b, op_a: str = "N", op_b: str = "N", alpha=1.0, beta=0.0, out=None,
               slices: int = 7) -> torch.Tensor:
    """Like ``gemm`` (complex128 in and out), computed on bf16 matrix cores from
    ``slices`` 7-bit digit slices of the operands (Ozaki-style split)."""
    a = as_device(a)
    b = as_device(b, a.device)
    ops = {"N": _lib.OP_N, "C": _lib.OP_C}
    oa, ob = ops[op_a], ops[op_b]
    m, k = (a.shape[-2], a.shape[-1]) if oa == _lib.OP_N else (a.shape[-1], a.shape[-2])
    kb, n = (b.shape[-2], b.shape[-1]) if ob == _lib.OP_N else (b.shape[-1], b.shape[-2])
    if k != kb:
        raise ValueError(f"inner dimensions differ: {k} != {kb}")
    if a.shape[:-2] != b.shape[:-2]:
        raise ValueError("batch shapes of a and b differ")
    a3, b3 = _flat(a), _flat(b)
    if out is None:
        if beta != 0:
            raise ValueError("beta != 0 needs out")
        out = torch.empty(a.shape[:-2] + (m, n), dtype=torch.complex128, device=a.device)
    c3 = _flat(out)
    al = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
    be = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
    h, st = _lib.handle(a.device), _lib.stream_ptr(a.device)
    rc = _lib.load().qx_zgemm_ozaki(
        h, st, a3.shape[0], m, n, k, oa, ob, al, a3.data_ptr(), a3.shape[-1], a3.stride(0),
        b3.data_ptr(), b3.shape[-1], b3.stride(0), be, c3.data_ptr(), c3.shape[-1],
        c3.stride(0), int(slices))
    _lib.check(rc, h, "gemm_ozaki")
    return out


def mm(*factors, precision: str = "c128") -> torch.Tensor:
    """Left-to-right product of several stacks (as ``a @ b @ c`` in NumPy)."""
    out = factors[0]
    for f in factors[1:]:
        out = gemm(out, f, precision=precision)
    return out


def inv(a, precision: str = "c128", check: bool = True) -> torch.Tensor:
    """Batched inverse with partial pivoting; raises LinAlgError if singular."""
    a = as_device(a)
    n = a.shape[-1]
    if a.shape[-2] != n:
        raise ValueError("inv needs square matrices")
    a3 = _flat(a)
    out = torch.empty_like(a)
    x3 = _flat(out)
    h, st = _lib.handle(a.device), _lib.stream_ptr(a.device)
    rc = _lib.load().qx_inv_batched(h, st, a3.shape[0], n, a3.data_ptr(), n, n * n, None, n,
                                    n * n, x3.data_ptr(), n, n * n, _lib.PRECISIONS[precision])
    _lib.check(rc, h, "inv")
    if check:
        _lib.sync_check(a.device, "inv")
    return out


def svd(a, max_sweeps: int = 30, tol: float = 1e-15):
    """Thin SVD a = u diag(s) v^H of (..., rows, cols) stacks, cols <= rows
    (one-sided Jacobi on the device); s descending.  Returns (u, s, v)."""
    a = as_device(a)
    rows, cols = a.shape[-2], a.shape[-1]
    if cols > rows:
        raise ValueError("svd needs cols <= rows")
    a3 = _flat(a)
    n = a3.shape[0]
    u = torch.empty_like(a)
    v = torch.empty(a.shape[:-2] + (cols, cols), dtype=torch.complex128, device=a.device)
    s = torch.empty(a.shape[:-2] + (cols,), dtype=torch.float64, device=a.device)
    h, st = _lib.handle(a.device), _lib.stream_ptr(a.device)
    rc = _lib.load().qx_svd_batched(h, st, n, rows, cols, a3.data_ptr(), cols, rows * cols,
                                    u.data_ptr(), s.data_ptr(), v.data_ptr(), int(max_sweeps),
                                    float(tol), None)
    _lib.check(rc, h, "svd")
    return u, s, v


def spectral_modes(ws, vs, av_ji, av_ii, av_ij, *, min_decay: float, max_decay: float,
                   min_propagation: float, residual_tolerance: float, eta_decay: float,
                   residual_normalization: bool = True, details: bool = False):
    """Reflected-mode mask of the Spectral OBC from eigenpairs (ws (n, m), vs
    (n, b, m)) and the products a_ji V, a_ii V, a_ij V, in one fused kernel
    (spectral.py:160-320).  Returns the boolean mask (n, m); with
    ``details`` also the residuals and dE/dk."""
    ws, vs = as_device(ws).contiguous(), as_device(vs).contiguous()
    avs = [as_device(a).contiguous() for a in (av_ji, av_ii, av_ij)]
    n, b, m = vs.shape
    if ws.shape != (n, m) or any(a.shape != vs.shape for a in avs):
        raise ValueError("spectral_modes: ws must be (n, m), vs and the products (n, b, m)")
    mask = torch.empty((n, m), dtype=torch.uint8, device=vs.device)
    res = torch.empty((n, m), dtype=torch.float64, device=vs.device) if details else None
    dk = torch.empty((n, m), dtype=torch.complex128, device=vs.device) if details else None
    h, st = _lib.handle(vs.device), _lib.stream_ptr(vs.device)
    rc = _lib.load().qx_spectral_modes(
        h, st, n, b, m, ws.data_ptr(), vs.data_ptr(), avs[0].data_ptr(), avs[1].data_ptr(),
        avs[2].data_ptr(), float(min_decay), float(max_decay), float(min_propagation),
        float(residual_tolerance), float(eta_decay), int(bool(residual_normalization)),
        mask.data_ptr(), res.data_ptr() if details else None, dk.data_ptr() if details else None)
    _lib.check(rc, h, "spectral_modes")
    mask = mask.bool()
    return (mask, res, dk) if details else mask


def dagger(x: torch.Tensor) -> torch.Tensor:
    return x.conj().transpose(-2, -1).resolve_conj().contiguous()
