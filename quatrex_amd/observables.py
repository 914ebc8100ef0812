"""Observables and system-matrix assembly on the HIP path.

- ``density(x, overlap=None)``: Im diag(G) or Im diag of the G S block
  products (quatrex/core/observables.py:57-126).  ``RGF.selected_solve(...,
  density=...)`` produces the no-overlap density of G^< inside the backward
  sweep without re-reading G.
- ``contact_currents(x_lesser, x_greater, obc_blocks)``:
  Tr[Sigma^>_c G^<_cc - G^>_cc Sigma^<_c] at both contacts
  (observables.py:129-190), block products on the MFMA GEMM.
- ``SystemMatrixAssembler``: A = (E + i eta)(I or S) - diag(V) - Sigma^R - H
  on the pattern of A (quatrex/electron/solver.py:390-407, ``_btd_subtract``
  :26-50).  The energy-independent part is mapped onto A's nnz pattern once;
  each call is one streaming kernel over the stack (HBM-bound: 16 B read of
  Sigma^R and 16 B written per nonzero and energy).
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sps
import torch

from quatrex_amd import _lib
from quatrex_amd.datastructures import DSDBCOO
from quatrex_amd.ops import as_device, gemm


def _btd_block(s: np.ndarray, off, i, j):
    return s[off[i]:off[i + 1], off[j]:off[j + 1]]


def density(x: DSDBCOO, overlap=None) -> torch.Tensor:
    """Im diag(G) over all orbitals, (*stack, n_orb) float64 on x's device."""
    n = x.num_blocks
    if overlap is None:
        return torch.cat([torch.diagonal(x.blocks[i, i], dim1=-2, dim2=-1).imag
                          for i in range(n)], dim=-1)
    off = np.concatenate([[0], np.cumsum(x.block_sizes)])
    s = sps.csr_matrix(overlap).toarray()
    dev = x.device
    lead = tuple(x.shape[:-2])

    def prod(i, j):
        blk = torch.as_tensor(np.ascontiguousarray(_btd_block(s, off, j, i)), device=dev)
        xb = x.blocks[i, j].contiguous()
        return gemm(xb, blk.expand(lead + blk.shape).contiguous())

    out = []
    for i in range(n):
        d = torch.diagonal(prod(i, i), dim1=-2, dim2=-1).clone()
        if i < n - 1:
            d += torch.diagonal(prod(i, i + 1), dim1=-2, dim2=-1)
        if i > 0:
            d += torch.diagonal(prod(i, i - 1), dim1=-2, dim2=-1)
        out.append(d.imag)
    return torch.cat(out, dim=-1)


def contact_currents(x_lesser: DSDBCOO, x_greater: DSDBCOO, obc_blocks):
    """(I_left, I_right) complex, one per stack entry."""
    n = x_lesser.num_blocks - 1
    dev = x_lesser.device

    def cur(j, k):
        sg = as_device(obc_blocks.greater[k], dev)
        sl = as_device(obc_blocks.lesser[k], dev)
        t = gemm(sg, x_lesser.blocks[j, j].contiguous())
        t = gemm(x_greater.blocks[j, j].contiguous(), sl, alpha=-1.0, beta=1.0, out=t)
        return torch.diagonal(t, dim1=-2, dim2=-1).sum(-1)

    return cur(0, 0), cur(n, -1)


def _pattern_values(m, rows, cols, shape) -> np.ndarray:
    """Values of a sparse/dense/DSDBCOO-free matrix at (rows, cols)."""
    if isinstance(m, np.ndarray):
        return m[rows, cols].astype(np.complex128)
    return np.asarray(sps.csr_matrix(m, shape=shape)[rows, cols]).ravel().astype(np.complex128)


def _btd_keep(rows, cols, offsets) -> np.ndarray:
    br = np.searchsorted(offsets, rows, side="right") - 1
    bc = np.searchsorted(offsets, cols, side="right") - 1
    return np.abs(br - bc) <= 1


class SystemMatrixAssembler:
    """A = (E + i eta)(I or S) - diag(V) - Sigma^R - H on the pattern of ``pattern``.

    ``hamiltonian`` and ``overlap`` are sparse (scipy) or dense host matrices
    of the device; only their block-tridiagonal entries on A's pattern enter,
    as ``_btd_subtract`` does.  ``potential`` is the on-site vector V.
    """

    def __init__(self, pattern: DSDBCOO, hamiltonian, overlap=None, potential=None):
        self.pattern = pattern
        rows, cols = pattern.rows, pattern.cols
        n = int(np.sum(pattern.block_sizes))
        offsets = np.concatenate([[0], np.cumsum(pattern.block_sizes)])
        keep = _btd_keep(rows, cols, offsets)
        diag = rows == cols
        s = np.where(diag, 1.0 + 0j, 0j) if overlap is None else _pattern_values(
            overlap, rows, cols, (n, n))
        t = -np.where(keep, _pattern_values(hamiltonian, rows, cols, (n, n)), 0)
        if potential is not None:
            t = t - np.where(diag, np.asarray(potential, dtype=np.float64)[rows], 0)
        dev = pattern.device
        self._s = torch.as_tensor(s, device=dev)
        self._t = torch.as_tensor(t, device=dev)
        self._keep = torch.as_tensor(keep, device=dev)

    def __call__(self, energies, eta: float, sse_retarded: DSDBCOO | None = None,
                 out: DSDBCOO | None = None) -> DSDBCOO:
        a = out if out is not None else DSDBCOO.zeros_like(self.pattern)
        _lib.require_device_tensor(a.data, "out.data")
        e = np.asarray(energies).reshape(-1)
        ne = a.data.shape[0] if a.data.dim() == 2 else int(np.prod(a.data.shape[:-1]))
        if e.size != ne:
            raise ValueError(f"{e.size} energies for a stack of {ne}")
        z = torch.as_tensor(e + 1j * eta, dtype=torch.complex128, device=a.device)
        data = a.data.reshape(ne, a.nnz)
        sig_ptr, sig_stride, keep_alive = None, 0, None
        if sse_retarded is not None:
            sig = sse_retarded.data.reshape(ne, -1)
            if not (np.array_equal(sse_retarded.rows, a.rows)
                    and np.array_equal(sse_retarded.cols, a.cols)):
                raise ValueError("sse_retarded must share the pattern of the system matrix")
            # only the block-tridiagonal entries are subtracted (_btd_subtract)
            if not bool(self._keep.all()):
                sig = sig * self._keep
                keep_alive = sig
            sig_ptr, sig_stride = sig.data_ptr(), sig.stride(0)
        h, st = _lib.handle(a.device), _lib.stream_ptr(a.device)
        rc = _lib.load().qx_assemble_system_matrix(
            h, st, ne, a.nnz, z.data_ptr(), self._s.data_ptr(), self._t.data_ptr(),
            ctypes.c_void_p(sig_ptr) if sig_ptr else None, sig_stride, data.data_ptr(),
            data.stride(0))
        _lib.check(rc, h, "assemble_system_matrix")
        del keep_alive
        return a
