"""Green's-function solvers: drop-in ``GFSolver`` / ``RGF`` on the HIP path.

Mirrors src/qttools/greens_function_solver/solver.py:12-113 (``OBCBlocks``,
``GFSolver``) and rgf.py:10-414 (``RGF``): same signatures, same batching
over the first stack axis (``get_batches``), same outputs and errors:

- ``selected_inv(a, obc_blocks=None, out=None)`` returns the selected inverse
  (diagonal and both off-diagonal block bands) or writes ``out``.
- ``selected_solve(a, sigma_lesser, sigma_greater, obc_blocks=None, out=None,
  return_retarded=False, return_current=False)`` returns ``(xl, xg[, xr]
  [, current])`` or writes ``out=(xl, xg[, xr])`` and returns ``current`` or
  None; ``ValueError("Invalid number of output matrices.")`` as rgf.py:185-186.
  Diagonal lesser/greater blocks are the anti-Hermitian projection
  0.5 (X - X^H), lower blocks are -(upper)^H unless the output container is
  symmetric-stored, and only diagonal retarded blocks are written (unless
  ``RGF(retarded_offdiag=True)``).

Every block product and inverse runs in libquatrex_amd on the tensors'
HIP device; the containers' fully populated blocks are read and written in
place.  An exactly singular block raises ``numpy.linalg.LinAlgError`` as the
reference's NumPy backend does.
"""

from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod

import numpy as np
import torch

from quatrex_amd import _lib
from quatrex_amd.datastructures import DSDBCOO
from quatrex_amd.utils import get_batches


class OBCBlocks:
    """Per-block boundary self-energies (solver.py:12-28)."""

    def __init__(self, num_blocks: int):
        self.retarded: list = [None] * num_blocks
        self.lesser: list = [None] * num_blocks
        self.greater: list = [None] * num_blocks


class GFSolver(ABC):
    """Abstract Green's-function solver (solver.py:31-113)."""

    @abstractmethod
    def selected_inv(self, a, obc_blocks=None, out=None):
        ...

    @abstractmethod
    def selected_solve(self, a, sigma_lesser, sigma_greater, obc_blocks=None, out=None,
                       return_retarded=False, return_current=False):
        ...


class _Bound:
    """One container's layout for one stack batch, possibly via a staging buffer."""

    def __init__(self, mat: DSDBCOO, s0: int, s1: int, *, output: bool,
                 write_lower: bool = True, write_offdiag: bool = True):
        if not isinstance(mat, DSDBCOO):
            raise TypeError("quatrex_amd solvers take quatrex_amd.datastructures.DSDBCOO")
        _lib.require_device_tensor(mat.data, "DSDBCOO.data")
        if mat.dtype != torch.complex128:
            raise TypeError("quatrex_amd solvers store complex128 data "
                            f"(got {mat.dtype}); choose the arithmetic with RGF(precision=...)")
        self.mat, self.s0, self.s1, self.output = mat, s0, s1, output
        n = mat.num_blocks
        inner = int(np.prod(mat.stack_shape[1:], dtype=np.int64))
        self.S = (s1 - s0) * inner
        bs = mat.block_sizes
        # Which blocks are read or written, and where: a function of the
        # pattern, the symmetry and the flags only, kept in the pattern cache
        # that the container shares with its views and zeros_like copies
        # (with the layouts built from it), so repeated calls on one pattern
        # skip the per-block walk.
        plan_key = ("_qx_plan", bool(output), bool(write_lower), bool(write_offdiag),
                    bool(mat.symmetry))
        plan = mat._slices.get(plan_key)
        if plan is None:
            plan = self._plan(mat, write_lower, write_offdiag)
            mat._slices[plan_key] = plan
        want, direct, staged = plan["want"], plan["direct"], plan["staged"]
        self.staging = None

        def offsets():
            diag = np.full(n, -1, np.int64)
            upper = np.full(max(n - 1, 0), -1, np.int64)
            lower = np.full(max(n - 1, 0), -1, np.int64)
            return diag, upper, lower, {"d": diag, "u": upper, "l": lower}

        # A read-only stack broadcast over its only stack axis (stride 0, e.g.
        # an energy-independent self-energy expanded to the batch) is passed
        # with stack stride 0: every matrix of the batch reads the same data.
        # Any other non-contiguous input (a broadcast over the first of
        # several stack axes repeats with the period of the inner axes, which
        # one stack stride cannot express) is copied contiguously for this
        # batch; outputs must be contiguous.
        data = mat.data
        bcast = (not output and data.dim() == 2 and data.shape[0] > 1 and data.stride(0) == 0
                 and data.stride(1) == 1)
        self._copy = None
        if not staged:
            if bcast:
                self.base, stride = data.data_ptr(), 0
            elif data.is_contiguous():
                self.base = data.data_ptr() + s0 * inner * mat.nnz * data.element_size()
                stride = mat.nnz
            elif output:
                raise ValueError("an output DSDBCOO.data must be contiguous")
            else:
                self._copy = data[s0:s1].reshape(self.S, mat.nnz).contiguous()
                self.base, stride = self._copy.data_ptr(), mat.nnz
            lay = plan["layouts"].get((self.S, stride))
            if lay is None:
                diag, upper, lower, arrs = offsets()
                for (kind, i), off in direct.items():
                    arrs[kind][i] = off
                lay = plan["layouts"][(self.S, stride)] = _lib.make_layout(
                    bs, self.S, stride, diag, upper, lower)
            self.layout, self._keep = lay
        else:
            # all blocks of this container go through one dense staging buffer
            diag, upper, lower, arrs = offsets()
            keys = sorted(set(direct) | set(staged), key=lambda k: (k[1], k[0]))
            off, offs = 0, {}
            for key in keys:
                i, j = want[key]
                offs[key] = off
                off += int(bs[i]) * int(bs[j])
            self.staging = torch.zeros((self.S, max(off, 1)), dtype=torch.complex128,
                                       device=mat.device)
            self._staged_blocks = {}
            view = mat.stack[s0:s1]
            for key in keys:
                i, j = want[key]
                arrs[key[0]][key[1]] = offs[key]
                seg = self.staging[:, offs[key]:offs[key] + int(bs[i]) * int(bs[j])]
                self._staged_blocks[key] = (i, j, seg)
                if not output:
                    seg.copy_(view.blocks[i, j].reshape(self.S, -1))
            self.base = self.staging.data_ptr()
            stride = self.staging.shape[1]
            self.layout, self._keep = _lib.make_layout(bs, self.S, stride, diag, upper, lower)

    @staticmethod
    def _plan(mat, write_lower, write_offdiag):
        n = mat.num_blocks
        want = {}
        for i in range(n):
            want[("d", i)] = (i, i)
            if i + 1 < n and write_offdiag:
                want[("u", i)] = (i, i + 1)
                if write_lower:
                    want[("l", i)] = (i + 1, i)
        direct, staged = {}, []
        for key, (i, j) in want.items():
            if mat.symmetry and (j < i or i == j):
                # symmetric storage: diagonal blocks hold a triangle, lower
                # blocks are implied -> densify through a staging buffer
                present = mat.block_slice(min(i, j), max(i, j)) is not None
                if present:
                    staged.append(key)
                continue
            sl = mat.block_slice(i, j)
            if sl is None:
                continue
            if mat.is_dense_block(i, j):
                direct[key] = sl.start
            else:
                staged.append(key)
        return {"want": want, "direct": direct, "staged": staged, "layouts": {}}

    def ptr(self):
        return ctypes.c_void_p(self.base)

    def lay(self):
        return ctypes.byref(self.layout)

    def write_back(self):
        if self.staging is None or not self.output:
            return
        view = self.mat.stack[self.s0:self.s1]
        lead = (self.s1 - self.s0,) + tuple(self.mat.stack_shape[1:])
        for key, (i, j, seg) in self._staged_blocks.items():
            bi, bj = int(self.mat.block_sizes[i]), int(self.mat.block_sizes[j])
            view.blocks[i, j] = seg.reshape(lead + (bi, bj))


def _obc_array(blocks, s0, s1, n, device, keep):
    """Host array of device pointers to the (S, b, b) slices, or None."""
    if blocks is None or all(b is None for b in blocks):
        return None
    arr = (ctypes.c_void_p * n)()
    for j in range(n):
        b = blocks[j]
        if b is None:
            arr[j] = None
            continue
        t = torch.as_tensor(b)
        if t.device.type != "cuda":
            t = t.to(device)
        t = t[s0:s1].to(torch.complex128).contiguous()
        keep.append(t)
        arr[j] = t.data_ptr()
    return arr


class RGF(GFSolver):
    """Recursive Green's function solver on the HIP path (rgf.py:10-23).

    Parameters
    ----------
    max_batch_size : int
        Largest number of first-axis stack entries per batch (as the
        reference; batches may exceed it, see ``get_batches``).
    precision : {"c128", "c128-3m", "c64", "mixed", "oz2-8" ... "oz2-20", "bf16x2" ... "bf16x7"}
        Arithmetic of the block products and inverses: complex128 on FP64
        matrix cores (4M complex split; "c128-3m": 3M split, a quarter less
        matrix-core work, rounding error a small factor larger); complex64
        on FP32 matrix cores; "mixed", complex64 products and FP64 block
        inverses (with QX_MIXED_INVERSE=c64: complex64 inverses refined by
        one FP64 Newton-Schulz step, FP64 re-inversion of failed blocks); or
        "oz2-n", block products on int8 matrix cores by modular arithmetic
        with n moduli (relative error ~2^-T per product of the line maxima,
        T = 42 for n = 12 at blocks <= 256) and FP64 block inverses; or
        "bf16xs", block products on bf16 matrix cores from s 7-bit digit
        slices per operand line (relative error ~2^(-7 s)) and FP64 block
        inverses.  Inputs and outputs stay complex128 in every mode.
    retarded_offdiag : bool
        Also write the off-diagonal retarded blocks in ``selected_solve``.
    """

    def __init__(self, max_batch_size: int = 100, precision: str = "c128",
                 retarded_offdiag: bool = False) -> None:
        self.max_batch_size = max_batch_size
        if precision not in _lib.RGF_PRECISIONS:
            raise ValueError(f"Unknown precision {precision!r}")
        self.precision = precision
        self.retarded_offdiag = retarded_offdiag

    # rgf.py:25-113
    def selected_inv(self, a: DSDBCOO, obc_blocks: OBCBlocks | None = None,
                     out: DSDBCOO | None = None):
        lib = _lib.load()
        _lib.require_device_tensor(a.data, "a.data")
        dev = a.device
        if obc_blocks is None:
            obc_blocks = OBCBlocks(num_blocks=a.num_blocks)
        x = out if out is not None else a.__class__.zeros_like(a)
        sizes, slices = get_batches(a.shape[0], self.max_batch_size)
        h, st = _lib.handle(dev), _lib.stream_ptr(dev)
        prec = _lib.RGF_PRECISIONS[self.precision]
        for b in range(len(sizes)):
            s0, s1 = int(slices[b]), int(slices[b + 1])
            keep: list = []
            A = _Bound(a, s0, s1, output=False)
            X = _Bound(x, s0, s1, output=True, write_lower=not x.symmetry)
            obc = _obc_array(obc_blocks.retarded, s0, s1, a.num_blocks, dev, keep)
            rc = lib.qx_rgf_selected_inv(h, st, A.lay(), A.ptr(), obc, X.lay(), X.ptr(), prec)
            _lib.check(rc, h, "RGF.selected_inv")
            X.write_back()
        _lib.sync_check(dev, "RGF.selected_inv")
        if out is None:
            return x
        return None

    # rgf.py:115-414
    def selected_solve(self, a: DSDBCOO, sigma_lesser: DSDBCOO, sigma_greater: DSDBCOO,
                       obc_blocks: OBCBlocks | None = None, out=None,
                       return_retarded: bool = False, return_current: bool = False,
                       *, corner: torch.Tensor | None = None,
                       density: torch.Tensor | None = None,
                       ldos: torch.Tensor | None = None):
        """rgf.py:115-414.

        Extensions (keyword-only, filled in place):
        ``corner`` (complex128 device tensor (*stack, b_0, b_{N-1})) receives
        G^R_{0,N-1} for the transmission; ``density`` (float64 device tensor
        (*stack, sum of block sizes)) receives Im diag(G^<), the electron
        density of quatrex/core/observables.py:57-86 (no overlap), written as
        each diagonal block is finished; ``ldos`` (same shape) receives the
        electron LDOS -Im diag(G^R) / (2 pi) of quatrex/core/scba.py:572-576
        (no overlap), whether or not ``return_retarded`` is set.
        """
        lib = _lib.load()
        _lib.require_device_tensor(a.data, "a.data")
        dev = a.device
        n = a.num_blocks
        if obc_blocks is None:
            obc_blocks = OBCBlocks(num_blocks=n)
        current = None
        if return_current:
            current = torch.zeros((*a.shape[:-2], n - 1), dtype=a.dtype, device=dev)
        sizes, slices = get_batches(a.shape[0], self.max_batch_size)
        if out is not None:
            xl, xg, *xr = out
            if return_retarded:
                if len(xr) != 1:
                    raise ValueError("Invalid number of output matrices.")
                xr = xr[0]
        else:
            xl = a.__class__.zeros_like(a)
            xg = a.__class__.zeros_like(a)
            if return_retarded:
                xr = a.__class__.zeros_like(a)
        flags = 0
        if return_retarded:
            flags |= _lib.SOLVE_RETARDED
            if self.retarded_offdiag:
                flags |= _lib.SOLVE_RETARDED_OFFDIAG
        if return_current:
            flags |= _lib.SOLVE_CURRENT
        if xl.symmetry:
            flags |= _lib.SOLVE_SKIP_LOWER_L
        if xg.symmetry:
            flags |= _lib.SOLVE_SKIP_LOWER_G
        h, st = _lib.handle(dev), _lib.stream_ptr(dev)
        prec = _lib.RGF_PRECISIONS[self.precision]
        inner = int(np.prod(a.stack_shape[1:], dtype=np.int64))
        for b in range(len(sizes)):
            s0, s1 = int(slices[b]), int(slices[b + 1])
            keep: list = []
            A = _Bound(a, s0, s1, output=False)
            SL = _Bound(sigma_lesser, s0, s1, output=False)
            SG = _Bound(sigma_greater, s0, s1, output=False)
            XL = _Bound(xl, s0, s1, output=True, write_lower=not xl.symmetry)
            XG = _Bound(xg, s0, s1, output=True, write_lower=not xg.symmetry)
            XR = None
            if return_retarded:
                XR = _Bound(xr, s0, s1, output=True, write_lower=not xr.symmetry,
                            write_offdiag=self.retarded_offdiag)
            o_r = _obc_array(obc_blocks.retarded, s0, s1, n, dev, keep)
            o_l = _obc_array(obc_blocks.lesser, s0, s1, n, dev, keep)
            o_g = _obc_array(obc_blocks.greater, s0, s1, n, dev, keep)
            corner_ptr = None
            if corner is not None:
                nl, n0 = int(a.block_sizes[-1]), int(a.block_sizes[0])
                if (corner.dtype != torch.complex128 or not corner.is_contiguous()
                        or tuple(corner.shape) != tuple(a.shape[:-2]) + (n0, nl)):
                    raise ValueError("corner must be a contiguous complex128 tensor of shape "
                                     "(*stack, b_0, b_{N-1})")
                corner_ptr = ctypes.c_void_p(corner.data_ptr()
                                             + s0 * inner * nl * n0 * corner.element_size())
            n_orb = int(np.sum(a.block_sizes))

            def orbital_ptr(t, name):
                if t is None:
                    return None
                if (t.dtype != torch.float64 or not t.is_contiguous()
                        or tuple(t.shape) != tuple(a.shape[:-2]) + (n_orb,)):
                    raise ValueError(f"{name} must be a contiguous float64 tensor of shape "
                                     "(*stack, sum of block sizes)")
                return ctypes.c_void_p(t.data_ptr() + s0 * inner * n_orb * t.element_size())

            dens_ptr = orbital_ptr(density, "density")
            ldos_ptr = orbital_ptr(ldos, "ldos")
            cur_ptr = None
            if current is not None:
                cur_ptr = ctypes.c_void_p(current.data_ptr()
                                          + s0 * inner * (n - 1) * current.element_size())
            rc = lib.qx_rgf_selected_solve(
                h, st, A.lay(), A.ptr(), SL.lay(), SL.ptr(), SG.lay(), SG.ptr(), o_r, o_l, o_g,
                XL.lay(), XL.ptr(), XG.lay(), XG.ptr(),
                XR.lay() if XR else None, XR.ptr() if XR else None, cur_ptr, corner_ptr,
                dens_ptr, ldos_ptr, flags, prec)
            _lib.check(rc, h, "RGF.selected_solve")
            for X in (XL, XG, XR):
                if X is not None:
                    X.write_back()
        _lib.sync_check(dev, "RGF.selected_solve")
        if out is None:
            if return_retarded:
                if return_current:
                    return xl, xg, xr, current
                return xl, xg, xr
            return xl, xg
        if return_current:
            return current
        return None


class Inv(GFSolver):
    """Dense selected inversion / solve on the HIP path (inv.py:10-227).

    Densifies the stack, folds the boundary blocks into the dense matrices,
    inverts them with the batched HIP inverse, forms X^<> = X Sigma^<> X^H
    with the HIP GEMM and gathers the entries of the output pattern.  As the
    reference, it is meant for checks on small systems, not production:
    the dense size (sum of the block sizes) is limited to 512 here.

    Reference behaviour kept: ``selected_solve`` subtracts the lesser /
    greater boundary blocks (inv.py:199-202, where RGF adds them), always
    returns its outputs (inv.py:224-227), raises
    ``ValueError("Missing output for the retarded Green's function.")``
    (inv.py:172-176) and ``NotImplementedError`` for ``return_current``
    (inv.py:148-151).  Boundary blocks fold on the last two axes, so stacks
    of any dimension work.
    """

    MAX_DENSE = 512

    def __init__(self, max_batch_size: int = 100) -> None:
        self.max_batch_size = max_batch_size

    def _dense(self, m: DSDBCOO, sl: slice) -> torch.Tensor:
        if not isinstance(m, DSDBCOO):
            raise TypeError("quatrex_amd solvers take quatrex_amd.datastructures.DSDBCOO")
        _lib.require_device_tensor(m.data, "DSDBCOO.data")
        return m.stack_view(sl.start, sl.stop).to_dense()

    def _check(self, a: DSDBCOO) -> None:
        n = int(np.sum(a.block_sizes))
        if n > self.MAX_DENSE:
            raise ValueError(f"Inv densifies the matrix: dense size {n} exceeds "
                             f"{self.MAX_DENSE}; use RGF")

    @staticmethod
    def _fold(dense, blocks, offsets, sl, sign):
        for j, blk in enumerate(blocks):
            if blk is None:
                continue
            r = slice(int(offsets[j]), int(offsets[j + 1]))
            dense[..., r, r] += sign * torch.as_tensor(blk)[sl].to(dense.device, dense.dtype)

    @staticmethod
    def _gather(out: DSDBCOO, dense: torch.Tensor, sl: slice) -> None:
        r = torch.as_tensor(out.rows, dtype=torch.long, device=dense.device)
        c = torch.as_tensor(out.cols, dtype=torch.long, device=dense.device)
        out.data[sl] = dense[..., r, c]

    # inv.py:32-97
    def selected_inv(self, a: DSDBCOO, obc_blocks: OBCBlocks | None = None,
                     out: DSDBCOO | None = None):
        from quatrex_amd.ops import inv

        self._check(a)
        sizes, slices = get_batches(a.shape[0], self.max_batch_size)
        if obc_blocks is None:
            obc_blocks = OBCBlocks(num_blocks=a.num_blocks)
        return_out = out is None
        if out is None:
            out = a.__class__.zeros_like(a)
        offs = np.concatenate([[0], np.cumsum(a.block_sizes)])
        for k in range(len(sizes)):
            sl = slice(int(slices[k]), int(slices[k + 1]))
            ad = self._dense(a, sl)
            self._fold(ad, obc_blocks.retarded, offs, sl, -1.0)
            self._gather(out, inv(ad), sl)
        if return_out:
            return out
        return None

    # inv.py:99-227
    def selected_solve(self, a: DSDBCOO, sigma_lesser: DSDBCOO, sigma_greater: DSDBCOO,
                       obc_blocks: OBCBlocks | None = None, out=None,
                       return_retarded: bool = False, return_current: bool = False):
        from quatrex_amd.ops import gemm, inv

        if return_current:
            raise NotImplementedError("The computation of the current is not implemented.")
        self._check(a)
        sizes, slices = get_batches(a.shape[0], self.max_batch_size)
        if obc_blocks is None:
            obc_blocks = OBCBlocks(num_blocks=a.num_blocks)
        if out is None:
            xl, xg = a.__class__.zeros_like(a), a.__class__.zeros_like(a)
            xr = a.__class__.zeros_like(a) if return_retarded else None
        else:
            xl, xg, *rest = out
            xr = None
            if return_retarded:
                if len(rest) == 0:
                    raise ValueError("Missing output for the retarded Green's function.")
                xr = rest[0]
        offs = np.concatenate([[0], np.cumsum(a.block_sizes)])
        for k in range(len(sizes)):
            sl = slice(int(slices[k]), int(slices[k + 1]))
            ad = self._dense(a, sl)
            sld, sgd = self._dense(sigma_lesser, sl), self._dense(sigma_greater, sl)
            self._fold(ad, obc_blocks.retarded, offs, sl, -1.0)
            self._fold(sld, obc_blocks.lesser, offs, sl, -1.0)
            self._fold(sgd, obc_blocks.greater, offs, sl, -1.0)
            x = inv(ad)
            self._gather(xl, gemm(gemm(x, sld), x, op_b="C"), sl)
            self._gather(xg, gemm(gemm(x, sgd), x, op_b="C"), sl)
            if return_retarded:
                self._gather(xr, x, sl)
        if return_retarded:
            return xl, xg, xr
        return xl, xg


def transmission(corner, gamma_l, gamma_r) -> torch.Tensor:
    """T(E) = Tr[Gamma_L G^R_{0,N-1} Gamma_R (G^R_{0,N-1})^H] on the device.

    ``corner`` from ``RGF.selected_solve(..., corner=...)``; the broadening
    matrices are (*stack, b, b) stacks (Gamma = i (Sigma^R - Sigma^A)).
    """
    from quatrex_amd.ops import as_device, gemm

    g = as_device(corner)
    t = gemm(gemm(as_device(gamma_l, g.device), g), as_device(gamma_r, g.device))
    t = gemm(t, g, op_b="C")
    return torch.diagonal(t, dim1=-2, dim2=-1).sum(-1).real
