"""DSDBCOO: stack of block-sparse COO matrices held in device memory.

Drop-in for the single-rank use of ``qttools.datastructures.DSDBCOO``
(src/qttools/datastructures/dsdbcoo.py:23-1008, base class
dsdbsparse.py:109-1142).  Same storage: ``data`` has shape
``(*stack_shape, nnz)``, nonzeros are ordered by (block-row, block-col) and
row-major inside a block, so a fully populated block (i, j) is one
contiguous ``b_i x b_j`` slice of the last axis.  The HIP solvers read and
write such blocks in place through pointer offsets (no gather); blocks with
partial sparsity or symmetric storage are staged through dense buffers.

``data`` is a torch tensor (complex128 on a HIP device for the solvers);
``rows``/``cols`` are host int32 arrays, the block metadata is host-side.
Only the stack distribution of one rank is modelled: distributed
transposition (``dtranspose``) and MPI block partitioning are not part of
this path.
"""

from __future__ import annotations

from typing import Callable

import numpy as np
import scipy.sparse as sps
import torch


def _conj(x: torch.Tensor) -> torch.Tensor:
    return x.conj().resolve_conj()


def block_sort_index(rows: np.ndarray, cols: np.ndarray, block_sizes) -> np.ndarray:
    """Stable (block-row, block-col) ordering of COO entries.

    Same grouping as compute_block_sort_index
    (src/qttools/kernels/datastructure/numba/dsdbcoo.py:172-236).
    """
    offsets = np.hstack(([0], np.cumsum(block_sizes)))
    brow = np.searchsorted(offsets, rows, side="right") - 1
    bcol = np.searchsorted(offsets, cols, side="right") - 1
    return np.argsort(brow.astype(np.int64) * len(block_sizes) + bcol, kind="stable")


class _BlockIndexer:
    """``.blocks[i, j]`` accessor (dsdbsparse.py:1064-1142)."""

    def __init__(self, mat: "DSDBCOO", stack_index: tuple):
        self._m = mat
        self._idx = stack_index

    def _check(self, index):
        if len(index) != 2:
            raise IndexError("Exactly two block indices are required.")
        i, j = index
        if isinstance(i, slice) or isinstance(j, slice):
            raise NotImplementedError("Slicing is not supported.")
        if i < 0 or j < 0:
            raise IndexError("Negative block indices are not supported.")
        if i >= self._m.num_blocks or j >= self._m.num_blocks:
            raise IndexError("Block index out of bounds.")
        return int(i), int(j)

    def __getitem__(self, index):
        return self._m._get_block(self._idx, *self._check(index))

    def __setitem__(self, index, block):
        self._m._set_block(self._idx, *self._check(index), block)


class _StackView:
    """``a.stack[...]`` view (dsdbsparse.py:959-1061)."""

    def __init__(self, mat: "DSDBCOO", index):
        if not isinstance(index, tuple):
            index = (index,)
        self._dsdbsparse = mat
        self._stack_index = index
        self.symmetry = mat.symmetry
        self.blocks = _BlockIndexer(mat, index)


class _StackIndexer:
    def __init__(self, mat):
        self._m = mat

    def __getitem__(self, index):
        return _StackView(self._m, index)


class DSDBCOO:
    """Block-sorted COO stack on a torch device (see module docstring)."""

    def __init__(self, data: torch.Tensor, rows, cols, block_sizes, global_stack_shape,
                 return_dense: bool = True, symmetry: bool | None = False,
                 symmetry_op: Callable = _conj):
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.asarray(data))
        # A read-only stack may be one entry broadcast over the first axis
        # (stride 0); it is kept as is.  Everything else is made contiguous.
        bcast = data.dim() >= 2 and data.shape[0] > 1 and data.stride(0) == 0 and \
            data[0].is_contiguous()
        self._data = data if bcast else data.contiguous()
        self.rows = np.asarray(rows, dtype=np.int32)
        self.cols = np.asarray(cols, dtype=np.int32)
        if self.rows.shape != self.cols.shape or self.rows.shape[0] != data.shape[-1]:
            raise ValueError("rows, cols and the last data axis must have equal length.")
        self._block_sizes = np.asarray(block_sizes, dtype=np.int64)
        self._block_offsets = np.hstack(([0], np.cumsum(self._block_sizes)))
        self.global_stack_shape = tuple(global_stack_shape)
        self.return_dense = return_dense
        self.symmetry = symmetry
        self.symmetry_op = symmetry_op
        self.num_blocks = len(self._block_sizes)
        self._slices: dict = {}
        self._dense: dict = {}
        self._dev_idx: dict = {}
        self.stack = _StackIndexer(self)
        self._blocks = _BlockIndexer(self, (Ellipsis,))

    # -- properties -----------------------------------------------------
    @property
    def data(self) -> torch.Tensor:
        return self._data

    @data.setter
    def data(self, value):
        value = torch.as_tensor(value).to(self._data.device, self._data.dtype)
        self._data[...] = value

    @property
    def block_sizes(self) -> np.ndarray:
        return self._block_sizes

    @property
    def block_offsets(self) -> np.ndarray:
        return self._block_offsets

    @property
    def stack_shape(self) -> tuple:
        return tuple(self._data.shape[:-1])

    @property
    def shape(self) -> tuple:
        n = int(self._block_offsets[-1])
        return self.stack_shape + (n, n)

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def device(self) -> torch.device:
        return self._data.device

    @property
    def nnz(self) -> int:
        return int(self._data.shape[-1])

    @property
    def blocks(self) -> _BlockIndexer:
        return self._blocks

    def spy(self):
        return self.rows, self.cols

    # -- construction ------------------------------------------------------
    @classmethod
    def from_sparray(cls, sparray, block_sizes, global_stack_shape, symmetry=False,
                     symmetry_op: Callable = _conj, device=None, dtype=None) -> "DSDBCOO":
        """dsdbcoo.py:873-964 for one rank: canonicalise, block-sort, broadcast."""
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        coo = sps.coo_matrix(sparray)
        if symmetry:
            coo = sps.triu(coo, format="coo")
        if not coo.has_canonical_format:
            coo.sum_duplicates()
        order = block_sort_index(coo.row, coo.col, block_sizes)
        vals = torch.as_tensor(coo.data[order])
        if dtype is not None:
            vals = vals.to(dtype)
        data = vals.to(device or "cpu").reshape((1,) * len(global_stack_shape) + (-1,))
        data = data.expand(tuple(global_stack_shape) + (len(order),)).contiguous()
        return cls(data, coo.row[order], coo.col[order], block_sizes, global_stack_shape,
                   symmetry=symmetry, symmetry_op=symmetry_op)

    @classmethod
    def zeros_like(cls, other: "DSDBCOO") -> "DSDBCOO":
        """dsdbcoo.py:837-870: same pattern, zero data."""
        out = cls(torch.zeros_like(other.data), other.rows, other.cols, other.block_sizes,
                  other.global_stack_shape, other.return_dense, other.symmetry,
                  other.symmetry_op)
        out._slices = other._slices
        out._dense = other._dense
        return out

    def stack_view(self, start: int, stop: int) -> "DSDBCOO":
        """Container aliasing entries [start, stop) of the first stack axis (no copy)."""
        out = self.__class__(self._data[start:stop], self.rows, self.cols, self.block_sizes,
                             (stop - start,) + self.global_stack_shape[1:], self.return_dense,
                             self.symmetry, self.symmetry_op)
        out._slices, out._dense = self._slices, self._dense
        return out

    def to(self, device) -> "DSDBCOO":
        out = self.__class__(self._data.to(device), self.rows, self.cols, self.block_sizes,
                             self.global_stack_shape, self.return_dense, self.symmetry,
                             self.symmetry_op)
        out._slices, out._dense = self._slices, self._dense
        return out

    # -- block metadata ----------------------------------------------------
    def block_slice(self, i: int, j: int):
        """Contiguous nnz slice of block (i, j), or None when it stores nothing.

        Entries are block-sorted, so the slice is a bisection on the
        (block-row, block-col) key of every entry (computed once per pattern).
        """
        key = (i, j)
        if key not in self._slices:
            if "_keys" not in self._slices:
                o = self._block_offsets
                brow = np.searchsorted(o, self.rows, side="right") - 1
                bcol = np.searchsorted(o, self.cols, side="right") - 1
                keys = brow.astype(np.int64) * self.num_blocks + bcol
                if keys.size and not np.all(keys[1:] >= keys[:-1]):
                    raise ValueError("DSDBCOO entries must be sorted by (block-row, block-col)")
                self._slices["_keys"] = keys
            keys = self._slices["_keys"]
            k = i * self.num_blocks + j
            lo, hi = np.searchsorted(keys, k, "left"), np.searchsorted(keys, k, "right")
            self._slices[key] = None if hi == lo else slice(int(lo), int(hi))
        return self._slices[key]

    def is_dense_block(self, i: int, j: int) -> bool:
        """True when block (i, j) is stored as one full row-major slice."""
        key = (i, j)
        if key not in self._dense:
            sl = self.block_slice(i, j)
            ok = False
            if sl is not None:
                bi, bj = int(self._block_sizes[i]), int(self._block_sizes[j])
                if sl.stop - sl.start == bi * bj:
                    r = self.rows[sl] - self._block_offsets[i]
                    c = self.cols[sl] - self._block_offsets[j]
                    ok = bool(np.array_equal(r, np.repeat(np.arange(bi), bj))
                              and np.array_equal(c, np.tile(np.arange(bj), bi)))
            self._dense[key] = ok
        return self._dense[key]

    def _local_index(self, i, j):
        key = (i, j, self.device)
        if key not in self._dev_idx:
            sl = self.block_slice(i, j)
            r = torch.as_tensor(self.rows[sl] - self._block_offsets[i], dtype=torch.long)
            c = torch.as_tensor(self.cols[sl] - self._block_offsets[j], dtype=torch.long)
            self._dev_idx[key] = (r.to(self.device), c.to(self.device))
        return self._dev_idx[key]

    # -- block access (dsdbcoo.py:321-509) -------------------------------
    def _get_block(self, stack_index, i, j) -> torch.Tensor:
        if self.symmetry and j < i:
            blk = self._get_block(stack_index, j, i)
            return self.symmetry_op(blk.transpose(-1, -2)).contiguous()
        stack = self._data[stack_index]
        bi, bj = int(self._block_sizes[i]), int(self._block_sizes[j])
        sl = self.block_slice(i, j)
        if sl is None:
            return torch.zeros(stack.shape[:-1] + (bi, bj), dtype=self.dtype, device=self.device)
        if self.is_dense_block(i, j) and not (self.symmetry and i == j):
            return stack[..., sl].reshape(stack.shape[:-1] + (bi, bj)).clone()
        out = torch.zeros(stack.shape[:-1] + (bi, bj), dtype=self.dtype, device=self.device)
        r, c = self._local_index(i, j)
        out[..., r, c] = stack[..., sl]
        if self.symmetry and i == j:
            out = out + self.symmetry_op(out.transpose(-1, -2))
            d = torch.arange(bi, device=self.device)
            out[..., d, d] /= 2
        return out

    def _set_block(self, stack_index, i, j, block) -> None:
        block = torch.as_tensor(block).to(device=self.device, dtype=self.dtype)
        if self.symmetry and j < i:
            self._set_block(stack_index, j, i, self.symmetry_op(block.transpose(-1, -2)))
            return
        sl = self.block_slice(i, j)
        if sl is None:
            return
        stack = self._data[stack_index]
        if self.is_dense_block(i, j):
            stack[..., sl] = block.reshape(block.shape[:-2] + (-1,))
            return
        r, c = self._local_index(i, j)
        stack[..., sl] = block[..., r, c]

    def to_dense(self) -> torch.Tensor:
        """dsdbcoo.py:966-1008 for one rank."""
        arr = torch.zeros(self.shape, dtype=self.dtype, device=self.device)
        r = torch.as_tensor(self.rows, dtype=torch.long, device=self.device)
        c = torch.as_tensor(self.cols, dtype=torch.long, device=self.device)
        arr[..., r, c] = self._data
        if self.symmetry:
            arr = arr + self.symmetry_op(arr.transpose(-1, -2))
            d = torch.arange(self.shape[-1], device=self.device)
            arr[..., d, d] /= 2
        return arr


def btd_pattern(block_sizes, offdiag: bool = True):
    """Dense BTD sparsity pattern in the block-sorted order.

    Returns (rows, cols, offsets) where ``offsets[(i, j)]`` is the nnz offset
    of the row-major block (i, j).  ``offdiag=False`` gives a block-diagonal
    pattern.
    """
    bs = [int(b) for b in block_sizes]
    n = len(bs)
    off = np.hstack(([0], np.cumsum(bs)))
    rows, cols, offsets, pos = [], [], {}, 0
    for i in range(n):
        cand = [(i, i - 1), (i, i), (i, i + 1)] if offdiag else [(i, i)]
        for (bi_, bj_) in cand:
            if bj_ < 0 or bj_ >= n:
                continue
            bi, bj = bs[bi_], bs[bj_]
            rows.append(np.repeat(np.arange(bi), bj) + off[bi_])
            cols.append(np.tile(np.arange(bj), bi) + off[bj_])
            offsets[(bi_, bj_)] = pos
            pos += bi * bj
    return np.concatenate(rows).astype(np.int32), np.concatenate(cols).astype(np.int32), offsets


def empty_btd(block_sizes, stack_shape, offdiag: bool = True, device=None,
              dtype=torch.complex128, zero: bool = True) -> tuple[DSDBCOO, dict]:
    """Allocate a dense-block BTD (or block-diagonal) container; returns it and its block offsets."""
    rows, cols, offsets = btd_pattern(block_sizes, offdiag)
    stack_shape = (stack_shape,) if isinstance(stack_shape, int) else tuple(stack_shape)
    alloc = torch.zeros if zero else torch.empty
    data = alloc(stack_shape + (len(rows),), dtype=dtype, device=device)
    return DSDBCOO(data, rows, cols, block_sizes, stack_shape), offsets


def btd_from_blocks(diag, upper, lower, global_stack_shape=None, device=None) -> DSDBCOO:
    """DSDBCOO with fully dense BTD blocks built straight from block stacks.

    ``diag[i]`` has shape (*stack, b_i, b_i), ``upper[i]`` (*stack, b_i,
    b_{i+1}) and ``lower[i]`` (*stack, b_{i+1}, b_i); ``upper``/``lower`` may
    be None (no off-diagonal storage, e.g. a block-diagonal self-energy).
    The nnz order is the reference's block-sorted row-major order, i.e. what
    ``from_sparray`` produces for the same dense pattern.
    """
    n = len(diag)
    bs = [int(d.shape[-1]) for d in diag]
    stack = tuple(diag[0].shape[:-2]) if global_stack_shape is None else tuple(global_stack_shape)
    off = np.hstack(([0], np.cumsum(bs)))
    parts, rows, cols = [], [], []

    def add(blk, i, j):
        bi, bj = bs[i], bs[j]
        parts.append(blk.reshape(stack + (bi * bj,)))
        rows.append(np.repeat(np.arange(bi), bj) + off[i])
        cols.append(np.tile(np.arange(bj), bi) + off[j])

    for i in range(n):
        if i > 0 and lower is not None:
            add(lower[i - 1], i, i - 1)
        add(diag[i], i, i)
        if i + 1 < n and upper is not None:
            add(upper[i], i, i + 1)
    data = torch.cat(parts, dim=-1)
    if device is not None:
        data = data.to(device)
    return DSDBCOO(data, np.concatenate(rows), np.concatenate(cols), bs, stack)
