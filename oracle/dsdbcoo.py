"""Single-rank NumPy restatement of the DSDBCOO block-sorted COO stack.

Follows ``src/qttools/datastructures/dsdbcoo.py`` (``from_sparray`` :873-964,
``_get_block`` :321-400, ``_set_block`` :454-509, ``zeros_like`` :837-870,
``to_dense`` :966-1008) and the block-sort kernel
``src/qttools/kernels/datastructure/numba/dsdbcoo.py:172-236``.

Layout facts this restates (and that the HIP path relies on):

- ``data`` has shape ``(*stack_shape, nnz)``; the last axis is contiguous.
- Nonzeros are ordered block-row major, then block-column, and inside a block
  in the canonical (row, col) order of a scipy COO with ``sum_duplicates``,
  i.e. row-major.  A fully populated block (i, j) is therefore one contiguous
  row-major ``b_i x b_j`` slice of ``data``.
- With ``symmetry=True`` only the upper triangle (``triu``) is stored, lower
  blocks are read as ``symmetry_op(block.T)`` and a diagonal block is read as
  ``B + op(B.T)`` with the diagonal halved.
"""

import numpy as np
import scipy.sparse as sps


def block_sort_index(rows: np.ndarray, cols: np.ndarray, block_sizes) -> np.ndarray:
    """Permutation that orders COO entries by (block-row, block-col).

    Same ordering as ``compute_block_sort_index``
    (``kernels/datastructure/numba/dsdbcoo.py:172-236``): a stable grouping,
    entries keep their relative order inside a block.
    """
    offsets = np.hstack(([0], np.cumsum(block_sizes)))
    brow = np.searchsorted(offsets, rows, side="right") - 1
    bcol = np.searchsorted(offsets, cols, side="right") - 1
    key = brow.astype(np.int64) * len(block_sizes) + bcol
    return np.argsort(key, kind="stable").astype(np.int64)


class _Blocks:
    def __init__(self, mat: "DSDBCOO", stack_index):
        self._m = mat
        self._idx = stack_index

    def __getitem__(self, ij):
        return self._m._get_block(self._idx, *ij)

    def __setitem__(self, ij, block):
        self._m._set_block(self._idx, *ij, block)


class _StackView:
    def __init__(self, mat: "DSDBCOO", index):
        if not isinstance(index, tuple):
            index = (index,)
        self._m = mat
        self.symmetry = mat.symmetry
        self.blocks = _Blocks(mat, index)


class _StackIndexer:
    def __init__(self, mat):
        self._m = mat

    def __getitem__(self, index):
        return _StackView(self._m, index)


class DSDBCOO:
    """Single-rank block-sorted COO stack (see module docstring)."""

    def __init__(self, data, rows, cols, block_sizes, global_stack_shape,
                 symmetry=False, symmetry_op=np.conj):
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        self.data = np.ascontiguousarray(data)
        self.rows = np.asarray(rows, dtype=np.int32)
        self.cols = np.asarray(cols, dtype=np.int32)
        self.block_sizes = np.asarray(block_sizes, dtype=np.int64)
        self.block_offsets = np.hstack(([0], np.cumsum(self.block_sizes)))
        self.global_stack_shape = tuple(global_stack_shape)
        self.symmetry = symmetry
        self.symmetry_op = symmetry_op
        self.num_blocks = len(self.block_sizes)
        n = int(self.block_offsets[-1])
        self.shape = tuple(self.data.shape[:-1]) + (n, n)
        self.dtype = self.data.dtype
        self._slices = {}
        self.stack = _StackIndexer(self)
        self.blocks = _Blocks(self, (Ellipsis,))

    # -- construction -------------------------------------------------
    @classmethod
    def from_sparray(cls, sparray, block_sizes, global_stack_shape,
                     symmetry=False, symmetry_op=np.conj):
        """dsdbcoo.py:873-964 for one rank: canonicalise, block-sort, broadcast."""
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        coo = sps.coo_matrix(sparray)
        if symmetry:
            coo = sps.triu(coo, format="coo")
        if not coo.has_canonical_format:
            coo.sum_duplicates()
        order = block_sort_index(coo.row, coo.col, block_sizes)
        vals = coo.data[order]
        data = np.zeros(tuple(global_stack_shape) + (len(vals),), dtype=coo.data.dtype)
        data[..., :] = vals
        return cls(data, coo.row[order], coo.col[order], block_sizes,
                   global_stack_shape, symmetry, symmetry_op)

    @classmethod
    def zeros_like(cls, other: "DSDBCOO") -> "DSDBCOO":
        """dsdbcoo.py:837-870 (the block-slice cache is shared: same pattern)."""
        out = cls(np.zeros_like(other.data), other.rows, other.cols,
                  other.block_sizes, other.global_stack_shape, other.symmetry,
                  other.symmetry_op)
        out._slices = other._slices
        return out

    # -- block access --------------------------------------------------
    def _block_keys(self):
        if "_keys" not in self._slices:
            o = self.block_offsets
            brow = np.searchsorted(o, self.rows, side="right") - 1
            bcol = np.searchsorted(o, self.cols, side="right") - 1
            keys = brow.astype(np.int64) * self.num_blocks + bcol
            self._slices["_keys"] = keys if np.all(keys[1:] >= keys[:-1]) else None
        return self._slices["_keys"]

    def block_slice(self, i: int, j: int):
        """Contiguous nnz slice of block (i, j) or ``None`` when empty (dsdbcoo.py:285-319).

        Entries are block-sorted, so the slice is found by bisection on the
        (block-row, block-col) keys; unsorted input falls back to a mask.
        """
        key = (i, j)
        if key not in self._slices:
            keys = self._block_keys()
            if keys is not None:
                k = i * self.num_blocks + j
                lo, hi = np.searchsorted(keys, k, "left"), np.searchsorted(keys, k, "right")
                self._slices[key] = None if hi == lo else slice(int(lo), int(hi))
            else:
                o = self.block_offsets
                m = ((self.rows >= o[i]) & (self.rows < o[i + 1])
                     & (self.cols >= o[j]) & (self.cols < o[j + 1]))
                idx = np.nonzero(m)[0]
                self._slices[key] = None if idx.size == 0 else slice(int(idx[0]), int(idx[-1]) + 1)
        return self._slices[key]

    def _get_block(self, stack_index, i, j):
        """Dense copy of block (i, j) (dsdbcoo.py:321-400)."""
        if self.symmetry and j < i:
            blk = self._get_block(stack_index, j, i)
            return np.ascontiguousarray(self.symmetry_op(np.swapaxes(blk, -1, -2)))
        stack = self.data[stack_index]
        bi, bj = int(self.block_sizes[i]), int(self.block_sizes[j])
        out = np.zeros(stack.shape[:-1] + (bi, bj), dtype=self.dtype)
        sl = self.block_slice(i, j)
        if sl is None:
            return out
        r = self.rows[sl] - self.block_offsets[i]
        c = self.cols[sl] - self.block_offsets[j]
        out[..., r, c] = stack[..., sl]
        if self.symmetry and i == j:
            out += self.symmetry_op(np.swapaxes(out, -1, -2))
            d = np.arange(bi)
            out[..., d, d] /= 2
        return out

    def _set_block(self, stack_index, i, j, block):
        """Scatter a dense block into its stored entries (dsdbcoo.py:454-509)."""
        if self.symmetry and j < i:
            self._set_block(stack_index, j, i, self.symmetry_op(np.swapaxes(block, -1, -2)))
            return
        sl = self.block_slice(i, j)
        if sl is None:
            return
        r = self.rows[sl] - self.block_offsets[i]
        c = self.cols[sl] - self.block_offsets[j]
        stack = self.data[stack_index]
        stack[..., sl] = block[..., r, c]

    def to_dense(self) -> np.ndarray:
        """dsdbcoo.py:966-1008 for one rank."""
        arr = np.zeros(self.shape, dtype=self.dtype)
        arr[..., self.rows, self.cols] = self.data
        if self.symmetry:
            arr += self.symmetry_op(np.swapaxes(arr, -1, -2))
            d = np.arange(self.shape[-1])
            arr[..., d, d] /= 2
        return arr
