"""Dense selected inversion / selected solve (oracle restatement).

Follows ``src/qttools/greens_function_solver/inv.py``: ``selected_inv``
(:32-97) and ``selected_solve`` (:99-227).  The matrix is densified, the
boundary blocks are folded in, the dense inverse (and the congruence
X^<> = X Sigma^<> X^H) is formed, and the entries of the output pattern
are gathered.

Quirks kept from the reference: ``selected_solve`` subtracts the lesser /
greater boundary blocks (inv.py:199-202), where RGF adds them; it always
returns the outputs, also when ``out`` is given (inv.py:224-227); a
missing retarded output raises ``ValueError("Missing output for the
retarded Green's function.")`` (inv.py:172-176); ``return_current`` raises
``NotImplementedError`` (inv.py:148-151).  The boundary blocks are folded
on the last two axes (``[..., b, b]``), so stacks of more than one
dimension work; the reference's ``[:, b, b]`` assumes a 1-D stack.

This module is test infrastructure (the CPU checker); the product path
never imports it.
"""

import numpy as np

from oracle.batching import get_batches
from oracle.rgf import OBCBlocks


def _h(x):
    return np.conj(np.swapaxes(x, -2, -1))


def _offsets(a):
    return np.concatenate([[0], np.cumsum(a.block_sizes)])


class Inv:
    """NumPy dense solver (inv.py:10-30)."""

    def __init__(self, max_batch_size: int = 100):
        self.max_batch_size = max_batch_size

    def selected_inv(self, a, obc_blocks=None, out=None):
        """inv.py:32-97."""
        sizes, offs = get_batches(a.shape[0], self.max_batch_size)
        if obc_blocks is None:
            obc_blocks = OBCBlocks(a.num_blocks)
        return_out = out is None
        if out is None:
            out = a.__class__.zeros_like(a)
        rows, cols = out.rows, out.cols
        bo = _offsets(a)
        dense = a.to_dense()
        for b in range(len(sizes)):
            sl = slice(int(offs[b]), int(offs[b + 1]))
            ad = dense[sl].copy()
            for j, blk in enumerate(obc_blocks.retarded):
                if blk is not None:
                    r = slice(bo[j], bo[j + 1])
                    ad[..., r, r] -= blk[sl]
            out.data[sl] = np.linalg.inv(ad)[..., rows, cols]
        if return_out:
            return out
        return None

    def selected_solve(self, a, sigma_lesser, sigma_greater, obc_blocks=None, out=None,
                       return_retarded=False, return_current=False):
        """inv.py:99-227."""
        if return_current:
            raise NotImplementedError("The computation of the current is not implemented.")
        sizes, offs = get_batches(a.shape[0], self.max_batch_size)
        if obc_blocks is None:
            obc_blocks = OBCBlocks(a.num_blocks)
        if out is None:
            xl, xg = a.__class__.zeros_like(a), a.__class__.zeros_like(a)
            if return_retarded:
                xr = a.__class__.zeros_like(a)
        else:
            xl, xg, *xr = out
            if return_retarded:
                if len(xr) == 0:
                    raise ValueError("Missing output for the retarded Green's function.")
                xr = xr[0]
        bo = _offsets(a)
        ad_all, sld_all, sgd_all = a.to_dense(), sigma_lesser.to_dense(), sigma_greater.to_dense()
        for b in range(len(sizes)):
            sl = slice(int(offs[b]), int(offs[b + 1]))
            ad, sld, sgd = ad_all[sl].copy(), sld_all[sl].copy(), sgd_all[sl].copy()
            for j, (br, bl, bg) in enumerate(zip(obc_blocks.retarded, obc_blocks.lesser,
                                                 obc_blocks.greater)):
                r = slice(bo[j], bo[j + 1])
                if br is not None:
                    ad[..., r, r] -= br[sl]
                if bl is not None:
                    sld[..., r, r] -= bl[sl]
                if bg is not None:
                    sgd[..., r, r] -= bg[sl]
            x = np.linalg.inv(ad)
            xh = _h(x)
            xl.data[sl] = (x @ sld @ xh)[..., xl.rows, xl.cols]
            xg.data[sl] = (x @ sgd @ xh)[..., xg.rows, xg.cols]
            if return_retarded:
                xr.data[sl] = x[..., xr.rows, xr.cols]
        if return_retarded:
            return xl, xg, xr
        return xl, xg
