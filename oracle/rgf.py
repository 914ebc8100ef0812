"""RGF selected inversion and selected solve (oracle restatement).

Follows ``src/qttools/greens_function_solver/rgf.py`` op for op:
``selected_inv`` (:25-113) and ``selected_solve`` (:115-414), with the OBC
folding of ``solver.py:12-28`` and the batching of ``solvers_utils.py:7-30``.
Matrices are the single-rank ``oracle.dsdbcoo.DSDBCOO``; every block read is
a dense copy and every block write a scatter, as in the reference.

Extension beyond the reference (off by default): ``retarded_offdiag=True``
also writes the off-diagonal retarded blocks in ``selected_solve`` using the
selected-inversion formulas of ``rgf.py:105-107``.
"""

import numpy as np

from oracle.batching import get_batches


def _h(x):
    return np.conj(np.swapaxes(x, -2, -1))


class OBCBlocks:
    """Per-block boundary self-energies (solver.py:12-28)."""

    def __init__(self, num_blocks: int):
        self.retarded = [None] * num_blocks
        self.lesser = [None] * num_blocks
        self.greater = [None] * num_blocks


class RGF:
    """NumPy RGF (rgf.py:10-23)."""

    def __init__(self, max_batch_size: int = 100):
        self.max_batch_size = max_batch_size

    def selected_inv(self, a, obc_blocks=None, out=None):
        """rgf.py:25-113."""
        n = a.num_blocks
        if obc_blocks is None:
            obc_blocks = OBCBlocks(n)
        sizes, offs = get_batches(a.shape[0], self.max_batch_size)
        x = out if out is not None else a.__class__.zeros_like(a)
        for b in range(len(sizes)):
            sl = slice(int(offs[b]), int(offs[b + 1]))
            a_, x_ = a.stack[sl], x.stack[sl]
            xd = [None] * n
            obc = obc_blocks.retarded[0]
            a00 = a_.blocks[0, 0] if obc is None else a_.blocks[0, 0] - obc[sl]
            xd[0] = np.linalg.inv(a00)
            for i in range(n - 1):
                j = i + 1
                obc = obc_blocks.retarded[j]
                ajj = a_.blocks[j, j] if obc is None else a_.blocks[j, j] - obc[sl]
                xd[j] = np.linalg.inv(ajj - a_.blocks[j, i] @ xd[i] @ a_.blocks[i, j])
            x_.blocks[n - 1, n - 1] = xd[-1]
            for i in range(n - 2, -1, -1):
                j = i + 1
                xii, xjj, aij = xd[i], xd[j], a_.blocks[i, j]
                xji = -xjj @ a_.blocks[j, i] @ xii
                x_.blocks[j, i] = xji
                x_.blocks[i, j] = -xii @ aij @ xjj
                xd[i] = xii - xii @ aij @ xji
                x_.blocks[i, i] = xd[i]
        if out is None:
            return x
        return None

    def selected_solve(self, a, sigma_lesser, sigma_greater, obc_blocks=None,
                       out=None, return_retarded=False, return_current=False,
                       retarded_offdiag=False, corner=None):
        """rgf.py:115-414.

        Extension: ``corner`` (array (*stack, b_0, b_{N-1})) receives the
        retarded corner block G^R_{0,N-1} from the forward recursion
        g_{0,j} = -g_{0,j-1} A_{j-1,j} x^R_jj, g_{0,0} = x^R_00 (the
        left-connected first-row blocks; the last one is exact).
        """
        n = a.num_blocks
        if obc_blocks is None:
            obc_blocks = OBCBlocks(n)
        if return_current:
            current = np.zeros((*a.shape[:-2], n - 1), dtype=a.dtype)
        sizes, offs = get_batches(a.shape[0], self.max_batch_size)
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

        for b in range(len(sizes)):
            sl = slice(int(offs[b]), int(offs[b + 1]))
            a_ = a.stack[sl]
            sl_ = sigma_lesser.stack[sl]
            sg_ = sigma_greater.stack[sl]
            xl_, xg_ = xl.stack[sl], xg.stack[sl]
            if return_retarded:
                xr_ = xr.stack[sl]

            def folded(j):
                o_r, o_l, o_g = (obc_blocks.retarded[j], obc_blocks.lesser[j],
                                 obc_blocks.greater[j])
                ajj = a_.blocks[j, j] if o_r is None else a_.blocks[j, j] - o_r[sl]
                slj = sl_.blocks[j, j] if o_l is None else sl_.blocks[j, j] + o_l[sl]
                sgj = sg_.blocks[j, j] if o_g is None else sg_.blocks[j, j] + o_g[sl]
                return ajj, slj, sgj

            xrd, xld, xgd = [None] * n, [None] * n, [None] * n
            a00, sl00, sg00 = folded(0)
            xr0 = np.linalg.inv(a00)
            xr0h = _h(xr0)
            xrd[0], xld[0], xgd[0] = xr0, xr0 @ sl00 @ xr0h, xr0 @ sg00 @ xr0h
            g_c = xr0

            # Forward sweep (rgf.py:234-295).
            for i in range(n - 1):
                j = i + 1
                ajj, sljj, sgjj = folded(j)
                aji = a_.blocks[j, i]
                xrii = xrd[i]
                ajih = _h(aji)
                aji_xrii = aji @ xrii
                xrjj = np.linalg.inv(ajj - aji_xrii @ a_.blocks[i, j])
                xrjjh = _h(xrjj)
                xrd[j] = xrjj
                if corner is not None:
                    g_c = -(g_c @ a_.blocks[i, j]) @ xrjj
                t = aji_xrii @ sl_.blocks[i, j]
                xld[j] = xrjj @ (sljj + aji @ xld[i] @ ajih + _h(t) - t) @ xrjjh
                t = aji_xrii @ sg_.blocks[i, j]
                xgd[j] = xrjj @ (sgjj + aji @ xgd[i] @ ajih + _h(t) - t) @ xrjjh

            if corner is not None:
                corner[sl] = g_c
            xl_.blocks[n - 1, n - 1] = 0.5 * (xld[-1] - _h(xld[-1]))
            xg_.blocks[n - 1, n - 1] = 0.5 * (xgd[-1] - _h(xgd[-1]))
            if return_retarded:
                xr_.blocks[n - 1, n - 1] = xrd[-1]

            # Backward sweep (rgf.py:308-404).
            for i in range(n - 2, -1, -1):
                j = i + 1
                xrii, xrjj = xrd[i], xrd[j]
                aij, aji = a_.blocks[i, j], a_.blocks[j, i]
                xlii, xljj, xgii, xgjj = xld[i], xld[j], xgd[i], xgd[j]
                slij, sgij = sl_.blocks[i, j], sg_.blocks[i, j]

                xrjjh = _h(xrjj)
                xrii_aij = xrii @ aij
                aijh_xriih = _h(xrii_aij)
                xrjj_aji = xrjj @ aji
                ajih_xrjjh = _h(xrjj_aji)
                xrii_aij_xrjj = xrii_aij @ xrjj
                xrjjh_aijh_xriih = _h(xrii_aij_xrjj)
                xrii_aij_xrjj_aji = xrii_aij @ xrjj_aji

                for sig, xd, x_, xii, xjj in ((slij, xld, xl_, xlii, xljj),
                                              (sgij, xgd, xg_, xgii, xgjj)):
                    t1 = xrii_aij_xrjj_aji @ xii - xrii @ sig @ xrjjh_aijh_xriih
                    t1 = t1 - _h(t1)
                    t2 = xrii_aij @ xjj
                    xij = -t2 - xii @ ajih_xrjjh + xrii @ sig @ xrjjh
                    x_.blocks[i, j] = xij
                    if not x_.symmetry:
                        x_.blocks[j, i] = -_h(xij)
                    xd[i] = xii + t2 @ aijh_xriih + t1
                    x_.blocks[i, i] = 0.5 * (xd[i] - _h(xd[i]))

                if return_current:
                    aji_xrii = aji @ xrii
                    t = aji_xrii @ slij
                    sl_tilde = aji @ xlii @ _h(aji) + _h(t) - t
                    t = aji_xrii @ sgij
                    sg_tilde = aji @ xgii @ _h(aji) + _h(t) - t
                    current[sl, ..., i] = np.trace(
                        sg_tilde @ xld[j] - xgd[j] @ sl_tilde, axis1=-2, axis2=-1)

                xrd[i] = xrii + xrii_aij_xrjj_aji @ xrii
                if return_retarded:
                    xr_.blocks[i, i] = xrd[i]
                    if retarded_offdiag:
                        # rgf.py:105-107 with the full G^R_jj.
                        xr_.blocks[i, j] = -xrii_aij_xrjj
                        xr_.blocks[j, i] = -xrjj_aji @ xrii

        if out is None:
            if return_retarded:
                if return_current:
                    return xl, xg, xr, current
                return xl, xg, xr
            return xl, xg
        if return_current:
            return current
        return None


def transmission(g_1n, gamma_l, gamma_r):
    """T(E) = Tr[Gamma_L G^R_{0,N-1} Gamma_R (G^R_{0,N-1})^H] (Caroli/Landauer)."""
    return np.trace(gamma_l @ g_1n @ gamma_r @ _h(g_1n), axis1=-2, axis2=-1).real
