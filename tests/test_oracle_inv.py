"""Pins the NumPy dense ``Inv`` oracle (greens_function_solver/inv.py).

The dense solver is checked against ``numpy.linalg.inv`` on the dense
matrix (the pattern of the reference's tests/qttools/greens_function_solver/
test_selected_inv.py:37-70) and against the RGF oracle, which is pinned
independently by brute force; the reference's sign quirk on the lesser /
greater boundary blocks (subtracted, inv.py:199-202) is pinned by the
equivalence Inv(+S) == RGF(-S); the error paths follow inv.py:148-151 and
:172-176.
"""

import numpy as np
import pytest
import scipy.sparse as sps

from oracle.dsdbcoo import DSDBCOO
from oracle.inv import Inv
from oracle.rgf import RGF, OBCBlocks
from tests.conftest import bt_dense, rel_err

BLOCK_SIZES = [np.array([3] * 5), np.array([2, 4, 3, 2])]


def _system(block_sizes, stack, seed=0):
    rng = np.random.default_rng(seed)
    dense = bt_dense(block_sizes)
    coo = sps.coo_matrix(dense)
    a = DSDBCOO.from_sparray(coo, block_sizes, stack)
    a.data[:] = a.data * (1 + 0.1 * rng.standard_normal(a.data.shape))
    s = coo - coo.conj().T
    sl = DSDBCOO.from_sparray(sps.coo_matrix(s), block_sizes, stack)
    sg = DSDBCOO.from_sparray(sps.coo_matrix(2 * s), block_sizes, stack)
    return a, sl, sg


@pytest.mark.parametrize("block_sizes", BLOCK_SIZES, ids=["const", "mixed"])
@pytest.mark.parametrize("max_batch_size", [1, 3, 100])
def test_selected_inv_matches_dense_inverse(block_sizes, max_batch_size):
    a, _, _ = _system(block_sizes, (5,))
    ref = np.linalg.inv(a.to_dense())
    x = Inv(max_batch_size).selected_inv(a)
    mask = a.to_dense()[0] != 0
    assert np.allclose(x.to_dense() * mask, ref * mask, rtol=1e-12, atol=1e-14)
    out = DSDBCOO.zeros_like(a)
    assert Inv(max_batch_size).selected_inv(a, out=out) is None
    assert np.array_equal(out.data, x.data)


@pytest.mark.parametrize("block_sizes", BLOCK_SIZES, ids=["const", "mixed"])
def test_selected_inv_matches_rgf_with_obc(block_sizes):
    a, _, _ = _system(block_sizes, (4, 2))
    n, b0 = len(block_sizes), block_sizes[0]
    obc = OBCBlocks(n)
    rng = np.random.default_rng(5)
    obc.retarded[0] = 0.3 * (rng.standard_normal((4, 2, b0, b0)) + 1j)
    x_inv = Inv().selected_inv(a, obc_blocks=obc)
    x_rgf = RGF().selected_inv(a, obc_blocks=obc)
    assert rel_err(x_inv.data, x_rgf.data) < 1e-11


@pytest.mark.parametrize("block_sizes", BLOCK_SIZES, ids=["const", "mixed"])
@pytest.mark.parametrize("return_retarded", [True, False])
def test_selected_solve_matches_rgf(block_sizes, return_retarded):
    a, sl, sg = _system(block_sizes, (6,))
    n, bn = len(block_sizes), block_sizes[-1]
    rng = np.random.default_rng(2)
    obc = OBCBlocks(n)
    obc.retarded[n - 1] = 0.2j * np.broadcast_to(np.eye(bn), (6, bn, bn))
    lo = 0.1j * (rng.standard_normal((6, bn, bn)))
    lo = lo - np.conj(np.swapaxes(lo, -1, -2))
    obc_inv, obc_rgf = OBCBlocks(n), OBCBlocks(n)
    for o in (obc_inv, obc_rgf):
        o.retarded[n - 1] = obc.retarded[n - 1]
    # inv.py subtracts the lesser/greater boundary blocks, RGF adds them.
    obc_inv.lesser[n - 1], obc_rgf.lesser[n - 1] = lo, -lo
    obc_inv.greater[n - 1], obc_rgf.greater[n - 1] = 2 * lo, -2 * lo
    got = Inv(4).selected_solve(a, sl, sg, obc_blocks=obc_inv, return_retarded=return_retarded)
    ref = RGF(4).selected_solve(a, sl, sg, obc_blocks=obc_rgf, return_retarded=return_retarded)
    assert len(got) == (3 if return_retarded else 2)
    # RGF writes only the diagonal G^R blocks (rgf.py:304-305, 402-404)
    if return_retarded:
        for i in range(n):
            assert rel_err(got[2].blocks[i, i], ref[2].blocks[i, i]) < 1e-11
    for g, r in zip(got[:2], ref[:2]):
        assert rel_err(g.data, r.data) < 1e-11


def test_selected_solve_out_and_errors():
    a, sl, sg = _system(BLOCK_SIZES[0], (3,))
    xl, xg, xr = (DSDBCOO.zeros_like(a) for _ in range(3))
    res = Inv().selected_solve(a, sl, sg, out=(xl, xg, xr), return_retarded=True)
    assert res[0] is xl and res[1] is xg and res[2] is xr
    x = np.linalg.inv(a.to_dense())
    ref_l = x @ sl.to_dense() @ np.conj(np.swapaxes(x, -1, -2))
    mask = a.to_dense()[0] != 0
    assert np.allclose(xl.to_dense() * mask, ref_l * mask, rtol=1e-12, atol=1e-13)
    with pytest.raises(ValueError, match="Missing output for the retarded"):
        Inv().selected_solve(a, sl, sg, out=(xl, xg), return_retarded=True)
    with pytest.raises(NotImplementedError):
        Inv().selected_solve(a, sl, sg, return_current=True)
