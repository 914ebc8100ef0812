"""Dense ``Inv`` solver on the HIP path against the NumPy oracle.

Same inputs for both; patterns of the reference's tests/qttools/
greens_function_solver/test_selected_inv.py and test_selected_solve.py
(which run every solver, ``Inv`` included, against the dense inverse).
"""

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from oracle.dsdbcoo import DSDBCOO as ODSDBCOO
from oracle.inv import Inv as OInv
from oracle.rgf import OBCBlocks as OOBCBlocks
from quatrex_amd.datastructures import DSDBCOO
from quatrex_amd.greens_function_solver import Inv, OBCBlocks
from tests.conftest import bt_dense, rel_err

pytestmark = pytest.mark.gpu


def to_gpu(o, dev):
    return DSDBCOO(torch.as_tensor(o.data, device=dev), o.rows, o.cols, o.block_sizes,
                   o.global_stack_shape, symmetry=o.symmetry)


def _system(block_sizes, stack, seed=0):
    rng = np.random.default_rng(seed)
    dense = bt_dense(block_sizes)
    coo = sps.coo_matrix(dense)
    a = ODSDBCOO.from_sparray(coo, block_sizes, stack)
    a.data[:] = a.data * (1 + 0.1 * rng.standard_normal(a.data.shape))
    s = coo - coo.conj().T
    sl = ODSDBCOO.from_sparray(sps.coo_matrix(s), block_sizes, stack)
    sg = ODSDBCOO.from_sparray(sps.coo_matrix(2 * s), block_sizes, stack)
    return a, sl, sg


@pytest.mark.parametrize("block_sizes", [np.array([3] * 5), np.array([2, 4, 3, 2])],
                         ids=["const", "mixed"])
@pytest.mark.parametrize("stack", [(5,), (4, 2)], ids=["1D", "2D"])
@pytest.mark.parametrize("max_batch_size", [1, 100])
def test_selected_inv(device, block_sizes, stack, max_batch_size):
    a, _, _ = _system(block_sizes, stack)
    rng = np.random.default_rng(4)
    b0 = block_sizes[0]
    obc_ref = OOBCBlocks(len(block_sizes))
    obc_ref.retarded[0] = 0.3 * (rng.standard_normal(stack + (b0, b0)) + 1j)
    obc = OBCBlocks(len(block_sizes))
    obc.retarded[0] = torch.as_tensor(obc_ref.retarded[0], device=device)
    ref = OInv(max_batch_size).selected_inv(a, obc_blocks=obc_ref)
    got = Inv(max_batch_size).selected_inv(to_gpu(a, device), obc_blocks=obc)
    assert rel_err(got.data.cpu().numpy(), ref.data) < 1e-11
    out = DSDBCOO.zeros_like(to_gpu(a, device))
    assert Inv(max_batch_size).selected_inv(to_gpu(a, device), out=out) is None


@pytest.mark.parametrize("return_retarded", [True, False])
@pytest.mark.parametrize("use_out", [True, False])
def test_selected_solve(device, return_retarded, use_out):
    bs = np.array([2, 4, 3, 2])
    a, sl, sg = _system(bs, (6,))
    n, bn = len(bs), bs[-1]
    lo = 0.1j * np.random.default_rng(2).standard_normal((6, bn, bn))
    oobc, obc = OOBCBlocks(n), OBCBlocks(n)
    oobc.retarded[n - 1] = 0.2j * np.broadcast_to(np.eye(bn), (6, bn, bn)).copy()
    oobc.lesser[n - 1] = lo
    oobc.greater[n - 1] = 2 * lo
    for name in ("retarded", "lesser", "greater"):
        getattr(obc, name)[n - 1] = torch.as_tensor(getattr(oobc, name)[n - 1], device=device)
    ref = OInv(4).selected_solve(a, sl, sg, obc_blocks=oobc, return_retarded=return_retarded)
    ga, gsl, gsg = to_gpu(a, device), to_gpu(sl, device), to_gpu(sg, device)
    if use_out:
        outs = tuple(DSDBCOO.zeros_like(ga) for _ in range(3 if return_retarded else 2))
        got = Inv(4).selected_solve(ga, gsl, gsg, obc_blocks=obc, out=outs,
                                    return_retarded=return_retarded)
        assert all(g is o for g, o in zip(got, outs))
    else:
        got = Inv(4).selected_solve(ga, gsl, gsg, obc_blocks=obc, return_retarded=return_retarded)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert rel_err(g.data.cpu().numpy(), r.data) < 1e-11


def test_errors(device):
    a, sl, sg = _system(np.array([3] * 4), (2,))
    ga, gsl, gsg = to_gpu(a, device), to_gpu(sl, device), to_gpu(sg, device)
    with pytest.raises(NotImplementedError):
        Inv().selected_solve(ga, gsl, gsg, return_current=True)
    with pytest.raises(ValueError, match="Missing output for the retarded"):
        Inv().selected_solve(ga, gsl, gsg, out=(DSDBCOO.zeros_like(ga), DSDBCOO.zeros_like(ga)),
                             return_retarded=True)
    big = ODSDBCOO.from_sparray(sps.coo_matrix(bt_dense(np.array([130] * 4))),
                                np.array([130] * 4), (1,))
    with pytest.raises(ValueError, match="exceeds"):
        Inv().selected_inv(to_gpu(big, device))
