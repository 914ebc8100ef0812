"""CPU checks: the DSDBCOO container against the oracle container, the host
batching helpers, and the C ABI library (loads, exports every declared symbol)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from oracle.batching import get_batches as o_get_batches
from oracle.dsdbcoo import DSDBCOO as ODSDBCOO
from quatrex_amd.datastructures import DSDBCOO, btd_pattern, empty_btd
from quatrex_amd.utils import get_batches
from tests.conftest import bt_dense

ROOT = Path(__file__).resolve().parents[1]
BLOCK_SIZES = [np.array([2] * 10), np.array([2] * 3 + [4] * 2 + [2] * 3)]


@pytest.mark.parametrize("n,m", [(10, 1), (10, 2), (10, 3), (10, 100), (150, 100), (7, 7), (9, 4)])
def test_get_batches_matches_oracle(n, m):
    s1, o1 = get_batches(n, m)
    s2, o2 = o_get_batches(n, m)
    assert s1 == s2 and np.array_equal(o1, o2)


@pytest.mark.parametrize("block_sizes", BLOCK_SIZES, ids=["const", "mixed"])
@pytest.mark.parametrize("symmetry", [False, True])
@pytest.mark.parametrize("stack", [(3,), (2, 3)])
def test_container_matches_oracle(block_sizes, symmetry, stack):
    dense = bt_dense(block_sizes, seed=2)
    if symmetry:
        dense = dense + dense.conj().T
    coo = sps.coo_matrix(dense)
    o = ODSDBCOO.from_sparray(coo, block_sizes, stack, symmetry=symmetry)
    g = DSDBCOO.from_sparray(coo, block_sizes, stack, symmetry=symmetry)
    assert np.array_equal(g.rows, o.rows) and np.array_equal(g.cols, o.cols)
    assert np.array_equal(g.data.numpy(), o.data)
    assert np.allclose(g.to_dense().numpy(), o.to_dense())
    n = len(block_sizes)
    for i in range(n):
        for j in range(max(0, i - 1), min(n, i + 2)):
            assert np.allclose(g.blocks[i, j].numpy(), o.blocks[i, j])
            assert np.allclose(g.stack[1:].blocks[i, j].numpy(), o.stack[1:].blocks[i, j])
    # set then get round trip through a block (including symmetric lower set)
    rng = np.random.default_rng(0)
    b = rng.standard_normal((*stack, block_sizes[1], block_sizes[0])) + 0j
    o.stack[:].blocks[1, 0] = b
    g.stack[:].blocks[1, 0] = torch.as_tensor(b)
    assert np.allclose(g.data.numpy(), o.data)
    if not symmetry:
        assert np.allclose(dense, o.to_dense()[(0,) * len(stack)] * 0 + dense)


def test_dense_pattern_equals_from_sparray():
    bs = [3, 5, 2, 4]
    rows, cols, offsets = btd_pattern(bs)
    off = np.hstack(([0], np.cumsum(bs)))
    mask = np.zeros((off[-1], off[-1]))
    for i in range(len(bs)):
        for j in (i - 1, i, i + 1):
            if 0 <= j < len(bs):
                mask[off[i]:off[i + 1], off[j]:off[j + 1]] = 1
    o = ODSDBCOO.from_sparray(sps.coo_matrix(mask), bs, (1,))
    assert np.array_equal(rows, o.rows) and np.array_equal(cols, o.cols)
    g, offs = empty_btd(bs, (2,))
    for (i, j), start in offs.items():
        assert g.is_dense_block(i, j) and g.block_slice(i, j).start == start
    assert offs == {k: o.block_slice(*k).start for k in offs}


def test_stack_view_aliases():
    g, _ = empty_btd([2, 3], (5,))
    v = g.stack_view(1, 3)
    v.data[...] = 1.0
    assert torch.all(g.data[1:3] == 1) and torch.all(g.data[0] == 0) and torch.all(g.data[3:] == 0)


def _declared_symbols():
    text = (ROOT / "include" / "quatrex_amd.h").read_text()
    return sorted(set(re.findall(r"\b(qx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_declared_symbols():
    from quatrex_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail("libquatrex_amd.so is not built (run python -m quatrex_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTED_SYMBOLS)
    lib.qx_abi_version.restype = ctypes.c_int
    assert lib.qx_abi_version() == _lib.ABI_VERSION == 3


def test_solver_refuses_host_tensors():
    from quatrex_amd.greens_function_solver import RGF

    dense = bt_dense(BLOCK_SIZES[0])
    a = DSDBCOO.from_sparray(sps.coo_matrix(dense), BLOCK_SIZES[0], (2,))
    with pytest.raises((ValueError, ImportError)):
        RGF().selected_inv(a)
