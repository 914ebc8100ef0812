import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libquatrex_amd.so")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    return torch.device("cuda:0")


def random_block(rng, m, n):
    """Quasi-sparse complex block (50 % density), as the reference test fixture
    tests/qttools/greens_function_solver/conftest.py:58-62."""
    mask = rng.random((m, n)) < 0.5
    re = rng.random((m, n)) * mask
    im = rng.random((m, n)) * mask
    return re + 1j * im


def bt_dense(block_sizes, seed=0):
    """Random diagonally dominant BTD matrix (reference conftest.py:65-104)."""
    rng = np.random.default_rng(seed)
    off = np.hstack(([0], np.cumsum(block_sizes)))
    n = int(off[-1])
    arr = np.zeros((n, n), dtype=np.complex128)
    for i in range(len(block_sizes)):
        bi = int(block_sizes[i])
        arr[off[i]:off[i + 1], off[i]:off[i + 1]] = random_block(rng, bi, bi) + np.eye(bi)
        if i > 0:
            bj = int(block_sizes[i - 1])
            arr[off[i]:off[i + 1], off[i - 1]:off[i]] = random_block(rng, bi, bj)
            arr[off[i - 1]:off[i], off[i]:off[i + 1]] = random_block(rng, bj, bi)
    for i in range(n):
        arr[i, i] = (1 + 1j) + complex(np.sum(arr[i, :]))
    return arr


def rel_err(x, y):
    x = np.asarray(x)
    y = np.asarray(y)
    den = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (den if den > 0 else 1.0)


def oracle_btd(w, energies=None):
    """(A, Sigma^<, Sigma^>) of a workload as oracle DSDBCOO stacks built block
    by block (no dense matrix), for the energies given (default: all)."""
    from oracle.dsdbcoo import DSDBCOO as ODSDBCOO
    from quatrex_amd.datastructures import btd_pattern

    e = w.energies if energies is None else np.asarray(energies)
    bs = w.block_sizes
    rows, cols, aoff = btd_pattern(bs, offdiag=True)
    drows, dcols, doff = btd_pattern(bs, offdiag=False)
    diag, upper, lower = w.a_blocks(e)
    ne = len(e)
    a = np.empty((ne, len(rows)), complex)
    sl = np.empty((ne, len(drows)), complex)
    sg = np.empty((ne, len(drows)), complex)
    for i, b in enumerate(bs):
        o = aoff[(i, i)]
        a[:, o:o + b * b] = diag[i].reshape(ne, -1)
        o = doff[(i, i)]
        sl[:, o:o + b * b] = w.sl_diag[i].reshape(1, -1)
        sg[:, o:o + b * b] = w.sg_diag[i].reshape(1, -1)
        if i + 1 < len(bs):
            o = aoff[(i, i + 1)]
            a[:, o:o + b * bs[i + 1]] = upper[i].reshape(ne, -1)
            o = aoff[(i + 1, i)]
            a[:, o:o + b * bs[i + 1]] = lower[i].reshape(ne, -1)
    mk = lambda d, r, c: ODSDBCOO(d, r, c, bs, (ne,))  # noqa: E731
    return mk(a, rows, cols), mk(sl, drows, dcols), mk(sg, drows, dcols)
