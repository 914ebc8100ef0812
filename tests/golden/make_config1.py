"""Writes tests/golden/config1_rgf.npz: the NumPy oracle's selected solve of
the CPU-runnable configuration (16 energies, 8 blocks of 32, seed 0).

The oracle is pinned independently (dense brute force, identities: see
tests/test_oracle_rgf.py); these vectors freeze its outputs so that the
HIP path and later oracle changes are compared against fixed numbers.

Stored: energies; density Im diag G^< (16 x 256); transmission T(E) from
the corner block G^R_{0,7}; Meir-Wingreen layer currents (16 x 7); the
Frobenius norm of every stored block of G^<, G^>, G^R (16 x 22 each); the
diagonal G^< blocks and the G^R corner block of energies 0 and 15.

Run from the repository root: python -m tests.golden.make_config1
"""

from pathlib import Path

import numpy as np

import oracle.rgf as orgf
from oracle.dsdbcoo import DSDBCOO
from quatrex_amd.datastructures import btd_pattern
from quatrex_amd.workloads import make_workload

OUT = Path(__file__).resolve().parent / "config1_rgf.npz"


def config1_containers(w):
    bs = w.block_sizes
    rows, cols, aoff = btd_pattern(bs, offdiag=True)
    drows, dcols, doff = btd_pattern(bs, offdiag=False)
    diag, upper, lower = w.a_blocks()
    ne = len(w.energies)
    a = np.zeros((ne, len(rows)), complex)
    sl = np.zeros((ne, len(drows)), complex)
    sg = np.zeros((ne, len(drows)), complex)
    for i, b in enumerate(bs):
        a[:, aoff[(i, i)]:aoff[(i, i)] + b * b] = diag[i].reshape(ne, -1)
        sl[:, doff[(i, i)]:doff[(i, i)] + b * b] = w.sl_diag[i].reshape(1, -1)
        sg[:, doff[(i, i)]:doff[(i, i)] + b * b] = w.sg_diag[i].reshape(1, -1)
        if i + 1 < len(bs):
            o = aoff[(i, i + 1)]
            a[:, o:o + b * bs[i + 1]] = upper[i].reshape(ne, -1)
            o = aoff[(i + 1, i)]
            a[:, o:o + b * bs[i + 1]] = lower[i].reshape(ne, -1)
    mk = lambda d, r, c: DSDBCOO(d, r, c, bs, (ne,))
    return mk(a, rows, cols), mk(sl, drows, dcols), mk(sg, drows, dcols)


def block_keys(n):
    return [(i, j) for i in range(n) for j in (i - 1, i, i + 1) if 0 <= j < n]


def compute():
    w = make_workload(seed=0)
    a, sl, sg = config1_containers(w)
    n, ne = w.n_blocks, len(w.energies)
    corner = np.zeros((ne, w.block_sizes[0], w.block_sizes[-1]), complex)
    xl, xg, xr, cur = orgf.RGF(16).selected_solve(a, sl, sg, return_retarded=True,
                                                  return_current=True, retarded_offdiag=True,
                                                  corner=corner)
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    keys = block_keys(n)
    norms = np.stack([np.stack([np.linalg.norm(x.blocks[i, j], axis=(-2, -1)) for i, j in keys], -1)
                      for x in (xl, xg, xr)])
    return {
        "energies": w.energies,
        "density": np.concatenate([np.diagonal(xl.blocks[i, i], axis1=-2, axis2=-1).imag
                                   for i in range(n)], -1),
        "transmission": orgf.transmission(corner, gl, gr),
        "current": cur,
        "block_norms": norms,
        "block_keys": np.array(keys),
        "lesser_diag_e0": np.stack([xl.blocks[i, i][0] for i in range(n)]),
        "lesser_diag_e15": np.stack([xl.blocks[i, i][15] for i in range(n)]),
        "corner_e0": corner[0],
        "corner_e15": corner[15],
    }


if __name__ == "__main__":
    np.savez_compressed(OUT, **compute())
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")
