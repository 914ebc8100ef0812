"""Observables and system-matrix assembly (oracle restatement).

- ``density`` follows quatrex/core/observables.py:57-126 (imaginary part of
  the diagonal of G, or of the G S products with an overlap S).
- ``contact_currents`` follows observables.py:129-190:
  I_c = Tr[Sigma^>_c G^<_cc - G^>_cc Sigma^<_c] at both contacts.
- ``assemble_system_matrix`` follows quatrex/electron/solver.py:390-407
  with ``_btd_subtract`` (:26-50): A = (E + i eta) (I or S) - diag(V)
  - Sigma^R (BTD blocks) - H (BTD blocks), on the pattern of A.

This module is test infrastructure (the CPU checker); the product path
never imports it.
"""

import numpy as np
import scipy.sparse as sps


def density(x, overlap=None):
    """Im diag(G) (overlap None) or Im diag of the G S block products."""
    n = x.num_blocks
    if overlap is None:
        return np.concatenate([np.diagonal(x.blocks[i, i], axis1=-2, axis2=-1).imag
                               for i in range(n)], axis=-1)
    off = np.concatenate([[0], np.cumsum(x.block_sizes)])
    s = sps.csr_matrix(overlap).toarray()
    blk = lambda i, j: s[off[i]:off[i + 1], off[j]:off[j + 1]]
    out = []
    for i in range(n):
        d = np.diagonal(x.blocks[i, i] @ blk(i, i), axis1=-2, axis2=-1).copy()
        if i < n - 1:
            d += np.diagonal(x.blocks[i, i + 1] @ blk(i + 1, i), axis1=-2, axis2=-1)
        if i > 0:
            d += np.diagonal(x.blocks[i, i - 1] @ blk(i - 1, i), axis1=-2, axis2=-1)
        out.append(d.imag)
    return np.concatenate(out, axis=-1)


def contact_currents(x_lesser, x_greater, obc):
    """(I_left, I_right), complex traces (observables.py:129-190)."""
    n = x_lesser.num_blocks - 1
    tr = lambda m: np.trace(m, axis1=-2, axis2=-1)
    i_l = tr(obc.greater[0] @ x_lesser.blocks[0, 0] - x_greater.blocks[0, 0] @ obc.lesser[0])
    i_r = tr(obc.greater[-1] @ x_lesser.blocks[n, n] - x_greater.blocks[n, n] @ obc.lesser[-1])
    return i_l, i_r


def _btd_subtract(a, b):
    """electron/solver.py:26-50 (single rank)."""
    for i in range(a.num_blocks):
        a.blocks[i, i] = a.blocks[i, i] - b.blocks[i, i]
        j = i + 1
        if j >= a.num_blocks:
            continue
        a.blocks[i, j] = a.blocks[i, j] - b.blocks[i, j]
        a.blocks[j, i] = a.blocks[j, i] - b.blocks[j, i]


def assemble_system_matrix(a, energies, eta, hamiltonian, sse_retarded, overlap=None,
                           potential=None):
    """Fill ``a`` (DSDBCOO, stack = energies) in place (electron/solver.py:390-407)."""
    a.data[:] = 0.0
    diag = a.rows == a.cols
    if overlap is None:
        a.data[..., diag] = 1.0
    else:
        s = sps.csr_matrix(overlap)
        a.data[:] += np.asarray(s[a.rows, a.cols]).ravel()
    z = np.asarray(energies) + 1j * eta
    a.data[:] *= z.reshape(z.shape + (1,) * (a.data.ndim - z.ndim))
    if potential is not None:
        a.data[..., diag] -= np.asarray(potential)[a.rows[diag]]
    _btd_subtract(a, sse_retarded)
    _btd_subtract(a, hamiltonian)
    return a
