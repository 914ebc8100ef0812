"""Lead-block preparation and contact self-energies (oracle restatement).

- ``periodic_superblocks`` follows ``quatrex/core/utils.py:157-212``
  (``get_periodic_superblocks``): the principal layer of the lead is built
  by repeating the outermost sub-layer of the device's first block layer.
- ``contact_obc_blocks`` follows ``quatrex/electron/solver.py:304-386``
  (``ElectronSolver._compute_obc``): left and right lead surface Green's
  functions from an OBC system, the right one through the reference's
  flip of both matrix axes, then Sigma^R_B = m_10 g m_01,
  Gamma = i (Sigma - Sigma^H), Sigma^<_B = i f Gamma, Sigma^>_B =
  i (f - 1) Gamma (``scale_stack``, qttools/utils/stack_utils.py:6-30).

This module is test infrastructure (the CPU checker); the product path
never imports it.
"""

import numpy as np

from oracle.obc import block_view
from oracle.rgf import OBCBlocks


def _h(x):
    return np.conj(np.swapaxes(x, -2, -1))


def periodic_superblocks(a_ii, a_ij, a_ji, block_sections: int):
    """(m_10, m_00, m_01) of the periodic lead layer (core/utils.py:157-212)."""
    s = block_sections
    view_ij = block_view(np.concatenate((a_ii, a_ij), -1), -2, s)
    view_ij = block_view(np.ascontiguousarray(view_ij[0]), -1, 2 * s)
    view_ji = block_view(np.concatenate((a_ii, a_ji), -2), -1, s)
    view_ji = block_view(np.ascontiguousarray(view_ji[0]), -2, 2 * s)
    layer = np.concatenate((view_ji[s::-1], view_ij[1:]), axis=0)
    sub = a_ii.shape[:-2] + (a_ii.shape[-1] // s,) * 2
    blocks = np.zeros((s, 3 * s) + sub, dtype=a_ii.dtype)
    for i in range(s):
        blocks[i] = np.roll(layer, i, axis=0)
    full = np.concatenate(np.concatenate(blocks, -2), -1)
    m = block_view(np.ascontiguousarray(full), -1, 3)
    return m[0].copy(), m[1].copy(), m[2].copy()


def _scale(gamma, f):
    f = np.asarray(f)
    return gamma * f.reshape(f.shape + (1,) * (gamma.ndim - f.ndim))


def contact_obc_blocks(a, obc, left_occupancies, right_occupancies, block_sections=1,
                       overlap=None, eta_obc=0.0):
    """OBCBlocks of both contacts (electron/solver.py:304-386).

    ``a`` is the system matrix (DSDBCOO, stack = energies), ``obc`` an OBC
    system ``obc((a_ii, a_ij, a_ji), contact) -> (x, ...)``; ``overlap``
    optionally maps the block keys (0, 0), (0, 1), (1, 0), (-1, -1),
    (-1, -2), (-2, -1) to (b, b) overlap blocks, scaled by i * eta_obc.
    """
    n_blocks = a.num_blocks
    out = OBCBlocks(n_blocks)
    flip = lambda x: np.flip(x, axis=(-2, -1))

    def s_blk(key, b):
        if overlap is None or key not in overlap:
            return 0.0
        return 1j * eta_obc * overlap[key]

    # left contact (:307-337)
    b0 = a.block_sizes[0]
    m_10, m_00, m_01 = periodic_superblocks(a.blocks[0, 0], a.blocks[0, 1], a.blocks[1, 0],
                                            block_sections)
    g_00 = obc((m_00 + s_blk((0, 0), b0), m_01 + s_blk((0, 1), b0), m_10 + s_blk((1, 0), b0)),
               "left")[0]
    sigma = m_10 @ g_00 @ m_01
    out.retarded[0] = sigma
    gamma = 1j * (sigma - _h(sigma))
    out.lesser[0] = 1j * _scale(gamma, left_occupancies)
    out.greater[0] = 1j * _scale(gamma, np.asarray(left_occupancies) - 1)

    # right contact (:338-386)
    n, m = n_blocks - 1, n_blocks - 2
    bn = a.block_sizes[n]
    m_mn, m_nn, m_nm = periodic_superblocks(flip(a.blocks[n, n]), flip(a.blocks[n, m]),
                                            flip(a.blocks[m, n]), block_sections)
    m_nn, m_nm, m_mn = flip(m_nn), flip(m_nm), flip(m_mn)
    g_nn = obc((flip(m_nn + s_blk((-1, -1), bn)), flip(m_nm + s_blk((-1, -2), bn)),
                flip(m_mn + s_blk((-2, -1), bn))), "right")[0]
    g_nn = flip(g_nn)
    sigma = m_mn @ g_nn @ m_nm
    out.retarded[n] = sigma
    gamma = 1j * (sigma - _h(sigma))
    out.lesser[n] = 1j * _scale(gamma, right_occupancies)
    out.greater[n] = 1j * _scale(gamma, np.asarray(right_occupancies) - 1)
    return out
