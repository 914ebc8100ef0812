"""Pins the lead preparation / contact self-energy oracle.

``periodic_superblocks`` (core/utils.py:157-212): the identity for one
section, and exact reproduction of a lead layer that is itself periodic at
the sub-layer level (the construction repeats the outermost sub-layer).
``contact_obc_blocks`` (electron/solver.py:304-386): the surface Green's
functions satisfy their fixed points, the right contact of a mirrored
device equals the flipped left contact of the original (pins the flips),
and Sigma^> - Sigma^< = Sigma^R - Sigma^A for any occupancies.
"""

import numpy as np
import pytest
import scipy.sparse as sps

from oracle.dsdbcoo import DSDBCOO
from oracle.leads import contact_obc_blocks, periodic_superblocks
from oracle.obc import OBCSystem, SanchoRubio
from tests.conftest import rel_err


def _cn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def _periodic_layer(rng, s, sub, ne):
    h0 = _cn(rng, (sub, sub))
    h0 = (h0 + h0.conj().T) / 2
    t = _cn(rng, (sub, sub))
    n = s * sub
    a_ii = np.zeros((n, n), complex)
    a_ij = np.zeros((n, n), complex)
    a_ji = np.zeros((n, n), complex)
    for k in range(s):
        a_ii[k * sub:(k + 1) * sub, k * sub:(k + 1) * sub] = h0
        if k + 1 < s:
            a_ii[k * sub:(k + 1) * sub, (k + 1) * sub:(k + 2) * sub] = t
            a_ii[(k + 1) * sub:(k + 2) * sub, k * sub:(k + 1) * sub] = t.conj().T
    a_ij[(s - 1) * sub:, :sub] = t
    a_ji[:sub, (s - 1) * sub:] = t.conj().T
    st = lambda x: np.broadcast_to(x, (ne, n, n)).copy()
    return st(a_ii), st(a_ij), st(a_ji)


def test_superblocks_one_section_is_identity():
    rng = np.random.default_rng(0)
    a_ii, a_ij, a_ji = (_cn(rng, (3, 6, 6)) for _ in range(3))
    m_10, m_00, m_01 = periodic_superblocks(a_ii, a_ij, a_ji, 1)
    assert np.array_equal(m_10, a_ji) and np.array_equal(m_00, a_ii)
    assert np.array_equal(m_01, a_ij)


@pytest.mark.parametrize("s", [2, 3, 4])
def test_superblocks_reproduce_periodic_layer(s):
    rng = np.random.default_rng(s)
    a_ii, a_ij, a_ji = _periodic_layer(rng, s, 3, 2)
    m_10, m_00, m_01 = periodic_superblocks(a_ii, a_ij, a_ji, s)
    assert np.allclose(m_00, a_ii) and np.allclose(m_01, a_ij) and np.allclose(m_10, a_ji)


def _device(rng, nb, b, ne):
    """BTD system matrix (E + i eta) I - H as a DSDBCOO stack."""
    e = rng.uniform(-1, 1, ne) + 1e-3j
    hd = [_cn(rng, (b, b)) / np.sqrt(b) for _ in range(nb)]
    hd = [(x + x.conj().T) / 2 for x in hd]
    hu = [_cn(rng, (b, b)) / np.sqrt(b) for _ in range(nb - 1)]
    n = nb * b
    h = np.zeros((n, n), complex)
    for i in range(nb):
        h[i * b:(i + 1) * b, i * b:(i + 1) * b] = hd[i]
        if i + 1 < nb:
            h[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = hu[i]
            h[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = hu[i].conj().T
    mask = sps.coo_matrix(np.where(h != 0, 1.0 + 0j, 0))
    a = DSDBCOO.from_sparray(mask, [b] * nb, (ne,))
    dense = e[:, None, None] * np.eye(n) - h
    a.data[:] = dense[:, a.rows, a.cols]
    return a, dense


def _obc():
    return OBCSystem(SanchoRubio(max_iterations=500, convergence_tol=1e-12))


def test_contact_self_energies_fixed_points_and_identity():
    rng = np.random.default_rng(1)
    a, _ = _device(rng, 4, 4, 5)
    fl, fr = rng.uniform(0, 1, 5), rng.uniform(0, 1, 5)
    obc = contact_obc_blocks(a, _obc(), fl, fr)
    n = a.num_blocks - 1
    for j in (0, n):
        s_r, s_l, s_g = obc.retarded[j], obc.lesser[j], obc.greater[j]
        assert rel_err(s_g - s_l, s_r - np.conj(np.swapaxes(s_r, -1, -2))) < 1e-12
    # left: g = inv(m00 - m10 g m01), sigma = m10 g m01
    m00, m01, m10 = a.blocks[0, 0], a.blocks[0, 1], a.blocks[1, 0]
    sig = obc.retarded[0]
    g = np.linalg.solve(m10, sig @ np.linalg.inv(m01))
    assert rel_err(g, np.linalg.inv(m00 - m10 @ g @ m01)) < 1e-8


def test_right_contact_is_mirrored_left_contact():
    rng = np.random.default_rng(2)
    a, dense = _device(rng, 4, 3, 4)
    mirrored = np.flip(dense, axis=(-2, -1)).copy()
    am = DSDBCOO.from_sparray(sps.coo_matrix(np.where(mirrored[0] != 0, 1.0 + 0j, 0)),
                             [3] * 4, (4,))
    am.data[:] = mirrored[:, am.rows, am.cols]
    f = rng.uniform(0, 1, 4)
    o = contact_obc_blocks(a, _obc(), f, f)
    om = contact_obc_blocks(am, _obc(), f, f)
    flip = lambda x: np.flip(x, axis=(-2, -1))
    assert rel_err(om.retarded[3], flip(o.retarded[0])) < 1e-9
    assert rel_err(om.lesser[3], flip(o.lesser[0])) < 1e-9


@pytest.mark.parametrize("s", [1, 2, 3, 4])
@pytest.mark.parametrize("stack", [(), (3,), (2, 2)])
def test_device_superblocks_match_oracle(s, stack):
    """quatrex_amd.leads.periodic_superblocks is data movement only (torch on CPU here)."""
    import torch

    from quatrex_amd.leads import periodic_superblocks as p_superblocks

    rng = np.random.default_rng(s)
    a = [_cn(rng, stack + (12, 12)) for _ in range(3)]
    ref = periodic_superblocks(*a, s)
    got = p_superblocks(*[torch.as_tensor(x) for x in a], s)
    for r, g in zip(ref, got):
        assert np.array_equal(r, g.numpy())
