"""Pins the observables / assembly oracle against dense linear algebra.

- density: Im diag of the dense G (no overlap); with a BTD overlap S,
  Im diag of G_btd S where G_btd is G restricted to the block tridiagonal
  (observables.py:57-126 sums exactly those block products);
- contact currents: in a ballistic device (self-energies from the leads
  only) I_L = -I_R and |I_L| = |f_L - f_R| T(E) with the Caroli
  transmission T = Tr[Gamma_L G^R Gamma_R G^A] of the dense inverse;
- assembly: A = (E + i eta) I - diag(V) - Sigma^R - H on the pattern.
"""

import numpy as np
import scipy.sparse as sps

from oracle.dsdbcoo import DSDBCOO
from oracle.observables import assemble_system_matrix, contact_currents, density
from oracle.rgf import RGF, OBCBlocks
from tests.conftest import rel_err


def _cn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def _btd_mask(bs):
    off = np.concatenate([[0], np.cumsum(bs)])
    n = off[-1]
    m = np.zeros((n, n), bool)
    for i in range(len(bs)):
        for j in (i - 1, i, i + 1):
            if 0 <= j < len(bs):
                m[off[i]:off[i + 1], off[j]:off[j + 1]] = True
    return m


def _container(dense, bs):
    mask = _btd_mask(bs)
    c = DSDBCOO.from_sparray(sps.coo_matrix(mask.astype(complex)), bs, dense.shape[:1])
    c.data[:] = dense[:, c.rows, c.cols]
    return c


def test_density_matches_dense_diagonal():
    rng = np.random.default_rng(0)
    bs = [3, 4, 2]
    n = sum(bs)
    g = _cn(rng, (5, n, n))
    x = _container(g, bs)
    assert np.allclose(density(x), np.diagonal(g, axis1=-2, axis2=-1).imag)
    s = _cn(rng, (n, n))
    s = np.where(_btd_mask(bs), s + s.conj().T, 0)
    g_btd = np.where(_btd_mask(bs), g, 0)
    ref = np.diagonal(g_btd @ s, axis1=-2, axis2=-1).imag
    assert np.allclose(density(x, sps.csr_matrix(s)), ref)


def test_contact_currents_ballistic_landauer():
    rng = np.random.default_rng(1)
    bs = [4, 4, 4, 4]
    nb, b, n = len(bs), 4, 16
    ne = 3
    h = _cn(rng, (n, n)) / 2
    h = np.where(_btd_mask(bs), h + h.conj().T, 0)
    sl_ = -1j * np.eye(b) * 0.3 + _cn(rng, (b, b)) * 0.05
    sr_ = -1j * np.eye(b) * 0.2
    sig_l = np.broadcast_to((sl_ + sl_.conj().T) / 2 - 0.3j * np.eye(b), (ne, b, b)).copy()
    sig_r = np.broadcast_to(sr_, (ne, b, b)).copy()
    e = np.array([-0.4, 0.1, 0.6])
    a_dense = e[:, None, None] * np.eye(n) - h
    a = _container(a_dense, bs)
    fl, fr = np.array([0.9, 0.7, 0.4]), np.array([0.1, 0.2, 0.3])
    obc = OBCBlocks(nb)
    obc.retarded[0], obc.retarded[-1] = sig_l, sig_r
    gam = lambda s: 1j * (s - np.conj(np.swapaxes(s, -1, -2)))
    f3 = lambda f: f[:, None, None]
    obc.lesser[0], obc.greater[0] = 1j * f3(fl) * gam(sig_l), 1j * f3(fl - 1) * gam(sig_l)
    obc.lesser[-1], obc.greater[-1] = 1j * f3(fr) * gam(sig_r), 1j * f3(fr - 1) * gam(sig_r)
    zero = DSDBCOO.zeros_like(a)
    xl, xg = RGF().selected_solve(a, zero, DSDBCOO.zeros_like(a), obc_blocks=obc)
    i_l, i_r = contact_currents(xl, xg, obc)
    full = a_dense.copy()
    full[:, :b, :b] -= sig_l
    full[:, -b:, -b:] -= sig_r
    gr = np.linalg.inv(full)
    g1n = gr[:, :b, -b:]
    t = np.trace(gam(sig_l) @ g1n @ gam(sig_r) @ np.conj(np.swapaxes(g1n, -1, -2)),
                 axis1=-2, axis2=-1).real
    assert np.allclose(i_l, -i_r, atol=1e-10)
    assert np.allclose(np.abs(i_l), np.abs(fl - fr) * t, rtol=1e-9)


def test_assemble_system_matrix_matches_dense():
    rng = np.random.default_rng(2)
    bs = [3, 3, 2]
    n, ne = sum(bs), 4
    h = _cn(rng, (n, n))
    h = np.where(_btd_mask(bs), h + h.conj().T, 0)
    sse = np.where(_btd_mask(bs), _cn(rng, (ne, n, n)), 0)
    v = rng.standard_normal(n)
    e = rng.uniform(-1, 1, ne)
    a = _container(np.zeros((ne, n, n), complex), bs)
    hc = _container(np.broadcast_to(h, (ne, n, n)).copy(), bs)
    sc = _container(sse, bs)
    assemble_system_matrix(a, e, 1e-3, hc, sc, potential=v)
    ref = (e + 1e-3j)[:, None, None] * np.eye(n) - np.diag(v) - sse - h
    assert rel_err(a.to_dense(), np.where(_btd_mask(bs), ref, 0)) < 1e-14
    s = np.where(_btd_mask(bs), np.eye(n) + 0.1 * _cn(rng, (n, n)), 0)
    a2 = _container(np.zeros((ne, n, n), complex), bs)
    assemble_system_matrix(a2, e, 1e-3, hc, sc, overlap=sps.csr_matrix(s), potential=v)
    ref2 = (e + 1e-3j)[:, None, None] * s - np.diag(v) - sse - h
    assert rel_err(a2.to_dense(), np.where(_btd_mask(bs), ref2, 0)) < 1e-14
