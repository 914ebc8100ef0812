"""Pins the NumPy OBC / NEVP oracle.

Reference tests mirrored: tests/qttools/boundary_conditions/obc/
test_sancho_rubio.py (fixed-point residual, max_iterations warning,
memoizer), test_nevp.py (eigen-residual after a robust z-score filter, Full
with reduction), test_spectral.py (recursion residual < 5e-3, closed forms of
_upscale_eigenmodes and _compute_dE_dk), tests/qttools/kernels/
test_operator.py (operator_inverse vs a direct inverse).  The carbon-nanotube
lead of the reference fixture is a Git-LFS stub here, so synthetic leads are
used.  Beyond the reference: Sancho-Rubio against the closed-form surface
Green's function of a 1D chain (scalar and unitarily rotated blocks), Beyn
eigenvalues against the companion linearisation, and Spectral x_ii against
Sancho-Rubio.
"""

import warnings

import numpy as np
import pytest
import scipy.sparse as sps

from oracle.nevp import Beyn, Full, operator_inverse
from oracle.obc import OBCSystem, SanchoRubio, Spectral
from quatrex_amd.workloads import cn
from tests.conftest import rel_err


def _h(x):
    return np.conj(np.swapaxes(x, -2, -1))


def chain_surface_g(z, eps0, t):
    """Retarded surface GF of a semi-infinite chain: root of t^2 g^2 - (z-eps0) g + 1 = 0
    with Im g < 0."""
    d = z - eps0
    s = np.sqrt(d * d - 4 * t * t)
    g1, g2 = (d - s) / (2 * t * t), (d + s) / (2 * t * t)
    return np.where(g1.imag < 0, g1, g2)


def test_sancho_rubio_scalar_chain():
    z = np.array([-1.5, -0.3, 0.0, 0.7, 2.5]) + 1e-3j
    eps0, t = 0.2, 0.8
    a_ii = (z - eps0)[:, None, None]
    a_ij = np.full((len(z), 1, 1), -t, complex)
    g = SanchoRubio(max_iterations=200, convergence_tol=1e-13)(a_ii, a_ij, a_ij.copy())
    assert np.allclose(g[:, 0, 0], chain_surface_g(z, eps0, t), rtol=1e-9, atol=1e-12)


def test_sancho_rubio_rotated_block_chain():
    rng = np.random.default_rng(1)
    b = 6
    q, _ = np.linalg.qr(cn(rng, (b, b)))
    eps = rng.uniform(-1, 1, b)
    t = rng.uniform(0.3, 1.0, b)
    z = np.array([-0.4, 0.1, 0.9]) + 1e-3j
    h00 = q @ np.diag(eps) @ q.conj().T
    h01 = q @ np.diag(t) @ q.conj().T
    a_ii = z[:, None, None] * np.eye(b) - h00
    a_ij = np.broadcast_to(-h01, (len(z), b, b)).copy()
    a_ji = np.broadcast_to(-h01.conj().T, (len(z), b, b)).copy()
    g = SanchoRubio(max_iterations=200, convergence_tol=1e-13)(a_ii, a_ij, a_ji)
    ref = np.stack([q @ np.diag(chain_surface_g(zz, eps, t)) @ q.conj().T for zz in z])
    assert rel_err(g, ref) < 1e-9


def lead(b=12, energies=(-0.6, 0.0, 0.5), eta=1e-3, seed=0, rank=None):
    """Synthetic lead; ``rank`` < b gives a rank-deficient coupling H01 (as real
    leads have), so that at most 2 rank eigenvalues are finite and nonzero."""
    rng = np.random.default_rng(seed)
    x = cn(rng, (b, b)) / np.sqrt(b)
    h00 = (x + x.conj().T) / 2
    if rank is None:
        h01 = cn(rng, (b, b)) / np.sqrt(b)
    else:
        h01 = cn(rng, (b, rank)) @ cn(rng, (rank, b)) / b
    e = np.asarray(energies) + 1j * eta
    a_ii = e[:, None, None] * np.eye(b) - h00
    a_ij = np.broadcast_to(-h01, (len(e), b, b)).copy()
    a_ji = np.broadcast_to(-h01.conj().T, (len(e), b, b)).copy()
    return a_ji, a_ii, a_ij


def test_sancho_rubio_fixed_point_and_warning():
    a_ji, a_ii, a_ij = lead()
    x = SanchoRubio(convergence_tol=1e-12)(a_ii, a_ij, a_ji)
    assert rel_err(x, np.linalg.inv(a_ii - a_ji @ x @ a_ij)) < 1e-9
    with pytest.warns(RuntimeWarning):
        SanchoRubio(max_iterations=1, convergence_tol=1e-8)(a_ii, a_ij, a_ji)


class _Spy:
    def __init__(self, solver):
        self.solver, self.calls = solver, 0

    def __call__(self, *args):
        self.calls += 1
        return self.solver(*args)


@pytest.mark.parametrize("mode", ["off", "auto", "force-after-first", "force"])
def test_memoizer_modes(mode):
    a_ji, a_ii, a_ij = lead()
    spy = _Spy(SanchoRubio(convergence_tol=1e-12))
    obc = OBCSystem(spy, memoization_mode=mode)
    scale = 1 + 1e-6
    if mode == "force":
        # no solve ever: starting guess inv(a_ii) plus the residual step only
        with pytest.warns(RuntimeWarning, match="High error"):
            obc((a_ii, a_ij, a_ji), "left")
        assert spy.calls == 0
        return
    x, rel, abs_ = obc((a_ii, a_ij, a_ji), "left")
    assert rel_err(x, np.linalg.inv(a_ii - a_ji @ x @ a_ij)) < 1e-8
    x2, rel2, _ = obc((a_ii * scale, a_ij * scale, a_ji * scale), "left")
    assert np.allclose(x2, np.linalg.inv(a_ii * scale - a_ji * scale @ x2 @ a_ij * scale),
                       atol=1e-5)
    expected = {"off": 2, "force-after-first": 1, "force": 0,
                # reference quirk: a cache that passes the tolerances on more than
                # agreement_threshold of the (single-rank) stack triggers a re-solve
                "auto": 2}[mode]
    assert spy.calls == expected


def test_num_ref_iterations_coerced():
    with pytest.warns(RuntimeWarning):
        o = OBCSystem(SanchoRubio(), num_ref_iterations=1)
    assert o.num_ref_iterations == 2


@pytest.mark.parametrize("nq,nb", [(4, 1), (7, 2)])
def test_operator_inverse(nq, nb):
    rng = np.random.default_rng(0)
    a_xx = tuple(rng.random((3, 1, 5, 5)) + 1j * rng.random((3, 1, 5, 5)) for _ in range(2 * nb + 1))
    z = rng.random((1, nq, 1, 1)) + 1j * rng.random((1, nq, 1, 1))
    got = operator_inverse(a_xx, z, z.dtype, z.dtype)
    ref = np.linalg.inv(sum(z ** n * a for a, n in zip(a_xx, range(-nb, nb + 1))))
    assert np.allclose(got, ref)


def _residuals(a_xx, ws, vs):
    a_ji, a_ii, a_ij = a_xx
    res = []
    for e in range(ws.shape[0]):
        for k in range(ws.shape[1]):
            w = ws[e, k]
            v = vs[e, :, k] / np.linalg.norm(vs[0, :, k])
            with np.errstate(divide="ignore", invalid="ignore"):
                res.append(np.linalg.norm((a_ji[e] / w + a_ii[e] + a_ij[e] * w) @ v)
                           / np.linalg.norm(w))
    res = np.nan_to_num(np.array(res))
    med = np.median(res)
    mad = np.median(np.abs(res - med))
    z = 0.6745 * (res - med) / mad
    keep = ~(np.abs(z) > 30)
    return res, keep


@pytest.mark.parametrize("use_qr", [False, True])
def test_beyn_residual_and_eigenvalues(use_qr):
    # Beyn with the moments A0, A1 resolves at most b eigenvalues inside the
    # annulus: a rank-b/2 coupling keeps the finite nonzero spectrum at b.
    a_xx = lead(b=10, rank=5)
    beyn = Beyn(r_o=3.0, r_i=1 / 3.0, m_0=10, num_quad_points=64, use_qr=use_qr)
    ws, vs = beyn(a_xx)
    res, keep = _residuals(a_xx, ws, vs)
    assert keep.any() and res[keep].max() < 1e-5
    wf, _ = Full()(a_xx)
    for e in range(ws.shape[0]):
        inside = wf[e][(np.abs(wf[e]) > 1 / 3.0 * 1.05) & (np.abs(wf[e]) < 3.0 * 0.95)]
        found = ws[e][np.abs(ws[e]) > 0]
        for lam in inside:
            assert np.min(np.abs(found - lam)) < 1e-6 * max(1.0, abs(lam))


@pytest.mark.parametrize("reduce", [False, True])
def test_full_residual(reduce):
    a_xx = tuple(a.copy() for a in lead(b=8))
    if reduce:
        a_xx[0][..., :4] = 0
        a_xx[2][..., 4:] = 0
    ws, vs = Full(reduce=reduce, a_xx_sparsity=tuple(sps.csc_matrix(a[0]) for a in a_xx))(a_xx)
    res, keep = _residuals(a_xx, ws, vs)
    assert keep.any() and res[keep].max() < 1e-5


def test_spectral_full_matches_sancho_rubio():
    a_ji, a_ii, a_ij = lead(b=8, eta=1e-3)
    x_sr = SanchoRubio(max_iterations=500, convergence_tol=1e-13)(a_ii, a_ij, a_ji)
    x_sp = Spectral(Full(), residual_tolerance=1e-6)(a_ii, a_ij, a_ji)
    assert rel_err(x_sp, x_sr) < 1e-7


def test_spectral_beyn_recursion_residual():
    a_ji, a_ii, a_ij = lead(b=8, eta=1e-3, rank=4)
    x = Spectral(Beyn(r_o=10.0, r_i=0.99, m_0=8, num_quad_points=64),
                 residual_tolerance=1e-1, max_decay=20)(a_ii, a_ij, a_ji)
    r = np.linalg.norm(x - np.linalg.inv(a_ii - a_ji @ x @ a_ij), axis=(-1, -2)) / \
        np.linalg.norm(x, axis=(-1, -2))
    assert np.all(r < 5e-3)


@pytest.mark.parametrize("block_sections", [1, 3])
def test_upscale_eigenmodes_closed_form(block_sections):
    rng = np.random.default_rng(3)
    bsz, batch = 18, 2
    sp = Spectral(Full(), block_sections=block_sections)
    ws = rng.random((batch, bsz)) + 1j * rng.random((batch, bsz))
    vs = rng.random((batch, bsz // block_sections, bsz)) + 1j * rng.random((batch, bsz // block_sections, bsz))
    _, vu = sp._upscale_eigenmodes(ws, vs)
    ref = np.zeros((batch, bsz, bsz), complex)
    for i in range(batch):
        for j, w in enumerate(ws[i]):
            ref[i, :, j] = np.kron(np.array([w ** n for n in range(block_sections)]), vs[i, :, j])
            ref[i, :, j] /= np.linalg.norm(ref[i, :, j])
    assert np.allclose(vu, ref)


def test_compute_dE_dk_closed_form():
    a_xx = lead(b=6)
    rng = np.random.default_rng(4)
    batch, b = a_xx[0].shape[0], a_xx[0].shape[-1]
    ws = rng.random((batch, b)) + 1j * rng.random((batch, b))
    vrs = rng.random((batch, b, b)) + 1j * rng.random((batch, b, b))
    got = Spectral(Full())._compute_dE_dk(ws, vrs, a_xx)
    ref = np.zeros_like(ws)
    for i in range(batch):
        for j, w in enumerate(ws[i]):
            a = -sum((1j * n) * w ** n * a_xn[i] for a_xn, n in zip(a_xx, range(-1, 2)))
            ref[i, j] = vrs[i, :, j].conj() @ a @ vrs[i, :, j]
    assert np.allclose(got, ref)
