"""Open-boundary solvers on the HIP path against the NumPy oracle.

Sancho-Rubio (same batch-global stop rule and iteration count), the memoizing
OBCSystem, the Beyn contour moments and eigenpairs (same probe stream), Full
and Spectral.  Bar: 1e-10 relative Frobenius error for the complex128
quantities; eigenvalues compared as sets.
"""

import numpy as np
import pytest
import torch

import oracle.nevp as onevp
import oracle.obc as oobc
from quatrex_amd.boundary_conditions import OBCSystem, SanchoRubio, Spectral
from quatrex_amd.nevp import Beyn, Full, operator_inverse
from tests.conftest import rel_err
from tests.test_oracle_obc import chain_surface_g, lead

pytestmark = pytest.mark.gpu


def np_(t):
    return t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


@pytest.mark.parametrize("b,ne", [(12, 3), (64, 8)])
def test_sancho_rubio_matches_oracle(device, b, ne):
    a_ji, a_ii, a_ij = lead(b=b, energies=np.linspace(-0.8, 0.8, ne), eta=1e-3, seed=b)
    o = oobc.SanchoRubio(max_iterations=100, convergence_tol=1e-12)
    ref = o(a_ii, a_ij, a_ji)
    g = SanchoRubio(max_iterations=100, convergence_tol=1e-12)
    x = g(torch.as_tensor(a_ii, device=device), a_ij, a_ji)
    assert g.iterations == o.iterations
    assert rel_err(np_(x), ref) < 1e-10


def test_sancho_rubio_chain_closed_form(device):
    z = np.array([-1.5, -0.3, 0.0, 0.7, 2.5]) + 1e-3j
    a_ii = (z - 0.2)[:, None, None]
    a_ij = np.full((len(z), 1, 1), -0.8, complex)
    x = SanchoRubio(max_iterations=200, convergence_tol=1e-13)(a_ii, a_ij, a_ij.copy())
    assert np.allclose(np_(x)[:, 0, 0], chain_surface_g(z, 0.2, 0.8), rtol=1e-9, atol=1e-12)


def test_sancho_rubio_single_matrix_and_warning(device):
    a_ji, a_ii, a_ij = lead()
    x = SanchoRubio(convergence_tol=1e-10)(a_ii[0], a_ij[0], a_ji[0])
    assert x.shape == a_ii[0].shape
    xn = np_(x)
    assert np.allclose(xn, np.linalg.inv(a_ii[0] - a_ji[0] @ xn @ a_ij[0]))
    with pytest.warns(RuntimeWarning):
        SanchoRubio(max_iterations=1, convergence_tol=1e-8)(a_ii, a_ij, a_ji)


@pytest.mark.parametrize("mode", ["off", "auto", "force-after-first"])
def test_obc_system_matches_oracle(device, mode):
    a_ji, a_ii, a_ij = lead()
    ref_sys = oobc.OBCSystem(oobc.SanchoRubio(convergence_tol=1e-12), memoization_mode=mode)
    sys_ = OBCSystem(SanchoRubio(convergence_tol=1e-12), memoization_mode=mode)
    for scale in (1.0, 1 + 1e-6):
        xr, rr, ar = ref_sys((a_ii * scale, a_ij * scale, a_ji * scale), "left")
        x, r, a = sys_((a_ii * scale, a_ij * scale, a_ji * scale), "left")
        assert rel_err(np_(x), xr) < 1e-9
        # At convergence both residuals are rounding noise (~1e-12), hence the
        # absolute floor; the refinement-only memoized calls have real
        # residuals (~1e-5), equal up to the 1e-9 difference in x.
        assert np.allclose(np_(a), ar, rtol=1e-3, atol=1e-10)
        assert np.allclose(np_(r), rr, rtol=1e-3, atol=1e-9)


def test_operator_inverse(device):
    rng = np.random.default_rng(0)
    a_xx = tuple(rng.random((3, 1, 6, 6)) + 1j * rng.random((3, 1, 6, 6)) for _ in range(3))
    z = rng.random((1, 5, 1, 1)) + 1j * rng.random((1, 5, 1, 1))
    got = np_(operator_inverse(a_xx, z))
    ref = np.linalg.inv(sum(z ** n * a for a, n in zip(a_xx, range(-1, 2))))
    assert rel_err(got, ref) < 1e-12


@pytest.mark.parametrize("b,nq", [(10, 64), (64, 32)])
def test_beyn_moments_match_oracle(device, b, nq):
    a_xx = lead(b=b, rank=b // 2, seed=3)
    ob = onevp.Beyn(r_o=3.0, r_i=1 / 3.0, m_0=b, num_quad_points=nq)
    q0r, q1r = ob._contour_integrate(a_xx)
    gb = Beyn(r_o=3.0, r_i=1 / 3.0, m_0=b, num_quad_points=nq)
    q0, q1 = gb._contour_integrate([torch.as_tensor(a, device=device) for a in a_xx])
    assert rel_err(np_(q0), q0r) < 1e-10
    assert rel_err(np_(q1), q1r) < 1e-10


@pytest.mark.parametrize("use_qr", [False, True])
def test_beyn_eigenpairs_match_oracle(device, use_qr):
    a_xx = lead(b=10, rank=5)
    wr, vr = onevp.Beyn(r_o=3.0, r_i=1 / 3.0, m_0=10, num_quad_points=64, use_qr=use_qr)(a_xx)
    w, v = Beyn(r_o=3.0, r_i=1 / 3.0, m_0=10, num_quad_points=64, use_qr=use_qr)(a_xx)
    w = np_(w)
    a_ji, a_ii, a_ij = a_xx
    n_checked = 0
    for e in range(w.shape[0]):
        ref = wr[e][np.abs(wr[e]) > 0]
        got = w[e][np.abs(w[e]) > 0]
        assert len(ref) == len(got)
        # Only genuine eigenpairs inside the contour annulus are determined
        # by the moments (the rest span the noise subspace of a rank-deficient
        # lead and depend on rounding): select them by their residual, as
        # the reference tests do.
        for k, lam in enumerate(wr[e]):
            if not (1.05 / 3.0 < abs(lam) < 3.0 / 1.05):
                continue
            v = vr[e][:, k]
            res = np.linalg.norm((a_ji[e] / lam + a_ii[e] + a_ij[e] * lam) @ v)
            if res > 1e-8 * np.linalg.norm(v):
                continue
            n_checked += 1
            assert np.min(np.abs(got - lam)) < 1e-7 * max(1.0, abs(lam))
    assert n_checked > 0


def test_full_matches_oracle(device):
    a_xx = lead(b=8)
    wr, _ = onevp.Full()(a_xx)
    w, _ = Full()(a_xx)
    w = np_(w)
    for e in range(w.shape[0]):
        for lam in wr[e]:
            assert np.min(np.abs(w[e] - lam)) < 1e-8 * max(1.0, abs(lam))


def test_spectral_full_matches_sancho_rubio(device):
    a_ji, a_ii, a_ij = lead(b=8, eta=1e-3)
    x_ref = oobc.Spectral(onevp.Full(), residual_tolerance=1e-6)(a_ii, a_ij, a_ji)
    x = Spectral(Full(), residual_tolerance=1e-6)(a_ii, a_ij, a_ji)
    assert rel_err(np_(x), x_ref) < 1e-8
    x_sr = SanchoRubio(max_iterations=500, convergence_tol=1e-13)(a_ii, a_ij, a_ji)
    assert rel_err(np_(x), np_(x_sr)) < 1e-7


def _dev(x, device):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128, device=device)


@pytest.mark.parametrize("b", [8, 40])
def test_spectral_mode_classification_matches_oracle(device, b):
    """Same eigenpairs in, same reflected-mode mask and x_ii a_ij out."""
    a_xx = lead(b=b, eta=1e-3, seed=b)
    ws, vs = onevp.Full()(a_xx)
    osp = oobc.Spectral(onevp.Full(), residual_tolerance=1e-6)
    ows, ovs = osp._upscale_eigenmodes(ws, vs)
    omask = osp._find_reflected_modes(ows, ovs, list(a_xx))
    gsp = Spectral(Full(), residual_tolerance=1e-6)
    gws, gvs = gsp._upscale_eigenmodes(_dev(ws, device), _dev(vs, device))
    assert rel_err(np_(gvs), ovs) < 1e-13
    gmask = gsp._find_reflected_modes(gws, gvs, [_dev(a, device) for a in a_xx])
    assert np.array_equal(np_(gmask), omask)
    assert np.all(omask.sum(-1) == b)  # one reflected mode per orbital
    x_a = np.zeros((ws.shape[0], b, b), complex)
    for i in range(ws.shape[0]):
        vr, w = ovs[i][:, omask[i]], ows[i][omask[i]]
        x_a[i] = vr / w @ np.linalg.inv(vr.conj().T @ vr) @ vr.conj().T
    got = np_(gsp._x_ii_a_ij(gws, gvs, gmask))
    assert rel_err(got, x_a) < 1e-10


def test_spectral_beyn_matches_oracle(device):
    a_ji, a_ii, a_ij = lead(b=8, eta=1e-3, rank=4)
    kw = dict(r_o=10.0, r_i=0.99, m_0=8, num_quad_points=64)
    ref = oobc.Spectral(onevp.Beyn(**kw), residual_tolerance=1e-1, max_decay=20)(a_ii, a_ij, a_ji)
    got = Spectral(Beyn(**kw), residual_tolerance=1e-1, max_decay=20)(a_ii, a_ij, a_ji)
    assert rel_err(np_(got), ref) < 1e-9


@pytest.mark.parametrize("block_sections", [2])
def test_spectral_block_sections_matches_oracle(device, block_sections):
    from tests.test_oracle_leads import _periodic_layer

    rng = np.random.default_rng(9)
    a_ii, a_ij, a_ji = _periodic_layer(rng, block_sections, 4, 3)
    z = (np.array([-0.3, 0.2, 0.5]) + 1e-3j)[:, None, None]
    m_ii = z * np.eye(a_ii.shape[-1]) - a_ii
    ref = oobc.Spectral(onevp.Full(), block_sections=block_sections, residual_tolerance=1e-6)(
        m_ii, -a_ij, -a_ji)
    got = Spectral(Full(), block_sections=block_sections, residual_tolerance=1e-6)(
        m_ii, -a_ij, -a_ji)
    assert rel_err(np_(got), ref) < 1e-9


@pytest.mark.parametrize("shape", [(5, 256, 32), (3, 64, 40), (4, 33, 33), (2, 50, 7)],
                         ids=["256x32", "64x40", "33x33", "50x7"])
def test_device_svd_against_numpy(device, shape):
    """One-sided Jacobi SVD (Beyn's projection on the device): singular
    values against LAPACK's, reconstruction, orthonormal V and U columns,
    descending order; one column repeated makes the matrix rank-deficient."""
    from quatrex_amd.ops import svd

    n, d, m = shape
    rng = np.random.default_rng(d + m)
    a = rng.standard_normal((n, d, m)) + 1j * rng.standard_normal((n, d, m))
    a[:, :, min(3, m - 1)] = 2 * a[:, :, 0]
    u, s, v = (np_(t) for t in svd(torch.as_tensor(a, device=device)))
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(s - s_ref)) < 1e-13 * s_ref.max()
    assert np.all(np.diff(s, axis=-1) <= 0)
    rec = (u * s[:, None, :]) @ np.conj(np.swapaxes(v, -1, -2))
    assert rel_err(rec, a) < 1e-13
    eye = np.eye(m)
    assert np.abs(np.conj(np.swapaxes(v, -1, -2)) @ v - eye).max() < 1e-12
    keep = s[0] > 1e-10 * s[0, 0]
    uk = u[0][:, keep]
    assert np.abs(uk.conj().T @ uk - np.eye(keep.sum())).max() < 1e-12


def test_beyn_device_projection_matches_host(device):
    """The device projection (Jacobi SVD, U^H P1 V S^-1 on the device) gives
    the eigenpairs of the host projection: the projected matrices differ by a
    unitary similarity, so the eigenvalues agree and the eigenvectors V = U v
    are parallel."""
    a_xx = lead(b=24, rank=12, seed=4)
    kw = dict(r_o=3.0, r_i=1 / 3.0, m_0=24, num_quad_points=64)
    wh, vh = Beyn(project_compute_location="numpy", **kw)(a_xx)
    wd, vd = Beyn(project_compute_location="device", **kw)(a_xx)
    wh, vh, wd, vd = np_(wh), np_(vh), np_(wd), np_(vd)
    n_checked = 0
    for e in range(wh.shape[0]):
        assert np.sum(np.abs(wh[e]) > 0) == np.sum(np.abs(wd[e]) > 0)
        a_ji, a_ii, a_ij = (x[e] for x in a_xx)
        for k, lam in enumerate(wh[e]):
            if lam == 0 or not (1.05 / 3.0 < abs(lam) < 3.0 / 1.05):
                continue
            v = vh[e][:, k]
            if np.linalg.norm((a_ji / lam + a_ii + a_ij * lam) @ v) > 1e-8 * np.linalg.norm(v):
                continue
            j = int(np.argmin(np.abs(wd[e] - lam)))
            assert abs(wd[e][j] - lam) < 1e-9 * max(1.0, abs(lam))
            w2 = vd[e][:, j]
            cos = abs(np.vdot(v, w2)) / (np.linalg.norm(v) * np.linalg.norm(w2))
            assert cos > 1 - 1e-9
            n_checked += 1
    assert n_checked > 0
