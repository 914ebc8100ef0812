"""Fused reflected-mode classification of the Spectral OBC
(`ops.spectral_modes`, csrc/spectral.hip) against the NumPy oracle's
`Spectral._find_reflected_modes` and `_compute_dE_dk` (spectral.py:160-320)
on the eigenpairs of random leads: the same mask, dE/dk to rounding, and the
residuals of the oracle's formula.  Eigenvalues at zero and infinity (a
singular coupling block gives both) must be classified as the oracle does
(never reflected: their residual is not finite)."""

import numpy as np
import pytest
import torch

import oracle.nevp as onevp
import oracle.obc as oobc
from quatrex_amd.ops import spectral_modes
from quatrex_amd.workloads import cn

pytestmark = pytest.mark.gpu


def _lead(rng, ne, b, rank=None):
    x = cn(rng, (b, b)) / np.sqrt(b)
    h00 = (x + x.conj().T) / 2
    h01 = cn(rng, (b, b)) / np.sqrt(b)
    if rank is not None:  # low-rank coupling: eigenvalues at 0 and infinity
        h01[:, rank:] = 0.0
    e = rng.uniform(-1, 1, ne) + 1e-3j
    m_ii = e[:, None, None] * np.eye(b) - h00
    m_ij = np.broadcast_to(-h01, (ne, b, b)).copy()
    m_ji = np.broadcast_to(-h01.conj().T, (ne, b, b)).copy()
    return m_ji, m_ii, m_ij


@pytest.mark.parametrize("b,rank,norm", [(24, None, True), (40, None, False), (32, 12, True)])
def test_modes_match_oracle(device, b, rank, norm):
    rng = np.random.default_rng(b)
    m_ji, m_ii, m_ij = _lead(rng, 3, b, rank)
    ws, vs = onevp.Full()((m_ji, m_ii, m_ij))
    with np.errstate(divide="ignore", invalid="ignore"):
        vs = vs / np.linalg.norm(vs, axis=-2, keepdims=True)
    spec = oobc.Spectral(onevp.Full(), residual_normalization=norm)
    want = spec._find_reflected_modes(ws, vs, (m_ji, m_ii, m_ij))
    dk_o = np.nan_to_num(spec._compute_dE_dk(ws, vs, (m_ji, m_ii, m_ij)), nan=0, posinf=0,
                         neginf=0)
    avs = [a @ vs for a in (m_ji, m_ii, m_ij)]
    T = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=device)
    mask, res, dk = spectral_modes(
        T(ws), T(vs), *(T(a) for a in avs), min_decay=spec.min_decay,
        max_decay=spec.max_decay, min_propagation=spec.min_propagation,
        residual_tolerance=spec.residual_tolerance, eta_decay=spec.eta_decay,
        residual_normalization=norm, details=True)
    mask, res, dk = mask.cpu().numpy(), res.cpu().numpy(), dk.cpu().numpy()
    assert np.array_equal(mask, want)
    assert mask.any() and not mask.all()
    # the eigenvalues at zero and infinity of a rank-deficient coupling come
    # out as |w| ~ 1e-16 or inf: their dE/dk is rounding noise divided by w,
    # so values are compared on the others (the masks agree on all)
    fin = np.isfinite(ws) & (np.abs(ws) > 1e-8) & (np.abs(ws) < 1e8)
    fin &= np.isfinite(vs).all(axis=-2)
    if rank is not None:
        assert not fin.all()  # zero / infinite eigenvalues are present
    with np.errstate(divide="ignore", invalid="ignore"):
        prod = avs[0] / ws[:, None, :] + avs[1] + avs[2] * ws[:, None, :]
        res_o = np.linalg.norm(prod, axis=-2) / np.linalg.norm(vs, axis=-2)
        if norm:
            res_o = res_o / np.abs(ws)
    ok = fin & np.isfinite(res_o)
    assert np.allclose(res[ok], res_o[ok], rtol=1e-10, atol=1e-14)
    scale = np.abs(dk_o[fin]).max()
    assert np.abs(dk[fin] - dk_o[fin]).max() <= 1e-10 * scale
    assert not mask[~(np.isfinite(ws) & (ws != 0))].any()
