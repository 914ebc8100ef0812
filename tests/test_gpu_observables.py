"""Observables and system-matrix assembly on the HIP path against the oracle."""

import numpy as np
import pytest
import scipy.sparse as sps
import torch

import oracle.rgf as orgf
from oracle.observables import assemble_system_matrix as o_assemble
from oracle.observables import contact_currents as o_currents
from oracle.observables import density as o_density
from quatrex_amd.datastructures import DSDBCOO
from quatrex_amd.greens_function_solver import RGF, OBCBlocks
from quatrex_amd.observables import SystemMatrixAssembler, contact_currents, density
from tests.conftest import rel_err
from tests.test_gpu_rgf import _oracle_containers, to_gpu
from tests.test_oracle_observables import _btd_mask, _cn, _container
from quatrex_amd.workloads import make_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solved():
    w = make_workload(seed=4)
    oa, osl, osg = _oracle_containers(w)
    ref = orgf.RGF(16).selected_solve(oa, osl, osg)
    return w, (oa, osl, osg), ref


@pytest.mark.parametrize("max_batch_size", [16, 5])
def test_fused_density_output(device, solved, max_batch_size):
    w, (oa, osl, osg), (rl, rg) = solved
    n_orb = int(np.sum(w.block_sizes))
    dens = torch.zeros((oa.data.shape[0], n_orb), dtype=torch.float64, device=device)
    xl, xg = RGF(max_batch_size).selected_solve(to_gpu(oa, device), to_gpu(osl, device),
                                                to_gpu(osg, device), density=dens)
    ref = o_density(rl)
    assert rel_err(dens.cpu().numpy(), ref) < 1e-10
    assert rel_err(density(xl).cpu().numpy(), ref) < 1e-10


def test_density_with_overlap(device):
    rng = np.random.default_rng(0)
    bs = [3, 4, 2]
    n = sum(bs)
    g = _cn(rng, (5, n, n))
    x = _container(g, bs)
    s = _cn(rng, (n, n))
    s = np.where(_btd_mask(bs), s + s.conj().T, 0)
    gx = DSDBCOO(torch.as_tensor(x.data, device=device), x.rows, x.cols, x.block_sizes,
                 x.global_stack_shape)
    got = density(gx, sps.csr_matrix(s)).cpu().numpy()
    assert rel_err(got, o_density(x, sps.csr_matrix(s))) < 1e-13


def test_contact_currents(device, solved):
    w, (oa, osl, osg), _ = solved
    ne, b0, bn, nb = oa.data.shape[0], w.block_sizes[0], w.block_sizes[-1], w.n_blocks
    rng = np.random.default_rng(3)
    obc = orgf.OBCBlocks(nb)
    obc.lesser[0] = 0.1j * _cn(rng, (ne, b0, b0))
    obc.greater[0] = -0.1j * _cn(rng, (ne, b0, b0))
    obc.lesser[-1] = 0.2j * _cn(rng, (ne, bn, bn))
    obc.greater[-1] = -0.2j * _cn(rng, (ne, bn, bn))
    xl, xg = orgf.RGF(16).selected_solve(oa, osl, osg, obc_blocks=obc)
    ref = o_currents(xl, xg, obc)
    gobc = OBCBlocks(nb)
    for name in ("lesser", "greater"):
        for j in (0, -1):
            getattr(gobc, name)[j] = torch.as_tensor(getattr(obc, name)[j], device=device)
    gl, gg = to_gpu(xl, device), to_gpu(xg, device)
    got = contact_currents(gl, gg, gobc)
    for r, g in zip(ref, got):
        assert rel_err(g.cpu().numpy(), r) < 1e-12


@pytest.mark.parametrize("with_overlap", [False, True])
def test_system_matrix_assembly(device, with_overlap):
    rng = np.random.default_rng(5)
    bs = [3, 3, 2, 4]
    n, ne = sum(bs), 6
    h = _cn(rng, (n, n))
    h = np.where(_btd_mask(bs), h + h.conj().T, 0)
    sse = np.where(_btd_mask(bs), _cn(rng, (ne, n, n)), 0)
    v = rng.standard_normal(n)
    e = rng.uniform(-1, 1, ne)
    s = np.where(_btd_mask(bs), np.eye(n) + 0.1 * _cn(rng, (n, n)), 0) if with_overlap else None
    a = _container(np.zeros((ne, n, n), complex), bs)
    hc = _container(np.broadcast_to(h, (ne, n, n)).copy(), bs)
    sc = _container(sse, bs)
    ov = sps.csr_matrix(s) if with_overlap else None
    o_assemble(a, e, 1e-3, hc, sc, overlap=ov, potential=v)
    pattern = DSDBCOO(torch.zeros((ne, a.data.shape[-1]), dtype=torch.complex128, device=device),
                      a.rows, a.cols, a.block_sizes, a.global_stack_shape)
    gs = DSDBCOO(torch.as_tensor(sc.data, device=device), sc.rows, sc.cols, sc.block_sizes,
                 sc.global_stack_shape)
    asm = SystemMatrixAssembler(pattern, sps.csr_matrix(h), overlap=ov, potential=v)
    got = asm(e, 1e-3, gs)
    assert rel_err(got.data.cpu().numpy(), a.data) < 1e-14
