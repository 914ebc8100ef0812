"""RGF on the HIP path against the NumPy oracle (same seeded inputs).

Cases follow the reference tests (tests/qttools/greens_function_solver/
conftest.py:10-124: block sizes [2]*10 and mixed [2,2,2,4,4,2,2,2], stacks
(10,), (7,2), (9,2,4), max_batch_size 1/2/100, out given or not,
return_retarded) plus the benchmark workload (16 energies, 8 blocks of 32).
Bar: complex128 arithmetic within 1e-10 relative Frobenius error per block
(elementwise rtol 1e-10 on the reference patterns); complex64 arithmetic
within 1e-4 per block.
"""

import numpy as np
import pytest
import scipy.sparse as sps
import torch

import oracle.rgf as orgf
from oracle.dsdbcoo import DSDBCOO as ODSDBCOO
from quatrex_amd.datastructures import DSDBCOO
from quatrex_amd.greens_function_solver import RGF, OBCBlocks
from quatrex_amd.workloads import make_workload
from tests.conftest import bt_dense, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["fused", "per-product"])
def rgf_path(request, monkeypatch):
    """Every case runs on both solve paths: small-block stacks take the
    fused one-workgroup-per-energy kernel unless QX_RGF_FUSED=0."""
    monkeypatch.setenv("QX_RGF_FUSED", "1" if request.param == "fused" else "0")
    return request.param

BLOCK_SIZES = [np.array([2] * 10), np.array([2] * 3 + [4] * 2 + [2] * 3)]
STACKS = [(10,), (7, 2), (9, 2, 4)]


def to_gpu(o: ODSDBCOO, dev) -> DSDBCOO:
    return DSDBCOO(torch.as_tensor(o.data, device=dev), o.rows, o.cols, o.block_sizes,
                   o.global_stack_shape, symmetry=o.symmetry)


def close(g: DSDBCOO, o: ODSDBCOO, rtol):
    got = g.data.cpu().numpy()
    ref = o.data
    scale = np.max(np.abs(ref)) if ref.size else 1.0
    return np.allclose(got, ref, rtol=rtol, atol=rtol * scale)


@pytest.mark.parametrize("block_sizes", BLOCK_SIZES, ids=["const", "mixed"])
@pytest.mark.parametrize("stack", STACKS, ids=["1D", "2D", "3D"])
@pytest.mark.parametrize("max_batch_size", [1, 2, 100])
@pytest.mark.parametrize("out", [True, False])
def test_selected_inv(device, block_sizes, stack, max_batch_size, out):
    dense = bt_dense(block_sizes)
    oa = ODSDBCOO.from_sparray(sps.coo_matrix(dense), block_sizes, stack)
    ref = orgf.RGF(max_batch_size).selected_inv(oa)
    a = to_gpu(oa, device)
    solver = RGF(max_batch_size=max_batch_size)
    if out:
        x = DSDBCOO.zeros_like(a)
        assert solver.selected_inv(a, out=x) is None
    else:
        x = solver.selected_inv(a)
    assert close(x, ref, 1e-10)


@pytest.mark.parametrize("block_sizes", BLOCK_SIZES, ids=["const", "mixed"])
@pytest.mark.parametrize("stack", STACKS, ids=["1D", "2D", "3D"])
@pytest.mark.parametrize("max_batch_size", [1, 2, 100])
@pytest.mark.parametrize("return_retarded", [True, False])
@pytest.mark.parametrize("out", [True, False])
def test_selected_solve(device, block_sizes, stack, max_batch_size, return_retarded, out):
    dense = bt_dense(block_sizes)
    coo = sps.coo_matrix(dense)
    b = coo - coo.conj().T
    oa = ODSDBCOO.from_sparray(coo, block_sizes, stack)
    ob = ODSDBCOO.from_sparray(sps.coo_matrix(b), block_sizes, stack)
    refs = orgf.RGF(max_batch_size).selected_solve(oa, ob, ob, return_retarded=True)
    a, bl, bg = to_gpu(oa, device), to_gpu(ob, device), to_gpu(ob, device)
    solver = RGF(max_batch_size=max_batch_size)
    if out:
        xs = [DSDBCOO.zeros_like(a) for _ in range(3)]
        assert solver.selected_solve(a, bl, bg, out=xs, return_retarded=return_retarded) is None
    else:
        xs = solver.selected_solve(a, bl, bg, return_retarded=return_retarded)
    assert close(xs[0], refs[0], 1e-10)
    assert close(xs[1], refs[1], 1e-10)
    if return_retarded:
        assert close(xs[2], refs[2], 1e-10)


def test_invalid_number_of_outputs(device):
    dense = bt_dense(BLOCK_SIZES[0])
    a = to_gpu(ODSDBCOO.from_sparray(sps.coo_matrix(dense), BLOCK_SIZES[0], (2,)), device)
    with pytest.raises(ValueError, match="Invalid number of output matrices."):
        RGF().selected_solve(a, a, a, out=[DSDBCOO.zeros_like(a), DSDBCOO.zeros_like(a)],
                             return_retarded=True)


def _oracle_containers(w):
    from tests.test_oracle_rgf import _containers, _dense_workload

    A, SL, SG = _dense_workload(w)
    return _containers(w, A, SL, SG)


@pytest.fixture(scope="module")
def workload():
    w = make_workload(seed=0)
    oa, osl, osg = _oracle_containers(w)
    ref = orgf.RGF(16).selected_solve(oa, osl, osg, return_retarded=True, return_current=True,
                                      retarded_offdiag=True)
    return w, (oa, osl, osg), ref


def _per_block_worst(w, g, o):
    worst = 0.0
    for i in range(w.n_blocks):
        for j in (i - 1, i, i + 1):
            if 0 <= j < w.n_blocks:
                gb = g.blocks[i, j].cpu().numpy()
                worst = max(worst, rel_err(gb, o.blocks[i, j]))
    return worst


@pytest.mark.parametrize("max_batch_size", [16, 5])
def test_workload_parity_c128(device, workload, max_batch_size):
    w, (oa, osl, osg), (rl, rg, rr, rcur) = workload
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    xl, xg, xr, cur = RGF(max_batch_size, retarded_offdiag=True).selected_solve(
        a, sl, sg, return_retarded=True, return_current=True)
    for g, o in ((xl, rl), (xg, rg), (xr, rr)):
        assert _per_block_worst(w, g, o) < 1e-10
    assert rel_err(cur.cpu().numpy(), rcur) < 1e-10


def test_workload_parity_c128_3m(device, workload):
    """complex128 with the 3M product split: same bar on this workload."""
    w, (oa, osl, osg), (rl, rg, rr, rcur) = workload
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    xl, xg, xr, cur = RGF(16, precision="c128-3m", retarded_offdiag=True).selected_solve(
        a, sl, sg, return_retarded=True, return_current=True)
    for g, o in ((xl, rl), (xg, rg), (xr, rr)):
        assert _per_block_worst(w, g, o) < 1e-10
    assert rel_err(cur.cpu().numpy(), rcur) < 1e-10


@pytest.mark.parametrize("precision", ["c128", "c128-3m"])
def test_deep_recursion_parity(device, precision):
    """16 blocks of 128 at eta = 1e-3 (the benchmark's distributions): the
    left-connected G^R grows like 1 / eta, so the Hermitian rounding part of
    each x^< product grows by ~|x^R|^2 per block.  It is harmless while it
    stays Hermitian (the final (X - X^H) / 2 drops it) but not if any step
    mixes it into the skew-Hermitian solution; this depth and block size
    expose such a step (a product with mirrored upper tiles reached relative
    error ~10 here), which the shallow cases above do not."""
    w = make_workload(seed=2, n_energies=2, n_blocks=16, block_size=128)
    oa, osl, osg = _oracle_containers(w)
    rl, rg, rr = orgf.RGF(2).selected_solve(oa, osl, osg, return_retarded=True)
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    xl, xg, xr = RGF(2, precision=precision).selected_solve(a, sl, sg, return_retarded=True)
    for g, o in ((xl, rl), (xg, rg), (xr, rr)):
        assert _per_block_worst(w, g, o) < 1e-10
    # selected_inv on the same system (G^R only)
    xi = RGF(2, precision=precision).selected_inv(a)
    assert _per_block_worst(w, xi, orgf.RGF(2).selected_inv(oa)) < 1e-10


def test_deep_recursion_with_obc_blocks(device):
    """The deep case with boundary self-energies on the first and last
    blocks (retarded, and skew-Hermitian lesser / greater as the contact
    self-energies are) and every optional output, against the oracle."""
    w = make_workload(seed=4, n_energies=2, n_blocks=12, block_size=128)
    oa, osl, osg = _oracle_containers(w)
    n, ne, b = w.n_blocks, 2, 128
    rng = np.random.default_rng(9)
    def block(scale):
        return scale * (rng.standard_normal((ne, b, b)) + 1j * rng.standard_normal((ne, b, b)))
    gam0, gam1 = block(0.05), block(0.05)
    obc = {"retarded": {0: block(0.05), n - 1: block(0.05)},
           "lesser": {0: 1j * (gam0 @ gam0.conj().transpose(0, 2, 1))},
           "greater": {n - 1: -1j * (gam1 @ gam1.conj().transpose(0, 2, 1))}}
    oobc, gobc = orgf.OBCBlocks(n), OBCBlocks(n)
    for kind, blocks in obc.items():
        for i, v in blocks.items():
            getattr(oobc, kind)[i] = v
            getattr(gobc, kind)[i] = torch.as_tensor(v, device=device)
    rl, rg, rr, rcur = orgf.RGF(2).selected_solve(
        oa, osl, osg, obc_blocks=oobc, return_retarded=True, return_current=True,
        retarded_offdiag=True)
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    xl, xg, xr, cur = RGF(2, retarded_offdiag=True).selected_solve(
        a, sl, sg, obc_blocks=gobc, return_retarded=True, return_current=True)
    for g, o in ((xl, rl), (xg, rg), (xr, rr)):
        assert _per_block_worst(w, g, o) < 1e-10
    assert rel_err(cur.cpu().numpy(), rcur) < 1e-10
    # selected_inv with the retarded boundary blocks
    xi = RGF(2).selected_inv(a, obc_blocks=gobc)
    assert _per_block_worst(w, xi, orgf.RGF(2).selected_inv(oa, obc_blocks=oobc)) < 1e-10


@pytest.mark.parametrize("precision", ["c128", "c128-3m"])
def test_small_blocks_every_output(device, precision):
    """Blocks of mixed sizes up to 32 (the fused kernel's range, on the fused
    path), Σ^< with off-diagonal blocks and Σ^> block-diagonal (only one of
    the two chains has the Σ_ij terms), boundary blocks that differ between
    the chains (retarded at both ends, lesser at the first block, greater at
    the last), the current, the off-diagonal G^R blocks, the corner
    G^R_{0,N-1} and the density, against the oracle."""
    bs = np.array([8, 32, 16, 24, 32, 12, 32])
    n, ne = len(bs), 5
    off = np.hstack(([0], np.cumsum(bs)))
    rng = np.random.default_rng(17)
    oa = ODSDBCOO.from_sparray(sps.coo_matrix(bt_dense(bs, seed=21)), bs, (ne,))
    oa.data = oa.data * (1.0 + 0.05 * np.arange(ne))[:, None]
    s = bt_dense(bs, seed=22)
    d = bt_dense(bs, seed=23)
    mask = np.zeros(d.shape, dtype=bool)
    for i in range(n):
        mask[off[i]:off[i + 1], off[i]:off[i + 1]] = True
    d = np.where(mask, d, 0)
    osl = ODSDBCOO.from_sparray(sps.coo_matrix(s - s.conj().T), bs, (ne,))
    osg = ODSDBCOO.from_sparray(sps.coo_matrix(d - d.conj().T), bs, (ne,))
    assert osg.data.shape[-1] < osl.data.shape[-1]  # no off-diagonal Σ^> blocks

    def block(b, scale=0.05):
        return scale * (rng.standard_normal((ne, b, b)) + 1j * rng.standard_normal((ne, b, b)))
    g0, g1 = block(bs[0]), block(bs[-1])
    obc = {"retarded": {0: block(bs[0]), n - 1: block(bs[-1])},
           "lesser": {0: 1j * (g0 @ g0.conj().transpose(0, 2, 1))},
           "greater": {n - 1: -1j * (g1 @ g1.conj().transpose(0, 2, 1))}}
    oobc, gobc = orgf.OBCBlocks(n), OBCBlocks(n)
    for kind, blocks in obc.items():
        for i, v in blocks.items():
            getattr(oobc, kind)[i] = v
            getattr(gobc, kind)[i] = torch.as_tensor(v, device=device)
    ocorner = np.zeros((ne, bs[0], bs[-1]), dtype=np.complex128)
    rl, rg, rr, rcur = orgf.RGF(ne).selected_solve(
        oa, osl, osg, obc_blocks=oobc, return_retarded=True, return_current=True,
        retarded_offdiag=True, corner=ocorner)
    corner = torch.zeros((ne, bs[0], bs[-1]), dtype=torch.complex128, device=device)
    dens = torch.zeros((ne, int(off[-1])), dtype=torch.float64, device=device)
    a, sl, sg = (to_gpu(o, device) for o in (oa, osl, osg))
    xl, xg, xr, cur = RGF(ne, precision=precision, retarded_offdiag=True).selected_solve(
        a, sl, sg, obc_blocks=gobc, return_retarded=True, return_current=True, corner=corner,
        density=dens)
    tol = 1e-10 if precision == "c128" else 1e-9
    for g, o in ((xl, rl), (xg, rg), (xr, rr)):
        for i in range(n):
            for j in (i - 1, i, i + 1):
                if 0 <= j < n:
                    assert rel_err(g.blocks[i, j].cpu().numpy(), o.blocks[i, j]) < tol, (i, j)
    assert rel_err(cur.cpu().numpy(), rcur) < tol
    assert rel_err(corner.cpu().numpy(), ocorner) < tol
    from oracle.observables import density as o_density
    assert rel_err(dens.cpu().numpy(), o_density(rl)) < tol
    # the retarded inverse with the same boundary blocks
    xi = RGF(ne, precision=precision).selected_inv(a, obc_blocks=gobc)
    ri = orgf.RGF(ne).selected_inv(oa, obc_blocks=oobc)
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                assert rel_err(xi.blocks[i, j].cpu().numpy(), ri.blocks[i, j]) < tol, (i, j)


def test_sub_batch_offset_changes_only_timing(device, monkeypatch):
    """The odd sub-batches start after sub-batch 0's first block inverse
    (QX_RGF_STAGGER); the energies are independent, so the outputs are
    bitwise those of the sub-batches started together."""
    w = make_workload(seed=5, n_energies=16, n_blocks=6, block_size=64)
    oa, osl, osg = _oracle_containers(w)
    outs = {}
    for stg in ("1", "0"):
        monkeypatch.setenv("QX_RGF_STAGGER", stg)
        a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
        xl, xg, xr = RGF(16).selected_solve(a, sl, sg, return_retarded=True)
        outs[stg] = [t.data.cpu().numpy() for t in (xl, xg, xr)]
    for x1, x0 in zip(outs["1"], outs["0"]):
        assert np.array_equal(x1.view(np.uint64), x0.view(np.uint64))


@pytest.mark.parametrize("precision", ["c128", "c128-3m", "mixed"])
@pytest.mark.parametrize("sigma_pattern", ["shared", "distinct"])
def test_paired_lesser_greater_products_bitwise(device, monkeypatch, precision, sigma_pattern):
    """Each G^< product and its G^> counterpart run as one launch
    (QX_RGF_PAIR); every element is computed as by the separate launches, so
    the outputs are bitwise those of QX_RGF_PAIR=0.  Mixed block sizes above
    the fused-kernel limit, Σ with off-diagonal blocks, boundary blocks on
    both ends, the current and the off-diagonal G^R blocks; complex128 also
    against the oracle.  Σ^< and Σ^> share one sparsity pattern (their
    products pair) or not (different nnz layouts: two launches each)."""
    bs = np.array([48, 64, 64, 40, 64])
    n = len(bs)
    dense = bt_dense(bs, seed=3)
    coo = sps.coo_matrix(dense)
    oa = ODSDBCOO.from_sparray(coo, bs, (6,))
    rng = np.random.default_rng(11)
    s = sps.coo_matrix(bt_dense(bs, seed=4))
    s = sps.coo_matrix(s - s.conj().T)
    if sigma_pattern == "shared":
        vals = rng.standard_normal(s.nnz) + 1j * rng.standard_normal(s.nnz)
        s2 = sps.coo_matrix((vals, (s.row, s.col)), shape=s.shape)
    else:
        s2 = sps.coo_matrix(bt_dense(bs, seed=5))
        s2 = sps.coo_matrix(s2 - s2.conj().T)
    sig = [ODSDBCOO.from_sparray(m, bs, (6,)) for m in (s, s2)]
    assert (sig[0].data.shape == sig[1].data.shape) == (sigma_pattern == "shared")
    def block(b, scale=0.05):
        return scale * (rng.standard_normal((6, b, b)) + 1j * rng.standard_normal((6, b, b)))
    g0, g1 = block(bs[0]), block(bs[-1])
    obc = {"retarded": {0: block(bs[0]), n - 1: block(bs[-1])},
           "lesser": {0: 1j * (g0 @ g0.conj().transpose(0, 2, 1))},
           "greater": {n - 1: -1j * (g1 @ g1.conj().transpose(0, 2, 1))}}
    oobc, gobc = orgf.OBCBlocks(n), OBCBlocks(n)
    for kind, blocks in obc.items():
        for i, v in blocks.items():
            getattr(oobc, kind)[i] = v
            getattr(gobc, kind)[i] = torch.as_tensor(v, device=device)
    outs = {}
    for pair in ("1", "0"):
        monkeypatch.setenv("QX_RGF_PAIR", pair)
        a, sl, sg = (to_gpu(o, device) for o in (oa, *sig))
        res = RGF(4, precision=precision, retarded_offdiag=True).selected_solve(
            a, sl, sg, obc_blocks=gobc, return_retarded=True, return_current=True)
        outs[pair] = [t.data.cpu().numpy() if hasattr(t, "data") and not torch.is_tensor(t)
                      else t.cpu().numpy() for t in res]
    for x1, x0 in zip(outs["1"], outs["0"]):
        assert np.array_equal(x1.view(np.uint64), x0.view(np.uint64))
    if precision == "c128":
        refs = orgf.RGF(4).selected_solve(oa, *sig, obc_blocks=oobc, return_retarded=True,
                                          return_current=True, retarded_offdiag=True)
        for got, ref in zip(outs["1"][:3], refs[:3]):
            scale = np.max(np.abs(ref.data))
            assert np.allclose(got, ref.data, rtol=1e-10, atol=1e-10 * scale)
        assert rel_err(outs["1"][3], refs[3]) < 1e-10


@pytest.mark.parametrize("precision", ["c128", "c128-3m", "mixed"])
def test_paired_selected_inv_bitwise(device, monkeypatch, precision):
    """selected_inv runs x_jj a_ji with x_ii a_ij, and the two products that
    complete x_ji and x_ij, as one launch each where neighbouring blocks are
    of equal size (QX_RGF_PAIR); the outputs are bitwise those of
    QX_RGF_PAIR=0.  Mixed sizes (pairs and single launches), retarded
    boundary blocks; complex128 also against the oracle."""
    bs = np.array([48, 64, 64, 40, 40, 64])
    n = len(bs)
    oa = ODSDBCOO.from_sparray(sps.coo_matrix(bt_dense(bs, seed=6)), bs, (5,))
    rng = np.random.default_rng(12)
    obc = {0: 0.05 * (rng.standard_normal((5, bs[0], bs[0])) + 1j * rng.standard_normal((5, bs[0], bs[0]))),
           n - 1: 0.05 * (rng.standard_normal((5, bs[-1], bs[-1]))
                          + 1j * rng.standard_normal((5, bs[-1], bs[-1])))}
    oobc, gobc = orgf.OBCBlocks(n), OBCBlocks(n)
    for i, v in obc.items():
        oobc.retarded[i] = v
        gobc.retarded[i] = torch.as_tensor(v, device=device)
    outs = {}
    for pair in ("1", "0"):
        monkeypatch.setenv("QX_RGF_PAIR", pair)
        x = RGF(5, precision=precision).selected_inv(to_gpu(oa, device), obc_blocks=gobc)
        outs[pair] = x.data.cpu().numpy()
    assert np.array_equal(outs["1"].view(np.uint64), outs["0"].view(np.uint64))
    if precision == "c128":
        ref = orgf.RGF(5).selected_inv(oa, obc_blocks=oobc)
        scale = np.max(np.abs(ref.data))
        assert np.allclose(outs["1"], ref.data, rtol=1e-10, atol=1e-10 * scale)


def test_workload_device_assembly_matches(device, workload):
    """Workload.assemble (device, dense-block layout) equals the oracle containers."""
    w, (oa, osl, osg), _ = workload
    a, sl, sg = w.assemble(device)
    for g, o in ((a, oa), (sl, osl), (sg, osg)):
        assert np.array_equal(g.rows, o.rows) and np.array_equal(g.cols, o.cols)
        assert np.array_equal(g.data.cpu().numpy(), o.data)


def test_workload_c64_accuracy(device, workload):
    w, (oa, osl, osg), (rl, rg, rr, _) = workload
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    xl, xg, xr = RGF(16, precision="c64").selected_solve(a, sl, sg, return_retarded=True)
    for g, o in ((xl, rl), (xg, rg)):
        assert _per_block_worst(w, g, o) < 1e-4
    for i in range(w.n_blocks):
        assert rel_err(xr.blocks[i, i].cpu().numpy(), rr.blocks[i, i]) < 1e-4


def test_workload_mixed_accuracy(device, workload):
    """Mixed mode (complex64 products, FP64 Newton-Schulz on the inverses)."""
    w, (oa, osl, osg), (rl, rg, rr, _) = workload
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    xl, xg, xr = RGF(16, precision="mixed").selected_solve(a, sl, sg, return_retarded=True)
    for g, o in ((xl, rl), (xg, rg)):
        assert _per_block_worst(w, g, o) < 1e-4
    dens = np.stack([np.diagonal(xl.blocks[i, i].cpu().numpy(), axis1=-2, axis2=-1)
                     for i in range(w.n_blocks)])
    dens_ref = np.stack([np.diagonal(rl.blocks[i, i], axis1=-2, axis2=-1)
                         for i in range(w.n_blocks)])
    assert rel_err(dens, dens_ref) < 1e-5


@pytest.mark.parametrize("inverse", ["fp64", "c64"])
def test_mixed_newton_schulz_single_block(device, monkeypatch, inverse):
    """With one block the mixed mode is one block inverse: in FP64 by
    default, or (QX_MIXED_INVERSE=c64) one complex64 inverse plus one FP64
    Newton-Schulz step.  Either must be FP64-accurate, where plain complex64
    stays at FP32 accuracy."""
    if inverse == "c64":
        monkeypatch.setenv("QX_MIXED_INVERSE", "c64")
    rng = np.random.default_rng(3)
    n, ne = 96, 6
    m = (rng.standard_normal((ne, n, n)) + 1j * rng.standard_normal((ne, n, n))) / np.sqrt(n)
    m += 3.0 * np.eye(n)
    ref = np.linalg.inv(m)
    oa = ODSDBCOO.from_sparray(sps.coo_matrix(m[0]), [n], (ne,))
    oa.data[:] = m.reshape(ne, -1)
    errs = {}
    for prec in ("c64", "mixed"):
        x = RGF(precision=prec).selected_inv(to_gpu(oa, device))
        errs[prec] = rel_err(x.blocks[0, 0].cpu().numpy(), ref)
    assert errs["c64"] > 1e-9
    assert errs["mixed"] < 1e-12, errs


def test_obc_blocks_path(device, workload):
    w, (oa, osl, osg), _ = workload
    n, b, ne = w.n_blocks, w.block_sizes[0], oa.data.shape[0]
    rng = np.random.default_rng(7)
    obc = orgf.OBCBlocks(n)
    extra = 0.05 * (rng.standard_normal((ne, b, b)) + 1j * rng.standard_normal((ne, b, b)))
    obc.retarded[0] = extra
    obc.retarded[n - 1] = -1j * 0.01 * np.broadcast_to(np.eye(b), (ne, b, b))
    obc.lesser[0] = 1j * 0.1 * np.broadcast_to(np.eye(b), (ne, b, b))
    obc.greater[n - 1] = -1j * 0.1 * np.broadcast_to(np.eye(b), (ne, b, b))
    ref = orgf.RGF(16).selected_solve(oa, osl, osg, obc_blocks=obc, return_retarded=True)
    gobc = OBCBlocks(n)
    for name in ("retarded", "lesser", "greater"):
        for j, blk in enumerate(getattr(obc, name)):
            if blk is not None:
                getattr(gobc, name)[j] = torch.as_tensor(np.ascontiguousarray(blk), device=device)
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    got = RGF(16).selected_solve(a, sl, sg, obc_blocks=gobc, return_retarded=True)
    for g, o in zip(got, ref):
        assert _per_block_worst(w, g, o) < 1e-10


def test_selected_inv_workload_with_obc(device, workload):
    w, (oa, _, _), _ = workload
    n, b, ne = w.n_blocks, w.block_sizes[0], oa.data.shape[0]
    obc = orgf.OBCBlocks(n)
    obc.retarded[n - 1] = 0.02j * np.broadcast_to(np.eye(b), (ne, b, b)).copy()
    ref = orgf.RGF(16).selected_inv(oa, obc_blocks=obc)
    gobc = OBCBlocks(n)
    gobc.retarded[n - 1] = torch.as_tensor(obc.retarded[n - 1], device=device)
    x = RGF(16).selected_inv(to_gpu(oa, device), obc_blocks=gobc)
    assert _per_block_worst(w, x, ref) < 1e-10


def test_symmetric_output_containers(device):
    """Symmetric-stored outputs: lower blocks are not written (rgf.py:351-352)."""
    bs = BLOCK_SIZES[1]
    dense = bt_dense(bs)
    coo = sps.coo_matrix(dense)
    b = coo - coo.conj().T
    oa = ODSDBCOO.from_sparray(coo, bs, (3,))
    ob = ODSDBCOO.from_sparray(sps.coo_matrix(b), bs, (3,))
    osym = ODSDBCOO.from_sparray(sps.coo_matrix(b), bs, (3,), symmetry=True)
    oxl, oxg = ODSDBCOO.zeros_like(osym), ODSDBCOO.zeros_like(osym)
    orgf.RGF().selected_solve(oa, ob, ob, out=(oxl, oxg))
    a, bl = to_gpu(oa, device), to_gpu(ob, device)
    xl = DSDBCOO.zeros_like(to_gpu(osym, device))
    xg = DSDBCOO.zeros_like(to_gpu(osym, device))
    RGF().selected_solve(a, bl, bl, out=(xl, xg))
    assert close(xl, oxl, 1e-10) and close(xg, oxg, 1e-10)


def test_non_dense_blocks_staging(device):
    """Blocks with partial sparsity go through staging; results match the oracle."""
    bs = BLOCK_SIZES[0]
    dense = bt_dense(bs, seed=4)
    coo = sps.coo_matrix(dense)
    oa = ODSDBCOO.from_sparray(coo, bs, (4,))
    assert any(sl is not None and (sl.stop - sl.start) < 4
               for sl in (oa.block_slice(i, i) for i in range(len(bs))))
    ref = orgf.RGF().selected_inv(oa)
    x = RGF().selected_inv(to_gpu(oa, device))
    assert close(x, ref, 1e-10)


def test_corner_and_transmission(device, workload):
    from oracle.rgf import transmission as o_transmission
    from quatrex_amd.greens_function_solver import transmission

    w, (oa, osl, osg), _ = workload
    b0, bl, ne = w.block_sizes[0], w.block_sizes[-1], oa.data.shape[0]
    ref = np.zeros((ne, b0, bl), complex)
    orgf.RGF(16).selected_solve(oa, osl, osg, corner=ref)
    corner = torch.zeros((ne, b0, bl), dtype=torch.complex128, device=device)
    for mbs in (16, 3):
        RGF(mbs).selected_solve(to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device),
                                corner=corner)
        assert rel_err(corner.cpu().numpy(), ref) < 1e-10
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    t = transmission(corner, np.broadcast_to(gl, (ne, b0, b0)), np.broadcast_to(gr, (ne, bl, bl)))
    assert np.allclose(t.cpu().numpy(), o_transmission(ref, gl, gr), rtol=1e-10)


@pytest.mark.parametrize("precision", ["c128", "c128-3m"])
def test_fused_small_block_path_matches_general(device, monkeypatch, precision):
    """Stacks of blocks <= 32 run the whole sweep of one energy in one
    workgroup; its outputs (G^<, G^>, G^R with off-diagonal blocks, layer
    current, corner block, density) equal the per-product path's, with Sigma
    off-diagonal blocks, OBC blocks and mixed block sizes."""
    bs = np.array([4, 8, 8, 6, 8, 4, 8, 2, 8])
    dense = bt_dense(bs, seed=11)
    coo = sps.coo_matrix(dense)
    b = coo - coo.conj().T
    ne, n = 6, len(bs)
    oa = ODSDBCOO.from_sparray(coo, bs, (ne,))
    ob = ODSDBCOO.from_sparray(sps.coo_matrix(b), bs, (ne,))
    rng = np.random.default_rng(5)
    gobc = OBCBlocks(n)
    gobc.retarded[0] = torch.as_tensor(
        0.05 * (rng.standard_normal((ne, 4, 4)) + 1j * rng.standard_normal((ne, 4, 4))),
        device=device)
    gobc.lesser[0] = torch.as_tensor(0.1j * np.broadcast_to(np.eye(4), (ne, 4, 4)).copy(),
                                     device=device)
    gobc.greater[n - 1] = torch.as_tensor(-0.1j * np.broadcast_to(np.eye(8), (ne, 8, 8)).copy(),
                                          device=device)
    outs = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("QX_RGF_FUSED", fused)
        a, bl, bg = to_gpu(oa, device), to_gpu(ob, device), to_gpu(ob, device)
        corner = torch.zeros((ne, bs[0], bs[-1]), dtype=torch.complex128, device=device)
        dens = torch.zeros((ne, int(bs.sum())), dtype=torch.float64, device=device)
        xl, xg, xr, cur = RGF(100, precision=precision, retarded_offdiag=True).selected_solve(
            a, bl, bg, obc_blocks=gobc, return_retarded=True, return_current=True,
            corner=corner, density=dens)
        outs[fused] = [t.cpu().numpy() for t in (xl.data, xg.data, xr.data, cur, corner, dens)]
    for got, ref in zip(outs["1"], outs["0"]):
        scale = np.max(np.abs(ref))
        assert scale > 0
        assert np.max(np.abs(got - ref)) <= 1e-12 * scale


@pytest.mark.parametrize("block_sizes", [[7], [1, 1], [32, 5], [1, 32, 1], [16] * 16,
                                         [32] * 3 + [1]],
                         ids=["N1", "b1x2", "32-5", "1-32-1", "16x16", "32x3-1"])
def test_edge_block_structures(device, block_sizes):
    """Single blocks, 1 x 1 blocks, the largest fused block size and block
    count, against the oracle (selected_inv and selected_solve with the
    retarded output), on both solve paths."""
    bs = np.array(block_sizes)
    dense = bt_dense(bs, seed=len(bs) + int(bs.sum()))
    coo = sps.coo_matrix(dense)
    b = coo - coo.conj().T
    oa = ODSDBCOO.from_sparray(coo, bs, (3,))
    ob = ODSDBCOO.from_sparray(sps.coo_matrix(b), bs, (3,))
    ref_inv = orgf.RGF().selected_inv(oa)
    refs = orgf.RGF().selected_solve(oa, ob, ob, return_retarded=True)
    a, bl = to_gpu(oa, device), to_gpu(ob, device)
    assert close(RGF().selected_inv(a), ref_inv, 1e-10)
    got = RGF().selected_solve(a, bl, bl, return_retarded=True)
    for g, o in zip(got, refs):
        assert close(g, o, 1e-10)


@pytest.mark.parametrize("shape", [(16, 8, 32, 0), (2, 6, 96, 3)], ids=["8x32", "6x96"])
def test_ldos_against_oracle(device, shape):
    """LDOS -Im diag(G^R) / (2 pi) (quatrex/core/scba.py:572-576) written by
    the sweep, with and without the retarded output, against the oracle's
    diagonal G^R blocks."""
    ne, nb, b, seed = shape
    w = make_workload(seed=seed, n_energies=ne, n_blocks=nb, block_size=b)
    oa, osl, osg = _oracle_containers(w)
    rr = orgf.RGF(ne).selected_solve(oa, osl, osg, return_retarded=True)[2]
    ref = -np.concatenate([np.diagonal(rr.blocks[i, i], axis1=-2, axis2=-1).imag
                           for i in range(nb)], -1) / (2 * np.pi)
    a, sl, sg = to_gpu(oa, device), to_gpu(osl, device), to_gpu(osg, device)
    for ret in (True, False):
        ldos = torch.zeros((ne, nb * b), dtype=torch.float64, device=device)
        RGF(ne).selected_solve(a, sl, sg, return_retarded=ret, ldos=ldos)
        assert rel_err(ldos.cpu().numpy(), ref) < 1e-10
