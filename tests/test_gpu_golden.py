"""The HIP path against the committed config-1 fixture (tests/golden/config1_rgf.npz):
density, transmission, layer currents, every block norm of G^<, G^>, G^R and
full diagonal G^< blocks.  Bar: 1e-10 relative (complex128)."""

from pathlib import Path

import numpy as np
import pytest
import torch

from quatrex_amd.greens_function_solver import RGF, transmission
from quatrex_amd.workloads import make_workload
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "config1_rgf.npz"


@pytest.mark.parametrize("precision,tol", [("c128", 1e-10), ("c128-3m", 1e-10)])
def test_hip_path_matches_fixture(device, precision, tol):
    ref = dict(np.load(GOLDEN, allow_pickle=False))
    w = make_workload(seed=0)
    a, sl, sg = w.assemble(device)
    n, ne = w.n_blocks, len(w.energies)
    dens = torch.zeros((ne, int(np.sum(w.block_sizes))), dtype=torch.float64, device=device)
    corner = torch.zeros((ne, w.block_sizes[0], w.block_sizes[-1]), dtype=torch.complex128,
                         device=device)
    xl, xg, xr, cur = RGF(16, precision=precision, retarded_offdiag=True).selected_solve(
        a, sl, sg, return_retarded=True, return_current=True, corner=corner, density=dens)
    gl = 1j * (w.sigma_b_left - w.sigma_b_left.conj().T)
    gr = 1j * (w.sigma_b_right - w.sigma_b_right.conj().T)
    t = transmission(corner, np.broadcast_to(gl, (ne,) + gl.shape),
                     np.broadcast_to(gr, (ne,) + gr.shape)).cpu().numpy()
    assert rel_err(dens.cpu().numpy(), ref["density"]) < tol
    assert rel_err(t, ref["transmission"]) < tol
    assert rel_err(cur.cpu().numpy(), ref["current"]) < tol
    for q, x in enumerate((xl, xg, xr)):
        norms = np.stack([torch.linalg.norm(x.blocks[int(i), int(j)], dim=(-2, -1)).cpu().numpy()
                          for i, j in ref["block_keys"]], -1)
        assert rel_err(norms, ref["block_norms"][q]) < tol
    for e, key in ((0, "lesser_diag_e0"), (15, "lesser_diag_e15")):
        got = np.stack([xl.blocks[i, i][e].cpu().numpy() for i in range(n)])
        assert rel_err(got, ref[key]) < tol
    assert rel_err(corner[0].cpu().numpy(), ref["corner_e0"]) < tol
