"""Contact self-energies on the HIP path against the oracle (electron/solver.py:304-386)."""

import numpy as np
import pytest
import torch

from oracle.leads import contact_obc_blocks as o_contact
from oracle.obc import OBCSystem as OOBCSystem, SanchoRubio as OSanchoRubio
from oracle.obc import boundary_self_energies
from quatrex_amd.boundary_conditions import OBCSystem, SanchoRubio
from quatrex_amd.datastructures import DSDBCOO
from quatrex_amd.leads import contact_obc_blocks, contact_self_energy
from tests.conftest import rel_err
from tests.test_oracle_leads import _device

pytestmark = pytest.mark.gpu


def test_contact_self_energy_kernel(device):
    rng = np.random.default_rng(0)
    ne, b = 5, 40
    c = lambda: rng.standard_normal((ne, b, b)) + 1j * rng.standard_normal((ne, b, b))
    m10, g, m01 = c(), c(), c()
    f = rng.uniform(0, 1, ne)
    ref = boundary_self_energies(g, m10, m01, f)
    got = contact_self_energy(m10, g, m01, f)
    for r, x in zip(ref, got):
        assert rel_err(x.cpu().numpy(), r) < 1e-13


@pytest.mark.parametrize("nb,b", [(4, 4), (5, 16)])
def test_contact_obc_blocks_match_oracle(device, nb, b):
    rng = np.random.default_rng(nb + b)
    oa, _ = _device(rng, nb, b, 6)
    a = DSDBCOO(torch.as_tensor(oa.data, device=device), oa.rows, oa.cols, oa.block_sizes,
                oa.global_stack_shape)
    fl, fr = rng.uniform(0, 1, 6), rng.uniform(0, 1, 6)
    ref = o_contact(oa, OOBCSystem(OSanchoRubio(max_iterations=500, convergence_tol=1e-12)),
                    fl, fr)
    got = contact_obc_blocks(a, OBCSystem(SanchoRubio(max_iterations=500,
                                                      convergence_tol=1e-12)), fl, fr)
    for name in ("retarded", "lesser", "greater"):
        for j in (0, nb - 1):
            r = getattr(ref, name)[j]
            g = getattr(got, name)[j].cpu().numpy()
            assert rel_err(g, r) < 1e-9, (name, j)
