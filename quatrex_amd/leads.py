"""Lead-block preparation and contact self-energies on the HIP path.

- ``periodic_superblocks(a_ii, a_ij, a_ji, block_sections)`` restates
  quatrex/core/utils.py:157-212 (``get_periodic_superblocks``): the lead's
  principal layer repeats the outermost sub-layer of the device's first
  block layer; returns ``(m_10, m_00, m_01)``.  Pure data movement on the
  tensors' device (the identity for ``block_sections == 1``).
- ``contact_obc_blocks(a, obc, left_occupancies, right_occupancies, ...)``
  restates quatrex/electron/solver.py:304-386 (``_compute_obc``): the left
  and right lead surface Green's functions from an OBC system (the right one
  through the reference's flips of both matrix axes), then per contact one
  ``qx_contact_self_energy`` call: Sigma^R_B = m_10 g m_01 on the MFMA GEMM
  and Sigma^<_B = i f Gamma, Sigma^>_B = i (f - 1) Gamma with
  Gamma = i (Sigma^R_B - Sigma^A_B) in one fused kernel.  The result is an
  ``OBCBlocks`` ready for ``RGF.selected_solve``.
"""

from __future__ import annotations

import numpy as np
import torch

from quatrex_amd import _lib
from quatrex_amd.greens_function_solver import OBCBlocks
from quatrex_amd.ops import as_device


def periodic_superblocks(a_ii, a_ij, a_ji, block_sections: int):
    """(m_10, m_00, m_01) of the periodic lead layer (core/utils.py:157-212)."""
    a_ii, a_ij, a_ji = (torch.as_tensor(x) for x in (a_ii, a_ij, a_ji))
    s = int(block_sections)
    n = a_ii.shape[-1]
    if s < 1 or n % s:
        raise ValueError("The array shape is not divisible by the number of blocks.")
    sub = n // s
    row = torch.cat((a_ii, a_ij), -1)[..., :sub, :]          # first sub-row, 2n wide
    col = torch.cat((a_ii, a_ji), -2)[..., :, :sub]          # first sub-column, 2n tall
    row_blocks = [row[..., :, c * sub:(c + 1) * sub] for c in range(2 * s)]
    col_blocks = [col[..., r * sub:(r + 1) * sub, :] for r in range(2 * s)]
    layer = col_blocks[s::-1] + row_blocks[1:]               # 3 s sub-blocks
    lay = torch.stack(layer, 0)                              # (3s, *stack, sub, sub)
    idx = (torch.arange(3 * s)[None, :] - torch.arange(s)[:, None]) % (3 * s)
    blocks = lay[idx.to(lay.device)]                         # (s, 3s, *stack, sub, sub)
    nd = blocks.dim()
    # (s, 3s, *stack, sub, sub) -> (*stack, s, sub, 3s, sub) -> (*stack, n, 3n)
    perm = tuple(range(2, nd - 2)) + (0, nd - 2, 1, nd - 1)
    full = blocks.permute(perm).reshape(a_ii.shape[:-2] + (n, 3 * n))
    return (full[..., :, :n].contiguous(), full[..., :, n:2 * n].contiguous(),
            full[..., :, 2 * n:].contiguous())


def contact_self_energy(m10, g, m01, occupancies):
    """(Sigma^R_B, Sigma^<_B, Sigma^>_B) of one lead (electron/solver.py:324-337)."""
    g = as_device(g)
    dev = g.device
    m10, m01 = as_device(m10, dev), as_device(m01, dev)
    b = g.shape[-1]
    lead = g.shape[:-2]
    n = int(np.prod(lead, dtype=np.int64))
    f = torch.as_tensor(np.asarray(occupancies, dtype=np.float64) if not isinstance(
        occupancies, torch.Tensor) else occupancies, dtype=torch.float64, device=dev)
    f = f.reshape(f.shape + (1,) * (len(lead) - f.dim())).expand(lead).contiguous()
    m10 = m10.expand(lead + m10.shape[-2:]).contiguous()
    m01 = m01.expand(lead + m01.shape[-2:]).contiguous()
    sr, sl, sg = (torch.empty(lead + (b, b), dtype=torch.complex128, device=dev)
                  for _ in range(3))
    h, st = _lib.handle(dev), _lib.stream_ptr(dev)
    rc = _lib.load().qx_contact_self_energy(h, st, n, b, m10.data_ptr(), g.data_ptr(),
                                            m01.data_ptr(), f.data_ptr(), sr.data_ptr(),
                                            sl.data_ptr(), sg.data_ptr())
    _lib.check(rc, h, "contact_self_energy")
    return sr, sl, sg


def contact_obc_blocks(a, obc, left_occupancies, right_occupancies, block_sections: int = 1,
                       overlap: dict | None = None, eta_obc: float = 0.0) -> OBCBlocks:
    """OBCBlocks of both contacts of the system matrix ``a`` (electron/solver.py:304-386).

    ``obc`` is an OBC system, ``obc((a_ii, a_ij, a_ji), contact) -> (x, ...)``
    (``quatrex_amd.boundary_conditions.OBCSystem``); ``overlap`` optionally
    maps the block keys (0, 0), (0, 1), (1, 0), (-1, -1), (-1, -2), (-2, -1)
    to overlap blocks, scaled by i * eta_obc as the reference does.
    """
    n_blocks = a.num_blocks
    out = OBCBlocks(n_blocks)
    flip = lambda x: torch.flip(x, dims=(-2, -1))
    dev = a.device

    def s_blk(key):
        if overlap is None or key not in overlap:
            return 0.0
        return 1j * eta_obc * torch.as_tensor(overlap[key], device=dev)

    m_10, m_00, m_01 = periodic_superblocks(a.blocks[0, 0], a.blocks[0, 1], a.blocks[1, 0],
                                            block_sections)
    g_00 = obc((m_00 + s_blk((0, 0)), m_01 + s_blk((0, 1)), m_10 + s_blk((1, 0))), "left")[0]
    out.retarded[0], out.lesser[0], out.greater[0] = contact_self_energy(
        m_10, g_00, m_01, left_occupancies)

    n, m = n_blocks - 1, n_blocks - 2
    m_mn, m_nn, m_nm = periodic_superblocks(flip(a.blocks[n, n]), flip(a.blocks[n, m]),
                                            flip(a.blocks[m, n]), block_sections)
    m_nn, m_nm, m_mn = flip(m_nn), flip(m_nm), flip(m_mn)
    g_nn = obc((flip(m_nn + s_blk((-1, -1))).contiguous(),
                flip(m_nm + s_blk((-1, -2))).contiguous(),
                flip(m_mn + s_blk((-2, -1))).contiguous()), "right")[0]
    g_nn = flip(torch.as_tensor(g_nn)).contiguous()
    out.retarded[n], out.lesser[n], out.greater[n] = contact_self_energy(
        m_mn, g_nn, m_nm, right_occupancies)
    return out
