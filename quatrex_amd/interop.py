"""Zero-copy bridges between qttools-style containers and quatrex_amd.

``from_qttools`` wraps any object exposing the DSDBCOO attributes of the
reference (``data``, ``rows``, ``cols``, ``block_sizes``,
``global_stack_shape``, ``symmetry``; src/qttools/datastructures/dsdbcoo.py)
as a ``quatrex_amd.datastructures.DSDBCOO`` sharing the same device memory
through DLPack (CuPy-on-ROCm or torch arrays); host (NumPy) data is uploaded
once.  ``to_array_module`` hands a device tensor back to the caller's array
library the same way.
"""

from __future__ import annotations

import numpy as np
import torch

from quatrex_amd.datastructures import DSDBCOO


def as_torch(x, device=None) -> torch.Tensor:
    """torch view of a device array (DLPack) or an uploaded copy of host data."""
    if isinstance(x, torch.Tensor):
        t = x
    elif hasattr(x, "__dlpack__") and not isinstance(x, np.ndarray):
        t = torch.from_dlpack(x)
    else:
        t = torch.as_tensor(np.ascontiguousarray(x))
    if t.device.type != "cuda":
        t = t.to(device or "cuda")
    return t


def from_qttools(mat, device=None) -> DSDBCOO:
    """Wrap a qttools DSDBCOO (single rank) without copying its device data."""
    data = as_torch(mat.data, device)
    rows = np.asarray(mat.rows.get() if hasattr(mat.rows, "get") else mat.rows)
    cols = np.asarray(mat.cols.get() if hasattr(mat.cols, "get") else mat.cols)
    bs = np.asarray(mat.block_sizes.get() if hasattr(mat.block_sizes, "get") else mat.block_sizes)
    return DSDBCOO(data, rows, cols, bs, tuple(mat.global_stack_shape),
                   symmetry=bool(getattr(mat, "symmetry", False)))


def to_array_module(t: torch.Tensor, xp):
    """Hand a torch device tensor to ``xp`` (e.g. cupy) through DLPack."""
    if hasattr(xp, "from_dlpack"):
        return xp.from_dlpack(t)
    return np.asarray(t.cpu())
