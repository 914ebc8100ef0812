"""quatrex_amd: MI355X-native batched RGF + open-boundary hot path.

Drop-in replacements for the reference's solver API (qttools):

- ``quatrex_amd.datastructures.DSDBCOO``          block-sorted COO stacks
- ``quatrex_amd.greens_function_solver.RGF``      selected_inv / selected_solve
- ``quatrex_amd.greens_function_solver.Inv``      dense reference solver (small systems)
- ``quatrex_amd.boundary_conditions.SanchoRubio`` / ``Spectral`` / ``OBCSystem``
- ``quatrex_amd.nevp.Beyn`` / ``Full``

All numerical work runs in ``libquatrex_amd.so`` (hand-written gfx950 HIP
kernels behind the C ABI of ``include/quatrex_amd.h``); torch provides device
memory and streams.  Build the library with ``python -m quatrex_amd.build``.
"""

import torch  # noqa: F401  (binds the HIP runtime before libquatrex_amd loads)

__version__ = "0.1.0"
