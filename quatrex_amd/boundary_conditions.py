"""Open-boundary (lead) solvers on the HIP path.

Drop-ins for src/qttools/boundary_conditions:

- ``OBCSolver``    abstract solver, ``__call__(a_ii, a_ij, a_ji, contact)``
                   (obc/obc.py:11-49)
- ``SanchoRubio``  decimation surface Green's function (obc/sancho_rubio.py:10-86),
                   one C-ABI call (qx_sancho_rubio); non-convergence is a
                   ``RuntimeWarning`` as in the reference
- ``OBCSystem``    memoizing wrapper with fixed-point refinement, residuals and
                   high-residual warnings (obc/obc.py:52-202 on
                   boundary_system.py:48-373, identity system reducer)
- ``Spectral``     surface Green's function from NEVP eigenmodes
                   (obc/spectral.py:15-526)

Inputs are stacks (..., b, b) of torch HIP tensors (NumPy arrays are
uploaded); outputs are complex128 HIP tensors.  Block products and inverses
run in libquatrex_amd; the NEVP's small dense eigen/SVD problems and the
mode-filtering bookkeeping of ``Spectral`` run on the host in float64, as in
the reference's default configuration (``eig_compute_location="numpy"``).
"""

from __future__ import annotations

import ctypes
import warnings
from abc import ABC, abstractmethod

import numpy as np
import torch

from quatrex_amd import _lib
from quatrex_amd.ops import as_device, dagger, gemm, inv, mm, spectral_modes


class OBCSolver(ABC):
    """x_ii = (a_ii - a_ji x_ii a_ij)^{-1} (obc.py:11-49)."""

    @abstractmethod
    def __call__(self, a_ii, a_ij, a_ji, contact: str):
        ...


def _batched3(*xs):
    """Upload, flatten to (n, b, b); returns tensors, batch shape, squeeze flag."""
    ts = [as_device(x) for x in xs]
    ts = [as_device(t, ts[0].device) for t in ts]
    squeeze = ts[0].dim() == 2
    shape = ts[0].shape[:-2]
    b = ts[0].shape[-1]
    return [t.reshape(-1, b, b).contiguous() for t in ts], shape, squeeze


class SanchoRubio(OBCSolver):
    """Sancho-Rubio decimation (sancho_rubio.py:10-86).

    Stop criterion (batch-global, as the reference):
    max over the stack of || |alpha| + |beta| ||_F / 2 < convergence_tol.
    """

    def __init__(self, max_iterations: int = 100, convergence_tol: float = 1e-6):
        self.max_iterations = max_iterations
        self.convergence_tol = convergence_tol
        self.iterations = 0

    def __call__(self, a_ii, a_ij, a_ji, contact: str = ""):
        (aii, aij, aji), shape, _ = _batched3(a_ii, a_ij, a_ji)
        n, b = aii.shape[0], aii.shape[-1]
        x = torch.empty_like(aii)
        it = ctypes.c_int(0)
        h, st = _lib.handle(aii.device), _lib.stream_ptr(aii.device)
        rc = _lib.load().qx_sancho_rubio(h, st, n, b, aii.data_ptr(), aij.data_ptr(),
                                         aji.data_ptr(), x.data_ptr(), self.max_iterations,
                                         float(self.convergence_tol), ctypes.byref(it))
        _lib.check(rc, h, "SanchoRubio")
        _lib.sync_check(aii.device, "SanchoRubio")
        self.iterations = it.value
        if rc == _lib.QX_NOT_CONVERGED:
            warnings.warn("Surface Green's function did not converge.", RuntimeWarning)
        return x.reshape(shape + (b, b))


class IdentityReducer:
    """System reducer that does not reduce (system_reduction.py:101-180)."""

    def contract_system(self, boundary_system):
        return boundary_system

    def expand_solution(self, boundary_system, reduced_system, reduced_solution):
        return reduced_solution

    def expand_residuals(self, boundary_system, reduced_system, rel, abs_):
        return rel, abs_


class IdentityCompressor:
    """Cache compressor that copies (boundary_system.py:11-45)."""

    def compress(self, x):
        return x.clone() if isinstance(x, torch.Tensor) else x

    def decompress(self, x):
        return x


class OBCSystem:
    """Memoizing OBC wrapper (obc.py:52-202 / boundary_system.py:48-373).

    ``__call__((a_ii, a_ij, a_ji), contact) -> (x, rel_residuals, abs_residuals)``.
    Memoization modes: "off", "auto", "force-after-first", "force".  In
    "auto" mode the reference compares the number of stack entries whose
    cached solution passes the tolerances, divided by the stack communicator
    size (here one process), with ``agreement_threshold`` and re-solves from
    scratch when it is exceeded (boundary_system.py:280-296); this is
    reproduced as written.
    """

    def __init__(self, boundary_solver, cache_compressor=None, system_reducer=None,
                 num_ref_iterations: int = 2, relative_tol: float = 2e-1,
                 absolute_tol: float = 1e-6, warning_threshold: float = 1e-1,
                 memoization_mode: str = "auto", agreement_threshold: float = 0.999):
        self.boundary_solver = boundary_solver
        self.num_ref_iterations = num_ref_iterations
        self.relative_tol = relative_tol
        self.absolute_tol = absolute_tol
        self.warning_threshold = warning_threshold
        self.memoization_mode = memoization_mode
        self.agreement_threshold = agreement_threshold
        self.cache_compressor = cache_compressor or IdentityCompressor()
        for m in ("compress", "decompress"):
            if not hasattr(self.cache_compressor, m):
                raise ValueError(f"Cache compressor must have a '{m}' method.")
        if not (0.0 <= self.agreement_threshold <= 1.0):
            raise ValueError("Memoizing tolerance must be between 0 and 1.")
        self.system_reducer = system_reducer or IdentityReducer()
        for m in ("contract_system", "expand_solution", "expand_residuals"):
            if not hasattr(self.system_reducer, m):
                raise ValueError(f"System reducer must have an '{m}' method.")
        self._cache: dict = {}
        if self.num_ref_iterations < 2:
            warnings.warn("The number of refinement iterations should be at least 2. "
                          "Defaulting to 2.", RuntimeWarning)
            self.num_ref_iterations = 2

    @staticmethod
    def _fixed_point_step(system, x):
        a_ii, a_ij, a_ji = system
        t = mm(a_ji, x, a_ij)
        return inv(a_ii - t)

    @staticmethod
    def _get_starting_guess(system):
        return inv(system[0])

    def _get_residuals(self, system, x):
        ref = self._fixed_point_step(system, x)
        abs_res = torch.linalg.vector_norm(ref - x, dim=(-2, -1))
        rel_res = abs_res / torch.linalg.vector_norm(ref, dim=(-2, -1))
        return rel_res, abs_res, ref

    def _solve(self, system, contact, **kwargs):
        x = self.boundary_solver(*system, contact, **kwargs)
        if self.memoization_mode != "off":
            self._cache[contact] = self.cache_compressor.compress(x)
        return x

    def _memoized_solve(self, system, contact, **kwargs):
        if self.memoization_mode == "off":
            return self._solve(system, contact, **kwargs)
        x = self.cache_compressor.decompress(self._cache.get(contact, None))
        if x is None:
            if self.memoization_mode in ("auto", "force-after-first"):
                return self._solve(system, contact, **kwargs)
            elif self.memoization_mode == "force":
                x = self._get_starting_guess(system)
            else:
                raise RuntimeError(f"Invalid memoizing mode: {self.memoization_mode}")
        rel, abs_, x = self._get_residuals(system, x)
        if self.memoization_mode == "auto":
            local = int(torch.sum((abs_ < self.absolute_tol) | (rel < self.relative_tol)))
            if local / 1 > self.agreement_threshold:
                return self._solve(system, contact, **kwargs)
        for _ in range(self.num_ref_iterations - 2):
            x = self._fixed_point_step(system, x)
        return x

    def __call__(self, boundary_system, contact: str, **kwargs):
        system = tuple(as_device(t) for t in boundary_system)
        reduced = self.system_reducer.contract_system(system)
        x = self._memoized_solve(reduced, contact, **kwargs)
        rel, abs_, x = self._get_residuals(reduced, x)
        if bool(torch.any((rel > self.warning_threshold) & (abs_ > self.absolute_tol))):
            warnings.warn(
                f"High error for {contact} of {self.boundary_solver.__class__.__name__}:\n"
                f"  Relative recursion error: {float(torch.max(rel)):.3e}\n"
                f"  Absolute recursion error: {float(torch.max(abs_)):.3e}\n",
                RuntimeWarning)
        if self.memoization_mode != "off":
            self._cache[contact] = self.cache_compressor.compress(x)
        sol = self.system_reducer.expand_solution(system, reduced, x)
        rel, abs_ = self.system_reducer.expand_residuals(system, reduced, rel, abs_)
        return sol, rel, abs_


def _block_view(arr: np.ndarray, axis: int, num_blocks: int) -> np.ndarray:
    shape = list(arr.shape)
    if shape[axis] % num_blocks != 0:
        raise ValueError("The array shape is not divisible by the number of blocks.")
    shape[axis] //= num_blocks
    strides = (arr.strides[axis] * shape[axis],) + arr.strides
    return np.lib.stride_tricks.as_strided(arr, shape=(num_blocks,) + tuple(shape), strides=strides)


class Spectral(OBCSolver):
    """Spectral OBC (spectral.py:15-526).

    Device-resident after the NEVP: mode upscaling, residuals, group
    velocities and the decay classification are device tensors (products on
    the MFMA GEMM), the per-energy pseudo-inverse of the reference's Python
    loop is one batched masked normal-equation solve, and the refinement
    inversions run on the GPU.  Beyn's small SVD/eig stays on the host, as
    the reference's default ``eig_compute_location="numpy"``.
    """

    def __init__(self, nevp, block_sections: int = 1, min_decay: float = 1e-3,
                 max_decay: float = 6.9, num_ref_iterations: int = 2,
                 min_propagation: float = 0.01, residual_tolerance: float = 1e-3,
                 residual_normalization: bool = True, warning_threshold: float = 1e-1,
                 eta_decay: float = 1e-14):
        self.nevp = nevp
        self.block_sections = block_sections
        self.min_decay, self.max_decay = min_decay, max_decay
        self.num_ref_iterations = num_ref_iterations
        self.min_propagation = min_propagation
        self.residual_tolerance = residual_tolerance
        self.residual_normalization = residual_normalization
        self.warning_threshold = warning_threshold
        self.eta_decay = eta_decay
        if self.num_ref_iterations < 1:
            raise ValueError("Number of refinement iterations must be at least 1.")

    def _extract_subblocks(self, a_ji, a_ii, a_ij):
        """spectral.py:97-158 (host arrays)."""
        layer = (a_ji, a_ii, a_ij)
        if self.block_sections == 1:
            return layer
        bs = self.block_sections
        host = [t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t) for t in layer]
        view = _block_view(np.concatenate(host, axis=-1), -1, 3 * bs)
        view = _block_view(view, -2, bs)
        first = np.linalg.norm(view[0, :])
        errs = [np.linalg.norm(view[0, :] - np.roll(view[i, :], -i, axis=0)) / first
                for i in range(1, bs)]
        if max(errs) > 1e-3:
            warnings.warn(f"Requested block sectioning is not periodic. ({max(errs):.2e})",
                          RuntimeWarning)
        blocks = view[0, :-bs + 1]
        idx = np.where([np.any(b) for b in blocks])[0]
        if idx.size == 0 or len(blocks) <= 3:
            return tuple(np.ascontiguousarray(b) for b in blocks)
        n_data = min(idx[0], len(blocks) - 1 - idx[-1])
        n = min(n_data, (len(blocks) - 3) // 2)
        sel = blocks[n:-n] if n > 0 else blocks
        return tuple(np.ascontiguousarray(b) for b in sel)

    def _upscale_eigenmodes(self, ws, vs):
        """spectral.py:322-377 on device tensors."""
        if self.block_sections == 1:
            return ws, vs / torch.linalg.norm(vs, dim=-2, keepdim=True)
        batch = vs.shape[:-2]
        sub, modes = vs.shape[-2], vs.shape[-1]
        wsu = torch.stack([ws ** n for n in range(self.block_sections)], dim=len(batch))
        vsu = (wsu[..., :, None, :] * vs[..., None, :, :]).reshape(
            *batch, sub * self.block_sections, modes)
        vsu = vsu / torch.linalg.norm(vsu, dim=-2, keepdim=True)
        return ws ** self.block_sections, vsu

    def _find_reflected_modes(self, ws, vs, a_xx):
        """spectral.py:198-320 (no injection vectors) on device tensors.

        The products a_n V run on the MFMA GEMM; residuals, group velocities
        (spectral.py:160-196) and the decay classification are one fused
        kernel over the eigenpairs (``ops.spectral_modes``).
        """
        if len(a_xx) != 3:
            raise ValueError("the Spectral OBC takes the three blocks (a_ji, a_ii, a_ij)")
        n = vs.shape[:-2]
        vs3 = vs.reshape(-1, *vs.shape[-2:]).contiguous()
        ws2 = ws.reshape(-1, ws.shape[-1]).contiguous()
        a_vs = [gemm(a.contiguous(), vs.contiguous()).reshape(vs3.shape) for a in a_xx]
        mask = spectral_modes(ws2, vs3, *a_vs, min_decay=self.min_decay,
                              max_decay=self.max_decay, min_propagation=self.min_propagation,
                              residual_tolerance=self.residual_tolerance,
                              eta_decay=self.eta_decay,
                              residual_normalization=self.residual_normalization)
        return mask.reshape(*n, ws.shape[-1])

    @staticmethod
    def _x_ii_a_ij(ws, vs, mask):
        """x_ii a_ij = V_r diag(1/w_r) V_r^+ for every energy at once
        (spectral.py:412-421).

        The selected modes of each energy are packed to the front (at most b
        of them); the padding columns are zero, and the identity on their
        diagonal of V^H V leaves the pseudo-inverse of the selected columns
        unchanged and its padding rows zero.
        """
        n, b = vs.shape[0], vs.shape[-2]
        k_max = int(mask.sum(-1).max().item()) if mask.numel() else 0
        if k_max == 0:
            return torch.zeros((n, b, b), dtype=vs.dtype, device=vs.device)
        order = torch.argsort(mask.to(torch.int8), dim=-1, descending=True, stable=True)
        idx = order[..., :k_max]
        sel = torch.gather(mask, -1, idx)
        vr = torch.gather(vs, -1, idx[..., None, :].expand(n, b, k_max))
        vr = torch.where(sel[..., None, :], vr, torch.zeros_like(vr)).contiguous()
        w = torch.gather(ws, -1, idx)
        w_inv = torch.where(sel, 1.0 / torch.where(sel, w, torch.ones_like(w)),
                            torch.zeros_like(w))
        gram = gemm(vr, vr, op_a="C") + torch.diag_embed((~sel).to(vs.dtype))
        pinv = gemm(inv(gram), vr, op_b="C")
        return gemm((vr * w_inv[..., None, :]).contiguous(), pinv)

    def __call__(self, a_ii, a_ij, a_ji, contact: str = ""):
        (aii, aij, aji), shape, squeeze = _batched3(a_ii, a_ij, a_ji)
        blocks = self._extract_subblocks(aji, aii, aij)
        ws, vs = self.nevp(blocks)
        dev = aii.device
        ws = torch.as_tensor(ws).to(dev, torch.complex128)
        vs = torch.as_tensor(vs).to(dev, torch.complex128)
        ws, vs = self._upscale_eigenmodes(ws, vs)
        mask = self._find_reflected_modes(ws, vs, [aji, aii, aij])
        b = aii.shape[-1]
        xa = self._x_ii_a_ij(ws, vs, mask)
        x = inv(aii + gemm(aji, xa))
        for _ in range(self.num_ref_iterations - 1):
            x = inv(aii - mm(aji, x, aij))
        x = inv(aii - mm(aji, x, aij))
        return x.reshape(shape + (b, b))
