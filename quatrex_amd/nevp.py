"""Non-linear eigenvalue solvers for the lead modes.

Drop-ins for src/qttools/nevp (``NEVP`` nevp.py:8-36, ``Beyn`` beyn.py:14-379,
``Full`` full.py:19-206) and src/qttools/kernels/operator.py
(``operator_inverse`` :113-179).

``Beyn`` computes the contour moments q0, q1 on the GPU (qx_beyn_moments:
T(z) assembled per node and inverted by the batched Gauss-Jordan kernel,
weights accumulated in place) and forms the probed moments P = q Y with the
MFMA GEMM.  The projection follows ``project_compute_location``: "device"
(the default here) runs the SVD on the device (one-sided Jacobi,
``qx_svd_batched``) and forms the projected matrices U^H P1 V S^-1 there, so
only those m_0 x m_0 matrices reach the host for the small eigenproblems
(``eig_compute_location = "numpy"``); "numpy" (the reference's default) runs
the SVD or QR on the host, as the reference does.  Both give the same
eigenpairs to rounding; at lead layer 256 the device projection takes the
Spectral OBC on Beyn from 90 / 212 / 613 ms to 58 / 145 / 442 ms per 32
energies at m_0 = 38 / 68 / 120.  ``use_qr=True`` always projects on the
host.  The host eigenproblems of Beyn and ``Full`` (LAPACK, one per matrix)
run on a pool of host threads with one BLAS thread each (``host_eig``), as
the reference's host path runs them in a parallel loop: 442 -> 114 ms per 32
energies at m_0 = 120.  The probe matrix Y comes from the
same persistent ``numpy.random.default_rng(42)`` stream as the reference,
so runs are reproducible call by call.
"""

from __future__ import annotations

import ctypes
import os
import threading
import warnings
from abc import ABC, abstractmethod
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.sparse as sps
import torch

from quatrex_amd import _lib
from quatrex_amd.ops import as_device, gemm, inv
from quatrex_amd.utils import get_section_sizes


_eig_pool: ThreadPoolExecutor | None = None
_eig_lock = threading.Lock()
EIG_THREADS_MAX = 16  # one GPU's share of the host cores


def host_eig(mats):
    """``numpy.linalg.eig`` of every matrix of ``mats`` (a list of square
    arrays or one (n, m, m) array), the matrices spread over a pool of host
    threads with one BLAS thread each, as the reference's host path runs one
    LAPACK eig per matrix in a parallel loop (kernels/linalg/eig.py:24-50,
    numba prange).  numpy's LAPACK call releases the GIL.  Returns the list
    of (w, v) per matrix."""
    global _eig_pool
    n = len(mats)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    workers = max(1, min(EIG_THREADS_MAX, cores))
    if workers == 1 or n <= 1:
        return [np.linalg.eig(mats[i]) for i in range(n)]
    from threadpoolctl import threadpool_limits

    with _eig_lock:  # one caller at a time: the BLAS thread limit is process-wide
        if _eig_pool is None:
            _eig_pool = ThreadPoolExecutor(max_workers=workers, thread_name_prefix="qx-eig")
        with threadpool_limits(limits=1):
            return list(_eig_pool.map(np.linalg.eig, [mats[i] for i in range(n)]))


class NEVP(ABC):
    """sum_n a_n w^n v = 0, n = -b..b (nevp.py:8-36)."""

    @abstractmethod
    def __call__(self, a_xx):
        ...


def operator_inverse(a_xx, z, contour_type=None, in_type=None, num_threads: int = 1024):
    """inv(sum_n z^n a_n) at every point z (operator.py:113-179).

    a_xx: tuple of (batch, 1, d, d) (or broadcastable) stacks, z: (1, nq, 1, 1).
    Returns (batch, nq, d, d) complex128 on the device.
    """
    b = len(a_xx) // 2
    dev = None
    ts = []
    for a in a_xx:
        t = as_device(a, dev)
        dev = t.device
        ts.append(t)
    zt = torch.as_tensor(np.asarray(z), dtype=torch.complex128, device=dev)
    op = sum(zt ** n * a for a, n in zip(ts, range(-b, b + 1)))
    return inv(op.contiguous())


class Beyn(NEVP):
    """Beyn's contour-integral NEVP solver (beyn.py:14-94)."""

    def __init__(self, r_o: float, r_i: float, m_0: int, num_quad_points: int,
                 num_threads_contour: int = 1024, eig_compute_location: str = "numpy",
                 project_compute_location: str = "device", use_qr: bool = False,
                 contour_batch_size: int | None = None, use_pinned_memory: bool = True):
        self.r_o, self.r_i, self.m_0 = r_o, r_i, m_0
        self.num_quad_points = num_quad_points
        self.num_threads_contour = num_threads_contour
        self.eig_compute_location = eig_compute_location
        self.project_compute_location = project_compute_location
        self.use_qr = use_qr
        if contour_batch_size is None:
            contour_batch_size = num_quad_points
        self.contour_batch_size = contour_batch_size
        counts, _ = get_section_sizes(num_quad_points,
                                      int(np.ceil(num_quad_points / contour_batch_size)))
        self.contour_displacements = np.cumsum(np.concatenate(([0], np.array(counts))))
        self.use_pinned_memory = use_pinned_memory
        self.rng = np.random.default_rng(42)

    def _contour_integrate(self, a_xx):
        """q0, q1 (beyn.py:177-235) on the device."""
        n, d = a_xx[0].shape[0], a_xx[0].shape[-1]
        dev = a_xx[0].device
        if len(a_xx) == 3:
            aji, aii, aij = (t.contiguous() for t in a_xx)
            q0 = torch.empty_like(aii)
            q1 = torch.empty_like(aii)
            h, st = _lib.handle(dev), _lib.stream_ptr(dev)
            rc = _lib.load().qx_beyn_moments(h, st, n, d, self.num_quad_points, float(self.r_o),
                                             float(self.r_i), aji.data_ptr(), aii.data_ptr(),
                                             aij.data_ptr(), q0.data_ptr(), q1.data_ptr(),
                                             _lib.PREC_C128)
            _lib.check(rc, h, "Beyn moments")
            _lib.sync_check(dev, "Beyn moments")
            return q0, q1
        # higher-order polynomial (subblocked leads): operator inverse per node
        zeta = np.arange(self.num_quad_points) + 1
        phase = np.exp(2j * np.pi * (zeta + 0.5) / self.num_quad_points)
        w = 1.0 / self.num_quad_points
        q0 = torch.zeros_like(a_xx[0])
        q1 = torch.zeros_like(a_xx[0])
        a4 = [a[:, None] for a in a_xx]
        dsp = self.contour_displacements
        for j in range(len(dsp) - 1):
            zo = (self.r_o * phase[dsp[j]:dsp[j + 1]]).reshape(1, -1, 1, 1)
            zi = (self.r_i * phase[dsp[j]:dsp[j + 1]]).reshape(1, -1, 1, 1)
            io = operator_inverse(a4, zo)
            ii = operator_inverse(a4, zi)
            zo_t = torch.as_tensor(zo, device=dev)
            zi_t = torch.as_tensor(zi, device=dev)
            q0 += torch.sum(w * zo_t * io - w * zi_t * ii, dim=1)
            q1 += torch.sum(w * zo_t ** 2 * io - w * zi_t ** 2 * ii, dim=1)
        return q0, q1

    def _project_svd(self, p0, p1):
        """beyn.py:96-145 (host)."""
        d = p0.shape[-1]
        u, s, vh = np.linalg.svd(p0, full_matrices=False)
        a, u_out = [], []
        for i in range(p0.shape[0]):
            ui, si, vhi = u[i], s[i], vh[i]
            eps = si.max() * d * np.finfo(p0.dtype).eps
            inds = np.where(si > eps)[0]
            ui, si, vhi = ui[:, inds], si[inds], vhi[inds, :]
            u_out.append(ui)
            a.append(ui.conj().T @ p1[i] @ vhi.conj().T / si)
        return a, u_out

    def _project_svd_device(self, p0, p1):
        """beyn.py:96-145 with the SVD and the projection on the device:
        returns the host projected matrices (each cut to the singular values
        above s_max m_0 eps) and the device projector U."""
        from quatrex_amd.ops import svd

        d = p0.shape[-1]
        u, s, v = svd(p0)
        eps = s[:, :1] * d * np.finfo(np.float64).eps
        keep = (s > eps).sum(-1).cpu().numpy()
        # U^H P1 V with column j divided by s_j
        a_dev = gemm(gemm(u, p1, op_a="C"), v) / s[:, None, :].to(torch.complex128)
        a_host = a_dev.cpu().numpy()
        return [a_host[i, :k, :k] for i, k in enumerate(keep)], u

    def _solve_reduced_device(self, a, y, u):
        """beyn.py:237-287 with the small eigenproblems on the host and the
        back-projection V = U v on the device."""
        n, d, m0 = y.shape
        ws = np.zeros((n, m0), dtype=np.complex128)
        vsmall = np.zeros((n, m0, m0), dtype=np.complex128)
        found = np.zeros(n, dtype=np.int64)
        for i, (w, v) in enumerate(host_eig(a)):
            k = len(w)
            ws[i, :k] = w
            vsmall[i, :k, :k] = v
            found[i] = k
        vs = gemm(u, torch.as_tensor(vsmall, device=u.device)).cpu().numpy()
        # slots without an eigenpair keep the probe vectors, as the reference
        cols = np.arange(m0)[None, None, :] >= found[:, None, None]
        vs = np.where(cols, y, vs)
        return ws, vs

    def _project_qr(self, p0, p1):
        """beyn.py:147-175 (host)."""
        q, r = np.linalg.qr(p0)
        return np.swapaxes(q.conj(), -2, -1) @ p1 @ np.linalg.inv(r), q

    def _solve_reduced_system(self, a, y, p_back):
        """beyn.py:237-287 (host)."""
        if isinstance(a, list):
            ws = np.zeros((len(a), self.m_0), dtype=np.complex128)
            vs = y.copy()
            for i, (w, v) in enumerate(host_eig(a)):
                ws[i, :len(w)] = w
                vs[i, :, :len(w)] = p_back[i] @ v
            return ws, vs
        pairs = host_eig(a)
        w = np.stack([p[0] for p in pairs])
        v = np.stack([p[1] for p in pairs])
        return w, p_back @ v

    def __call__(self, a_xx):
        """beyn.py:289-379."""
        a_xx = [as_device(a) for a in a_xx]
        if a_xx[0].dim() == 2:
            a_xx = [a[None] for a in a_xx]
        batch_shape = tuple(a_xx[0].shape[:-2])
        d = a_xx[0].shape[-1]
        a_xx = [a.reshape(-1, d, d).contiguous() for a in a_xx]
        old = None
        if d < self.m_0 and self.use_qr:
            warnings.warn(f"Subspace guess {self.m_0} is larger than the dimension of the "
                          f"system {d}. Setting subspace guess to {d}.")
            old, self.m_0 = self.m_0, d
        n = a_xx[0].shape[0]
        y = self.rng.random((n, d, self.m_0)) + 1j * self.rng.random((n, d, self.m_0))
        q0, q1 = self._contour_integrate(a_xx)
        yd = as_device(y, q0.device)
        p0d, p1d = gemm(q0, yd), gemm(q1, yd)
        if self.project_compute_location != "numpy" and not self.use_qr:
            a, u = self._project_svd_device(p0d, p1d)
            ws, vs = self._solve_reduced_device(a, y, u)
        else:
            p0, p1 = p0d.cpu().numpy(), p1d.cpu().numpy()
            if self.use_qr:
                a, p_back = self._project_qr(p0, p1)
            else:
                a, p_back = self._project_svd(p0, p1)
            ws, vs = self._solve_reduced_system(a, y, p_back)
        ws = ws.reshape(*batch_shape, self.m_0)
        vs = vs.reshape(*batch_shape, d, self.m_0)
        if old is not None:
            self.m_0 = old
        dev = q0.device
        return torch.as_tensor(ws, device=dev), torch.as_tensor(vs, device=dev)


class Full(NEVP):
    """Companion linearisation + dense eig (full.py:19-206)."""

    def __init__(self, eig_compute_location: str = "numpy", use_pinned_memory: bool = True,
                 reduce: bool = False, a_xx_sparsity=None):
        self.eig_compute_location = eig_compute_location
        self.use_pinned_memory = use_pinned_memory
        self.reduce = reduce
        if reduce and a_xx_sparsity is not None:
            self.zero_inds, self.nonzero_inds, self.all_inds = self._find_zero_columns(a_xx_sparsity)
            if self.zero_inds is None:
                warnings.warn("No columns are zero in the first and last blocks. "
                              "Reduction has no effect.")
                self.reduce = False

    @staticmethod
    def _find_zero_columns(a_xx):
        """full.py:53-97."""
        if not all(a.shape[0] == a.shape[1] for a in a_xx):
            raise ValueError("All arrays in a_xx must be square.")
        if not all(a.shape[0] == a_xx[0].shape[0] for a in a_xx):
            raise ValueError("All arrays in a_xx must have the same shape.")
        z_first = np.where(np.diff(a_xx[0].indptr) == 0)[0]
        z_last = np.where(np.diff(a_xx[-1].indptr) == 0)[0]
        offset = a_xx[0].shape[1] * (len(a_xx) - 2)
        zero = np.concatenate((z_first, z_last + offset))
        allc = np.arange(a_xx[0].shape[1] * (len(a_xx) - 1))
        nonzero = np.setdiff1d(allc, zero)
        if len(zero) == 0:
            return None, None, None
        return zero, nonzero, np.concatenate((nonzero, zero))

    def __call__(self, a_xx):
        """full.py:120-206: inverse of sum(a_xx) on the device, eig on the host."""
        a_xx = [as_device(a) for a in a_xx]
        if a_xx[0].dim() == 2:
            a_xx = [a[None] for a in a_xx]
        batch_shape = tuple(a_xx[0].shape[:-2])
        if a_xx[0].dim() > 3:
            a_xx = [a.reshape(-1, *a.shape[-2:]) for a in a_xx]
        if self.reduce and not hasattr(self, "zero_inds"):
            self.zero_inds, self.nonzero_inds, self.all_inds = self._find_zero_columns(
                tuple(sps.csc_matrix(a[0].cpu().numpy()) for a in a_xx))
            if self.zero_inds is None:
                warnings.warn("No columns are zero in the first and last blocks. "
                              "Reduction has no effect.")
                self.reduce = False
        inverse = inv(sum(a_xx).contiguous())
        parts = [gemm(inverse, sum(a_xx[:i]).contiguous()) for i in range(1, len(a_xx) - 1)]
        parts.append(gemm(inverse, (-a_xx[-1]).contiguous()))
        row = torch.cat(parts, dim=-1).cpu().numpy()
        A = np.concatenate([row] * (len(a_xx) - 1), axis=-2)
        B = np.kron(np.tri(len(a_xx) - 2).T, np.eye(a_xx[0].shape[-1]))
        A[:, :B.shape[0], :B.shape[1]] -= B
        if self.reduce:
            A_b = A[:, self.zero_inds, :][:, :, self.nonzero_inds]
            A_c = A[:, self.zero_inds, :][:, :, self.zero_inds]
            A = A[:, self.nonzero_inds, :][:, :, self.nonzero_inds]
        pairs = host_eig(A)
        w = np.stack([p[0] for p in pairs])
        v = np.stack([p[1] for p in pairs])
        if self.reduce:
            v_zero = (A_b @ v) / (w[:, np.newaxis, :]
                                  - np.diagonal(A_c, axis1=-2, axis2=-1)[:, :, np.newaxis])
            tmp = np.concatenate([v, v_zero], axis=1)
            v = np.empty_like(tmp)
            v[:, self.all_inds, :] = tmp
        w = np.where(np.abs(w) == 0.0, -1.0, w)
        w = 1 / w + 1
        v = v[:, :a_xx[0].shape[-1]]
        dev = inverse.device
        return (torch.as_tensor(w.reshape(*batch_shape, -1), device=dev),
                torch.as_tensor(v.reshape(*batch_shape, v.shape[1], -1), device=dev))
