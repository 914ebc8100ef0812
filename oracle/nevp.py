"""Non-linear eigenvalue solvers (oracle restatement).

- ``operator_inverse``  follows ``src/qttools/kernels/operator.py:113-179``
  (NumPy branch: ``sum(z**n * a_n)`` then ``inv``).
- ``Beyn``              follows ``src/qttools/nevp/beyn.py:14-379``: persistent
  ``default_rng(42)`` probe stream, annulus quadrature with nodes
  ``r * exp(2 pi i (k + 1.5) / N)`` for k = 0..N-1 and weights ``1/N``,
  moments ``q0, q1``, SVD projection with cutoff ``s_max * d * eps`` (or QR),
  small eigenproblem, back-projection.
- ``Full``              follows ``src/qttools/nevp/full.py:120-206``:
  companion linearisation of the spectral transform ``w -> 1/(w - 1)``,
  dense ``eig``, optional zero-column reduction.
"""

import warnings

import numpy as np
import scipy.sparse as sps

from oracle.batching import get_section_sizes


def operator_inverse(a_xx, z, contour_type, in_type):
    """inv(sum_n z^n a_n), n = -b..b with b = len(a_xx)//2 (operator.py:113-179)."""
    b = len(a_xx) // 2
    op = sum(z ** n * a for a, n in zip(a_xx, range(-b, b + 1)))
    return np.linalg.inv(op.astype(contour_type)).astype(in_type)


class Beyn:
    """Beyn contour-integral NEVP (beyn.py:14-94)."""

    def __init__(self, r_o, r_i, m_0, num_quad_points, use_qr=False,
                 contour_batch_size=None, **_ignored):
        self.r_o, self.r_i, self.m_0 = r_o, r_i, m_0
        self.num_quad_points = num_quad_points
        self.use_qr = use_qr
        if contour_batch_size is None:
            contour_batch_size = num_quad_points
        self.contour_batch_size = contour_batch_size
        counts, _ = get_section_sizes(
            num_quad_points, int(np.ceil(num_quad_points / contour_batch_size)))
        self.contour_displacements = np.cumsum(np.concatenate(([0], np.array(counts))))
        self.rng = np.random.default_rng(42)

    def quadrature(self):
        """Nodes and weights of beyn.py:198-202."""
        zeta = np.arange(self.num_quad_points) + 1
        phase = np.exp(2j * np.pi * (zeta + 0.5) / self.num_quad_points)
        w = np.ones(self.num_quad_points) / self.num_quad_points
        return self.r_o * phase, self.r_i * phase, w

    def _contour_integrate(self, a_xx):
        """beyn.py:177-235."""
        in_type = a_xx[0].dtype
        z_o, z_i, w = self.quadrature()
        z_o, z_i, w = (v.reshape(1, -1, 1, 1) for v in (z_o, z_i, w))
        q0 = np.zeros_like(a_xx[0])
        q1 = np.zeros_like(a_xx[0])
        a4 = [a[:, np.newaxis, :, :] for a in a_xx]
        d = self.contour_displacements
        for j in range(len(d) - 1):
            zo, zi, wj = z_o[:, d[j]:d[j + 1]], z_i[:, d[j]:d[j + 1]], w[:, d[j]:d[j + 1]]
            inv_o = operator_inverse(a4, zo, in_type, in_type)
            inv_i = operator_inverse(a4, zi, in_type, in_type)
            q0 += np.sum(wj * zo * inv_o - wj * zi * inv_i, axis=1)
            q1 += np.sum(wj * zo ** 2 * inv_o - wj * zi ** 2 * inv_i, axis=1)
        return q0, q1

    def _project_svd(self, p0, p1):
        """beyn.py:96-145."""
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

    def _project_qr(self, p0, p1):
        """beyn.py:147-175."""
        q, r = np.linalg.qr(p0)
        return np.swapaxes(q.conj(), -2, -1) @ p1 @ np.linalg.inv(r), q

    def _solve_reduced_system(self, a, y, p_back):
        """beyn.py:237-287."""
        in_type = a[0].dtype
        if isinstance(a, list):
            ws = np.zeros((len(a), self.m_0), dtype=in_type)
            vs = y.copy()
            for i in range(len(a)):
                w, v = np.linalg.eig(a[i])
                ws[i, :len(w)] = w
                vs[i, :, :len(w)] = p_back[i] @ v
            return ws, vs
        w, v = np.linalg.eig(a)
        return w, p_back @ v

    def _one_sided(self, a_xx):
        """beyn.py:289-331."""
        d = a_xx[0].shape[-1]
        n = a_xx[0].shape[0]
        y = self.rng.random((n, d, self.m_0)) + 1j * self.rng.random((n, d, self.m_0))
        q0, q1 = self._contour_integrate(a_xx)
        p0, p1 = q0 @ y, q1 @ y
        if self.use_qr:
            a, p_back = self._project_qr(p0, p1)
        else:
            a, p_back = self._project_svd(p0, p1)
        return self._solve_reduced_system(a, y, p_back)

    def __call__(self, a_xx):
        """beyn.py:333-379."""
        if a_xx[0].ndim == 2:
            a_xx = tuple(a[np.newaxis] for a in a_xx)
        batch_shape = a_xx[0].shape[:-2]
        d = a_xx[0].shape[-1]
        a_xx = tuple(a.reshape(-1, d, d) for a in a_xx)
        old = None
        if d < self.m_0 and self.use_qr:
            warnings.warn(f"Subspace guess {self.m_0} is larger than the dimension "
                          f"of the system {d}. Setting subspace guess to {d}.")
            old, self.m_0 = self.m_0, d
        ws, vs = self._one_sided(a_xx)
        ws = ws.reshape(*batch_shape, self.m_0)
        vs = vs.reshape(*batch_shape, d, self.m_0)
        if old is not None:
            self.m_0 = old
        return ws, vs


class Full:
    """Companion-linearisation NEVP (full.py:19-206)."""

    def __init__(self, reduce=False, a_xx_sparsity=None, **_ignored):
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
        """full.py:120-206."""
        if a_xx[0].ndim == 2:
            a_xx = tuple(a[np.newaxis] for a in a_xx)
        batch_shape = a_xx[0].shape[:-2]
        if a_xx[0].ndim > 3:
            a_xx = tuple(a.reshape(-1, *a.shape[-2:]) for a in a_xx)
        if self.reduce and not hasattr(self, "zero_inds"):
            self.zero_inds, self.nonzero_inds, self.all_inds = self._find_zero_columns(
                tuple(sps.csc_matrix(a[0]) for a in a_xx))
            if self.zero_inds is None:
                warnings.warn("No columns are zero in the first and last blocks. "
                              "Reduction has no effect.")
                self.reduce = False
        inverse = np.linalg.inv(sum(a_xx))
        row = np.concatenate([inverse @ sum(a_xx[:i]) for i in range(1, len(a_xx) - 1)]
                             + [inverse @ -a_xx[-1]], axis=-1)
        A = np.concatenate([row] * (len(a_xx) - 1), axis=-2)
        B = np.kron(np.tri(len(a_xx) - 2).T, np.eye(a_xx[0].shape[-1]))
        A[:, :B.shape[0], :B.shape[1]] -= B
        if self.reduce:
            A_b = A[:, self.zero_inds, :][:, :, self.nonzero_inds]
            A_c = A[:, self.zero_inds, :][:, :, self.zero_inds]
            A = A[:, self.nonzero_inds, :][:, :, self.nonzero_inds]
        w, v = np.linalg.eig(A)
        if self.reduce:
            v_zero = (A_b @ v) / (w[:, np.newaxis, :]
                                  - np.diagonal(A_c, axis1=-2, axis2=-1)[:, :, np.newaxis])
            tmp = np.concatenate([v, v_zero], axis=1)
            v = np.empty_like(tmp)
            v[:, self.all_inds, :] = tmp
        w = np.where(np.abs(w) == 0.0, -1.0, w)
        w = 1 / w + 1
        v = v[:, :a_xx[0].shape[-1]]
        return w.reshape(*batch_shape, -1), v.reshape(*batch_shape, v.shape[1], -1)
