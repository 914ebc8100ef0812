"""Open-boundary solvers (oracle restatement).

- ``SanchoRubio``   ``src/qttools/boundary_conditions/obc/sancho_rubio.py:10-86``
- ``OBCSystem``     ``obc/obc.py:52-202`` on ``boundary_system.py:48-373`` with
                    the identity reducer of ``system_reduction.py:101-180``
- ``Spectral``      ``obc/spectral.py:15-526``
- ``boundary_self_energies``  ``quatrex/electron/solver.py:304-337`` (left
                    contact: ``Sigma = m10 g m01``, ``Gamma = i(Sigma - Sigma^H)``,
                    ``Sigma^< = i f Gamma``, ``Sigma^> = i (f - 1) Gamma``)
"""

import warnings

import numpy as np


def _h(x):
    return np.conj(np.swapaxes(x, -2, -1))


def block_view(arr, axis, num_blocks):
    """Strided block view (``dsdbsparse.py:35-70``)."""
    shape = list(arr.shape)
    if shape[axis] % num_blocks != 0:
        raise ValueError("The array shape is not divisible by the number of blocks.")
    shape[axis] //= num_blocks
    strides = (arr.strides[axis] * shape[axis],) + arr.strides
    return np.lib.stride_tricks.as_strided(arr, shape=(num_blocks,) + tuple(shape), strides=strides)


class SanchoRubio:
    """Decimation surface Green's function (sancho_rubio.py:28-86)."""

    def __init__(self, max_iterations: int = 100, convergence_tol: float = 1e-6):
        self.max_iterations = max_iterations
        self.convergence_tol = convergence_tol
        self.iterations = 0

    def __call__(self, a_ii, a_ij, a_ji, contact=""):
        eps, eps_s = a_ii.copy(), a_ii.copy()
        alpha, beta = a_ji.copy(), a_ij.copy()
        self.iterations = 0
        for _ in range(self.max_iterations):
            self.iterations += 1
            g = np.linalg.inv(eps)
            eps = eps - alpha @ g @ beta - beta @ g @ alpha
            eps_s = eps_s - alpha @ g @ beta
            alpha = alpha @ g @ alpha
            beta = beta @ g @ beta
            delta = np.linalg.norm(np.abs(alpha) + np.abs(beta), axis=(-2, -1)).max() / 2
            if delta < self.convergence_tol:
                break
        else:
            warnings.warn("Surface Green's function did not converge.", RuntimeWarning)
        return np.linalg.inv(eps_s)


class OBCSystem:
    """Memoizing OBC wrapper (obc.py:52-202, boundary_system.py:48-373).

    Single rank: the ``auto`` agreement vote divides the converged-energy
    count by a stack-communicator size of 1 (boundary_system.py:280-296),
    exactly as the reference does on one rank.
    """

    def __init__(self, boundary_solver, num_ref_iterations=2, relative_tol=2e-1,
                 absolute_tol=1e-6, warning_threshold=1e-1, memoization_mode="auto",
                 agreement_threshold=0.999, stack_comm_size=1):
        self.boundary_solver = boundary_solver
        self.num_ref_iterations = num_ref_iterations
        self.relative_tol = relative_tol
        self.absolute_tol = absolute_tol
        self.warning_threshold = warning_threshold
        self.memoization_mode = memoization_mode
        self.agreement_threshold = agreement_threshold
        self.stack_comm_size = stack_comm_size
        if not (0.0 <= agreement_threshold <= 1.0):
            raise ValueError("Memoizing tolerance must be between 0 and 1.")
        self._cache = {}
        if self.num_ref_iterations < 2:
            warnings.warn("The number of refinement iterations should be at least 2. "
                          "Defaulting to 2.", RuntimeWarning)
            self.num_ref_iterations = 2

    @staticmethod
    def _fixed_point_step(system, x):
        a_ii, a_ij, a_ji = system
        return np.linalg.inv(a_ii - a_ji @ x @ a_ij)

    @staticmethod
    def _get_starting_guess(system):
        return np.linalg.inv(system[0])

    def _get_residuals(self, system, x):
        ref = self._fixed_point_step(system, x)
        abs_res = np.linalg.norm(ref - x, axis=(-2, -1))
        rel_res = abs_res / np.linalg.norm(ref, axis=(-2, -1))
        return rel_res, abs_res, ref

    def _solve(self, system, contact):
        x = self.boundary_solver(*system, contact)
        if self.memoization_mode != "off":
            self._cache[contact] = x.copy()
        return x

    def _memoized_solve(self, system, contact):
        if self.memoization_mode == "off":
            return self._solve(system, contact)
        x = self._cache.get(contact, None)
        if x is None:
            if self.memoization_mode in ("auto", "force-after-first"):
                return self._solve(system, contact)
            elif self.memoization_mode == "force":
                x = self._get_starting_guess(system)
            else:
                raise RuntimeError(f"Invalid memoizing mode: {self.memoization_mode}")
        rel, abs_, x = self._get_residuals(system, x)
        if self.memoization_mode == "auto":
            local = int(np.sum((abs_ < self.absolute_tol) | (rel < self.relative_tol)))
            if local / self.stack_comm_size > self.agreement_threshold:
                return self._solve(system, contact)
        for _ in range(self.num_ref_iterations - 2):
            x = self._fixed_point_step(system, x)
        return x

    def __call__(self, system, contact):
        x = self._memoized_solve(system, contact)
        rel, abs_, x = self._get_residuals(system, x)
        if np.any((rel > self.warning_threshold) & (abs_ > self.absolute_tol)):
            warnings.warn(f"High error for {contact}: relative recursion error "
                          f"{np.max(rel):.3e}, absolute {np.max(abs_):.3e}", RuntimeWarning)
        if self.memoization_mode != "off":
            self._cache[contact] = x.copy()
        return x, rel, abs_


class Spectral:
    """Spectral OBC from NEVP eigenmodes (spectral.py:15-95)."""

    def __init__(self, nevp, block_sections=1, min_decay=1e-3, max_decay=6.9,
                 num_ref_iterations=2, min_propagation=0.01, residual_tolerance=1e-3,
                 residual_normalization=True, warning_threshold=1e-1, eta_decay=1e-14):
        self.nevp = nevp
        self.block_sections = block_sections
        self.min_decay, self.max_decay = min_decay, max_decay
        self.num_ref_iterations = num_ref_iterations
        self.min_propagation = min_propagation
        self.residual_tolerance = residual_tolerance
        self.residual_normalization = residual_normalization
        self.warning_threshold = warning_threshold
        self.eta_decay = eta_decay
        if num_ref_iterations < 1:
            raise ValueError("Number of refinement iterations must be at least 1.")

    def _extract_subblocks(self, a_ji, a_ii, a_ij):
        """spectral.py:97-158."""
        layer = (a_ji, a_ii, a_ij)
        if self.block_sections == 1:
            return layer
        bs = self.block_sections
        view = block_view(np.concatenate(layer, axis=-1), -1, 3 * bs)
        view = block_view(view, -2, bs)
        first = np.linalg.norm(view[0, :])
        errs = [np.linalg.norm(view[0, :] - np.roll(view[i, :], -i, axis=0)) / first
                for i in range(1, bs)]
        if max(errs) > 1e-3:
            warnings.warn(f"Requested block sectioning is not periodic. ({max(errs):.2e})",
                          RuntimeWarning)
        blocks = view[0, :-bs + 1]
        idx = np.where([np.any(b) for b in blocks])[0]
        if idx.size == 0 or len(blocks) <= 3:
            return tuple(blocks)
        n_data = min(idx[0], len(blocks) - 1 - idx[-1])
        n = min(n_data, (len(blocks) - 3) // 2)
        return tuple(blocks[n:-n]) if n > 0 else tuple(blocks)

    @staticmethod
    def _compute_dE_dk(ws, vs, a_xx):
        """spectral.py:160-196."""
        b = len(a_xx) // 2
        with np.errstate(divide="ignore", invalid="ignore"):
            return -sum((1j * n) * np.diagonal(_h(vs) @ a @ vs, axis1=-2, axis2=-1) * ws ** n
                        for a, n in zip(a_xx, range(-b, b + 1)))

    def _find_reflected_modes(self, ws, vs, a_xx):
        """spectral.py:198-320 (without injection vectors)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            prod = sum(a @ vs * ws[..., np.newaxis, :] ** (i - len(a_xx) // 2)
                       for i, a in enumerate(a_xx))
            res = np.linalg.norm(prod, axis=-2)
            res /= np.linalg.norm(vs, axis=-2)
            if self.residual_normalization:
                res /= np.abs(ws)
        dk = self._compute_dE_dk(ws, vs, a_xx)
        with np.errstate(divide="ignore", invalid="ignore"):
            ks = -1j * np.log(ws)
        dk = np.nan_to_num(dk, nan=0, posinf=0, neginf=0)
        ks = np.nan_to_num(ks, nan=0, posinf=0, neginf=0)
        eta = np.finfo(dk.dtype).eps
        ratio = np.abs(dk.real) / (np.abs(dk.imag) + eta)
        prop = (np.abs(ks.imag) < self.min_decay) & (self.min_propagation < ratio) & (dk.real < 0)
        dec = ks.imag < -self.min_decay
        dec |= (self.min_propagation >= ratio) & (ks.imag < -self.eta_decay)
        dec &= ks.imag > -self.max_decay
        return (prop | dec) & (res < self.residual_tolerance)

    def _upscale_eigenmodes(self, ws, vs):
        """spectral.py:322-377."""
        with np.errstate(divide="ignore", invalid="ignore"):
            if self.block_sections == 1:
                return ws, vs / np.linalg.norm(vs, axis=-2, keepdims=True)
            batch = vs.shape[:-2]
            nb = len(batch)
            sub, modes = vs.shape[-2], vs.shape[-1]
            wsu = np.moveaxis(np.array([ws ** n for n in range(self.block_sections)]), 0, nb)
            vsu = (wsu[..., :, np.newaxis, :] * vs[..., np.newaxis, :, :]).reshape(
                *batch, sub * self.block_sections, modes)
            vsu = vsu / np.linalg.norm(vsu, axis=-2, keepdims=True)
        return ws ** self.block_sections, vsu

    @staticmethod
    def _compute_x_ii(a_ii, a_ij, a_ji, ws, vs, mask):
        """spectral.py:379-423."""
        x_a = np.zeros(mask.shape[:-1] + a_ij.shape[-2:], dtype=a_ij.dtype)
        for i in np.ndindex(mask.shape[:-1]):
            m = mask[i]
            vr, w = vs[i][:, m], ws[i][m]
            v_inv = np.linalg.inv(vr.conj().T @ vr) @ vr.conj().T
            x_a[i] = vr / w @ v_inv
        return np.linalg.inv(a_ii + a_ji @ x_a)

    def __call__(self, a_ii, a_ij, a_ji, contact=""):
        """spectral.py:425-526 (return_injected=False)."""
        if a_ii.ndim == 2:
            a_ii, a_ij, a_ji = a_ii[np.newaxis], a_ij[np.newaxis], a_ji[np.newaxis]
        blocks = self._extract_subblocks(a_ji, a_ii, a_ij)
        ws, vs = self.nevp(blocks)
        ws, vs = self._upscale_eigenmodes(ws, vs)
        mask = self._find_reflected_modes(ws, vs, [a_ji, a_ii, a_ij])
        x = self._compute_x_ii(a_ii, a_ij, a_ji, ws, vs, mask)
        for _ in range(self.num_ref_iterations - 1):
            x = np.linalg.inv(a_ii - a_ji @ x @ a_ij)
        return np.linalg.inv(a_ii - a_ji @ x @ a_ij)


def boundary_self_energies(g, m10, m01, occupancies):
    """Retarded/lesser/greater boundary self-energies (electron/solver.py:324-337)."""
    sigma = m10 @ g @ m01
    gamma = 1j * (sigma - _h(sigma))
    f = np.asarray(occupancies).reshape(-1, *([1] * (gamma.ndim - 1)))
    return sigma, 1j * gamma * f, 1j * gamma * (f - 1)
