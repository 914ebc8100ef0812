"""CPU oracle for the batched RGF + open-boundary hot path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference algorithms (PaulFisch/quatrex, ``src/qttools``) that the HIP
implementation in ``quatrex_amd`` is checked against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the timed CPU baseline -- never as
the thing measured or shipped.  ``quatrex_amd`` never imports it.

Modules and what pins them (each check lives under ``tests/``):

- ``dsdbcoo``   block-sorted COO stack container (``datastructures/dsdbcoo.py``)
                -- pinned by dense round trips against scipy.
- ``rgf``       RGF selected inversion / selected solve (``rgf.py:25-414``)
                -- pinned by dense brute force (``inv`` and ``G Sigma G^H``),
                anti-Hermiticity, the ``G^> - G^< = G^R - G^A`` identity and
                the dense Meir-Wingreen layer current.
- ``obc``       Sancho-Rubio, OBCSystem memoizer, Spectral OBC
                (``sancho_rubio.py``, ``obc.py``, ``boundary_system.py``,
                ``spectral.py``) -- pinned by the fixed-point residual, the
                closed-form surface Green's function of a 1D chain, and
                Sancho-Rubio vs Spectral agreement.
- ``nevp``      Beyn contour NEVP, Full linearisation, operator_inverse
                (``beyn.py``, ``full.py``, ``operator.py``) -- pinned by
                eigen-residuals, eigenvalue agreement with the companion
                linearisation and a direct ``inv(sum z^n a_n)``.
- ``batching``  ``get_batches`` / ``get_section_sizes`` -- pinned by the
                documented examples of the reference docstring.

Parity unpinned: the exact mode-selection branch decisions of
``Spectral._find_reflected_modes`` at threshold edges on real lead
Hamiltonians, and the multi-rank agreement vote of the memoizer's ``auto``
mode (single-rank behaviour is pinned).
"""
