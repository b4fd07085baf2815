"""CPU restatement of the per-k projector assembly (test infrastructure only:
the oracle checks the device path, it is never the measured or shipped code).

* ``qe_cubic_interpolate`` -- ``qepy_pw/upf.py:880-912``: QE's four-point
  cubic (Lagrange) interpolation of a radial table on a uniform q grid,
  samples table[i0 .. i0+3] with i0 = floor(q / dq) (the value part).
* ``qe_real_harmonics``    -- ``qepy_pw/upf.py:722-772``: real spherical
  harmonics in QE ``ylmr2`` order (m = 0, then cos and sin of m = 1..l),
  sqrt(2) Re / sqrt(2) Im of the complex harmonic with the Condon-Shortley
  phase, the +z direction at q = 0.  Evaluated here from SciPy's complex
  harmonics for every l, independently of the device's closed forms.
* ``projector_basis``      -- ``qepy_pw/upf.py:483-541`` (general branch):
  beta[g, j] = table_j(|q_g|) Y_{l_j m_j}(q_g) (-i)^{l_j} for channel j of
  radial projector p_j.

Pinned by: the reference's own known answer (interpolation exact for a
cubic polynomial, ``tests/unit/core/test_upf.py:41-48``), orthonormality of
the harmonics by quadrature on the sphere, the closed forms the reference
states for l <= 2 (``upf.py:738-746``), and the analytic Gaussian projector
transforms of the synthetic systems.
"""

from __future__ import annotations

import numpy as np
from scipy.special import sph_harm_y


def qe_cubic_interpolate(table: np.ndarray, q: np.ndarray, dq: float) -> np.ndarray:
    """Values of the four-point cubic interpolant of table (nq, ...) at q."""
    q = np.asarray(q, dtype=float).reshape(-1)
    scaled = q / dq
    lower = np.floor(scaled).astype(int)
    px = scaled - lower
    ux, vx, wx = 1.0 - px, 2.0 - px, 3.0 - px
    weights = np.stack([ux * vx * wx / 6.0, px * vx * wx / 2.0,
                        -px * ux * wx / 2.0, px * ux * vx / 6.0], axis=1)
    samples = np.stack([table[lower + o] for o in range(4)], axis=1)
    expand = (slice(None), slice(None)) + (None,) * (np.ndim(table) - 1)
    return np.sum(samples * weights[expand], axis=1)


def qe_real_harmonics(l: int, vectors: np.ndarray) -> np.ndarray:
    """(n, 2l+1) real harmonics of the directions of ``vectors`` (QE order)."""
    v = np.asarray(vectors, dtype=float)
    q = np.linalg.norm(v, axis=1)
    safe = q > 1.0e-14
    theta = np.zeros(len(v))
    theta[safe] = np.arccos(np.clip(v[safe, 2] / q[safe], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    cols = [np.real(sph_harm_y(l, 0, theta, phi))]
    for m in range(1, l + 1):
        y = sph_harm_y(l, m, theta, phi)
        cols += [np.sqrt(2.0) * np.real(y), np.sqrt(2.0) * np.imag(y)]
    return np.column_stack(cols)


def projector_basis(gk: np.ndarray, table: np.ndarray, dq: float, channels) -> np.ndarray:
    """beta (npw, nch) complex for channels [(radial index p, l, m), ...]."""
    gk = np.asarray(gk, dtype=float)
    q = np.linalg.norm(gk, axis=1)
    radial = qe_cubic_interpolate(np.asarray(table, dtype=float), q, dq)
    beta = np.empty((len(gk), len(channels)), dtype=np.complex128)
    harm = {}
    for j, (p, l, m) in enumerate(channels):
        if l not in harm:
            harm[l] = qe_real_harmonics(l, gk)
        beta[:, j] = radial[:, p] * harm[l][:, m] * (-1j) ** l
    return beta
