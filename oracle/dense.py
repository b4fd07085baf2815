"""Dense brute-force Hamiltonian, the independent check of ``oracle.hpsi``.

H_GG' = V(G-G') + delta_GG' T_G + sum_terms B D B^H, with V(G) the forward
FFT of V(r) divided by N (``qepy_pw/basis.py:1685-1688`` ``potential_matrix``
and ``tests/unit/core/test_basis.py:76-89``).  It shares no code path with
the FFT sandwich in ``oracle.hpsi`` and is what pins it.
"""

from __future__ import annotations

import numpy as np


def potential_g_from_real(real_potential: np.ndarray) -> np.ndarray:
    return np.fft.fftn(real_potential) / real_potential.size


def potential_matrix(vg: np.ndarray, indices: np.ndarray) -> np.ndarray:
    delta = indices[:, None, :] - indices[None, :, :]
    return vg[tuple(delta[:, :, a] % vg.shape[a] for a in range(3))]


def dense_hamiltonian(indices, kinetic, real_potential, projector_terms=()):
    vg = potential_g_from_real(np.asarray(real_potential, float))
    h = potential_matrix(vg, np.asarray(indices)).astype(np.complex128)
    h[np.diag_indices_from(h)] += kinetic
    for term in projector_terms:
        if hasattr(term, "phases"):
            for a in range(term.phases.shape[1]):
                b = term.beta * term.phases[:, a, None]
                h += b @ term.coupling @ b.conj().T
        elif hasattr(term, "diagonal"):
            h += term.beta @ term.coupling @ term.beta.conj().T
        else:
            beta, coupling = term
            h += beta @ coupling @ beta.conj().T
    return h
