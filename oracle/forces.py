"""Oracle gradients of the band energy E = sum_n f_n <psi_n|H|psi_n>.

* ``nonlocal_energy_and_forces`` restates the force branch of
  ``_nonlocal_derivatives`` (``qepy_pw/pw/scf.py:1521-1631``): per atom,
  ``modulated = conj(phase) psi`` with ``phase = exp(-i gk.tau)``,
  ``overlap = conj(beta)^T modulated``, ``coupled = D overlap``,
  ``energy += w sum conj(overlap) coupled f`` and
  ``derivative_overlap = sum_g conj(beta) (1j gk) modulated``,
  ``F_a -= 2 w Re sum conj(derivative_overlap) coupled f`` (lines 1573-1599).
* ``band_energy_gradients`` adds the local part: dE/dV(r) = (1/N) sum_n f_n
  |psi_n(r)|^2 (psi(r) the unnormalised backward FFT, the density kernel of
  ``_native_fft.pyx:2592-2648``), and dE/dpsi* = f_n H psi_n (returned as
  torch's convention 2 dE/dpsi*).

Pinned by (tests/test_oracle_forces.py): central finite differences of the
oracle energy in tau (h = 1e-5 bohr) and V (exact: E is linear in V), and the
identity E_NL(oracle forces) = the nonlocal part of the oracle H expectation.
"""

from __future__ import annotations

import numpy as np

from .hpsi import OracleHamiltonian, accumulate_density, linear_slots


def nonlocal_energy_and_forces(vectors, occupations, gk, beta_species, couplings, positions,
                               weight=1.0):
    """(E_NL, F) for atoms with one shared species beta (npw, nch) and
    per-atom couplings; F = -dE_NL/dtau, shape (natom, 3)."""
    beta_c = np.conjugate(beta_species)
    f = np.asarray(occupations, dtype=float)
    energy = 0.0
    forces = np.zeros((len(positions), 3))
    for a, tau in enumerate(positions):
        phase = np.exp(-1j * (gk @ tau))
        modulated = np.conjugate(phase)[:, None] * vectors
        overlap = beta_c.T @ modulated
        coupled = couplings[a] @ overlap
        energy += weight * float(np.real(np.sum(np.conjugate(overlap) * coupled * f[None, :])))
        derivative = np.einsum("gp,gab->pab", beta_c, 1j * gk[:, :, None] * modulated[:, None, :],
                               optimize=True)
        forces[a] -= 2.0 * weight * np.real(
            np.einsum("pab,pb,b->a", np.conjugate(derivative), coupled, f, optimize=True))
    return energy, forces


def packed_projectors(gk, beta_species, couplings, positions):
    """Atom-packed (beta (npw, P), block-diagonal D) of the same projectors."""
    nch = beta_species.shape[1]
    natom = len(positions)
    beta = np.empty((len(gk), nch * natom), dtype=np.complex128)
    coupling = np.zeros((nch * natom, nch * natom))
    for a, tau in enumerate(positions):
        blk = slice(a * nch, (a + 1) * nch)
        beta[:, blk] = beta_species * np.exp(-1j * (gk @ tau))[:, None]
        coupling[blk, blk] = couplings[a]
    return beta, coupling


class _Packed:
    def __init__(self, beta, coupling):
        self.beta = beta
        self.coupling = coupling
        self.diagonal = np.zeros(beta.shape[0])


def band_energy_gradients(indices, kinetic, gk, shape, real_potential, beta_species,
                          couplings, positions, vectors, occupations):
    """E, dE/dV (n1, n2, n3), dE/dtau (natom, 3), torch-convention grad psi."""
    beta, coupling = packed_projectors(gk, beta_species, couplings, positions)
    ham = OracleHamiltonian(indices, kinetic, shape, real_potential, (_Packed(beta, coupling),))
    hpsi = ham.apply(vectors)
    f = np.asarray(occupations, dtype=float)
    energy = float(np.real(np.einsum("gb,gb,b->", vectors.conj(), hpsi, f)))
    n = int(np.prod(shape))
    dv = accumulate_density(vectors, f, linear_slots(indices, shape), shape) / n
    _, forces = nonlocal_energy_and_forces(vectors, f, gk, beta_species, couplings, positions)
    grad_psi = 2.0 * hpsi * f[None, :]
    return energy, dv, -forces, grad_psi
