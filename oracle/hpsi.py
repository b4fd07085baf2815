"""Oracle h_psi: serial restatement of the reference Hamiltonian application.

Follows, as text:

* ``apply_local``      -- ``LocalPotentialWorkspace.apply`` serial branch
  (``qepy_pw/basis.py:1354-1455``) and ``apply_serial``
  (``qepy_pw/_native_fft.pyx:2338-2455``): zero the grid, scatter psi_G into
  the C-order slot ``ravel_multi_index(m mod n)`` (``basis.py:515-523``),
  unnormalised backward FFT (sign +1), multiply by V(r), forward FFT (sign
  -1), gather ``grid[slot]/N + T_G psi_G``.
* ``prepare_potential`` -- ``basis.py:777-792`` / ``_native_fft.pyx:2188-2202``.
* ``OracleHamiltonian``  -- ``PlaneWaveHamiltonian.__init__/diagonal/_apply``
  (``qepy_pw/diagonalization.py:127-369``): kinetic fused at gather, then
  calbec ``becp = beta^H psi`` and add_vuspsi ``Hpsi += beta (D becp)`` for
  materialised ``(beta, D)`` tuples, ``PackedProjectorTerm`` and
  ``FactorizedProjectorTerm`` (per-atom ``conj(psi)*phase`` then ``beta^T``).
* ``accumulate_density`` -- ``_native_fft.pyx:2592-2648``:
  ``rho(r) += w_b |psi_b(r)|^2`` with psi(r) the unnormalised backward FFT.

Pinned by (tests/test_oracle_hpsi.py): the reference's known-answer test
``tests/unit/core/test_basis.py:76-89`` (atol 2e-14), dense brute force
``potential_matrix + diag(T) + B D B^H`` on the small workload, Hermiticity,
the identity-potential round trip, and plane-wave eigenstates for constant V.
"""

from __future__ import annotations

import os

import numpy as np
import scipy.fft as sfft

_WORKERS = max(1, int(os.environ.get("ORACLE_FFT_WORKERS", os.cpu_count() or 1)))


def set_workers(n: int) -> None:
    global _WORKERS
    _WORKERS = max(1, int(n))


def linear_slots(indices: np.ndarray, shape) -> np.ndarray:
    idx = np.asarray(indices)
    return np.ravel_multi_index(tuple(idx[:, a] % shape[a] for a in range(3)),
                                tuple(shape))


def prepare_potential(potential_g: np.ndarray) -> np.ndarray:
    values = np.array(potential_g, dtype=np.complex128, order="C", copy=True)
    grid = sfft.ifftn(values, norm="forward", workers=_WORKERS)
    return np.asarray(grid.real).copy()


def apply_local(real_potential: np.ndarray, slots: np.ndarray, shape,
                vectors: np.ndarray, diagonal: np.ndarray | None = None,
                batch: int = 8, native_layout: bool = False) -> np.ndarray:
    """V_loc psi + T psi for a (npw, nb) block, reference serial semantics."""
    vectors = np.asarray(vectors)
    if vectors.ndim == 1:
        return apply_local(real_potential, slots, shape, vectors[:, None],
                           diagonal, batch, native_layout)[:, 0]
    n = int(np.prod(shape))
    pot = np.asarray(real_potential, dtype=np.float64)
    if native_layout:  # (n3, n1, n2) -> (n1, n2, n3)
        pot = np.transpose(pot, (1, 2, 0))
    pot = pot.reshape(shape)
    npw, nb = vectors.shape
    result = np.empty((npw, nb), dtype=np.complex128, order="F")
    for start in range(0, nb, batch):
        stop = min(nb, start + batch)
        grid = np.zeros((stop - start, n), dtype=np.complex128)
        grid[:, slots] = vectors[:, start:stop].T
        grid = grid.reshape((stop - start,) + tuple(shape))
        grid = sfft.ifftn(grid, axes=(1, 2, 3), norm="forward", workers=_WORKERS)
        grid *= pot[None]
        grid = sfft.fftn(grid, axes=(1, 2, 3), workers=_WORKERS)
        gathered = grid.reshape(stop - start, n)[:, slots].T / n
        if diagonal is not None:
            gathered = gathered + diagonal[:, None] * vectors[:, start:stop]
        result[:, start:stop] = gathered
    return result


def accumulate_density(vectors: np.ndarray, weights: np.ndarray, slots, shape,
                       batch: int = 8) -> np.ndarray:
    """sum_b w_b |psi_b(r)|^2 on the (n1, n2, n3) grid (unnormalised FFT)."""
    n = int(np.prod(shape))
    rho = np.zeros(n)
    nb = vectors.shape[1]
    for start in range(0, nb, batch):
        stop = min(nb, start + batch)
        grid = np.zeros((stop - start, n), dtype=np.complex128)
        grid[:, slots] = vectors[:, start:stop].T
        grid = grid.reshape((stop - start,) + tuple(shape))
        grid = sfft.ifftn(grid, axes=(1, 2, 3), norm="forward", workers=_WORKERS)
        rho += np.einsum("b,bn->n", np.asarray(weights[start:stop], float),
                         (grid.real ** 2 + grid.imag ** 2).reshape(stop - start, n))
    return rho.reshape(shape)


def _term_kind(term):
    if hasattr(term, "phases"):
        return "factorized"
    if hasattr(term, "diagonal") and hasattr(term, "beta"):
        return "packed"
    return "materialized"


class OracleHamiltonian:
    """CPU restatement of ``PlaneWaveHamiltonian`` (serial)."""

    def __init__(self, indices, kinetic, shape, real_potential, projector_terms=(),
                 potential_average: float = 0.0, native_layout: bool = False):
        self.indices = np.asarray(indices)
        self.kinetic = np.asarray(kinetic, dtype=float)
        self.shape = tuple(shape)
        self.slots = linear_slots(self.indices, self.shape)
        self.real_potential = np.asarray(real_potential, dtype=float)
        self.native_layout = native_layout
        self.projector_terms = tuple(projector_terms)
        self.potential_average = float(potential_average)
        self.applications = 0

    @property
    def diagonal(self) -> np.ndarray:
        result = self.kinetic + self.potential_average
        for term in self.projector_terms:
            kind = _term_kind(term)
            if kind == "packed":
                result = result + term.diagonal
                continue
            if kind == "factorized":
                beta, coupling, mult = term.beta, term.coupling, term.phases.shape[1]
            else:
                (beta, coupling), mult = term, 1
            result = result + mult * np.real(
                np.einsum("gi,ij,gj->g", beta, coupling, beta.conj(), optimize=True))
        return result

    def apply(self, vectors: np.ndarray) -> np.ndarray:
        vec = np.asarray(vectors)
        was_vector = vec.ndim == 1
        if was_vector:
            vec = vec[:, None]
        self.applications += vec.shape[1]
        result = apply_local(self.real_potential, self.slots, self.shape, vec,
                             self.kinetic, native_layout=self.native_layout)
        for term in self.projector_terms:
            kind = _term_kind(term)
            if kind == "factorized":
                for atom in range(term.phases.shape[1]):
                    phase = term.phases[:, atom]
                    overlap = np.conjugate(term.beta.T @ (np.conjugate(vec) * phase[:, None]))
                    coupled = term.coupling @ overlap
                    result += (term.beta @ coupled) * phase[:, None]
            else:
                beta, coupling = ((term.beta, term.coupling) if kind == "packed"
                                  else term)
                overlap = beta.conj().T @ vec
                result += beta @ (coupling @ overlap)
        return result[:, 0] if was_vector else result

    def apply_into(self, vectors, out):
        out[...] = self.apply(vectors)
        return out
