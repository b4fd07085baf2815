"""A stand-in for the reference SCF's use of the eigensolver names.

``make_scf_module`` builds a module object holding the names that
``qepy_pw/pw/scf.py:29-40`` imports, bound *by value* to host
implementations with the reference's signatures (the oracle's h_psi and
Davidson behind a ``PlaneWaveHamiltonian`` constructor of the reference's
form).  ``solve_kpoint`` then drives one k-point the way the SCF does,
looking every name up in that module at call time, as the SCF's module
globals are looked up:

* the operator: ``PlaneWaveHamiltonian(basis, v_eff_g=[<V>], terms,
  local_workspace=..., timers=..., real_potential=<native layout>,
  native_potential_layout=True, potential_average=v_ion_average)``
  (``pw/scf.py:3282-3288, 3396-3411``);
* the starting rotation ``_rotate_starting_subspace``
  (``pw/scf.py:923-953``): H and S of the starting block through
  ``operator.apply_into`` and ``_lowest_generalized_eigh``, rotated with
  ``_matmul_f``;
* Davidson with the SCF's keywords and its retry loop
  (``pw/scf.py:3497-3556``): ``max_iterations=20``, ``residual_factor=None``,
  ``initial_is_ritz=True`` with the rotated eigenvalues and applied block,
  then up to five retries from the previous eigenvectors with
  ``initial_applied=None`` while any band is unconverged (the
  fixed-potential rule).

Test infrastructure only.
"""

from __future__ import annotations

import types

import numpy as np

from oracle import davidson as odav
from oracle.hpsi import OracleHamiltonian


class SerialMPI:
    """The attributes of the reference's serial ``MPIContext`` the path reads
    (``qepy_pw/mpi.py:14-31, 129-136``)."""

    rank = 0
    size = 1
    comm = None

    def sum_array(self, value):
        return np.ascontiguousarray(value)

    def sum_scalar(self, value):
        return float(value)


class ReferenceWorkspace:
    """Only the attributes of the reference's serial
    ``LocalPotentialWorkspace`` that ``PlaneWaveHamiltonian`` reads
    (``diagonalization.py:139-163``; ``basis.py:460-555``)."""

    def __init__(self, basis, shape):
        self.shape = tuple(int(n) for n in shape)
        self.mpi = SerialMPI()
        self.local_plane_wave_indices = np.arange(len(basis.indices))


class HostHamiltonian:
    """Reference-signature constructor over the oracle h_psi."""

    def __init__(self, basis, potential_g, projector_terms=(), local_workspace=None,
                 timers=None, real_potential=None, native_potential_layout=False,
                 potential_average=None):
        self.basis = basis
        self.potential_g = np.asarray(potential_g)
        self.local_workspace = local_workspace
        self.mpi = local_workspace.mpi
        self.local_rows = local_workspace.local_plane_wave_indices
        self.local_kinetic = np.asarray(basis.kinetic)
        average = (float(potential_average) if potential_average is not None
                   else float(np.real(self.potential_g.flat[0])))
        self._oracle = OracleHamiltonian(basis.indices, basis.kinetic, local_workspace.shape,
                                         real_potential, projector_terms, average,
                                         native_layout=native_potential_layout)
        self.projector_terms = tuple(projector_terms)

    @property
    def diagonal(self):
        return self._oracle.diagonal

    def apply(self, vectors):
        return self._oracle.apply(vectors)

    def apply_into(self, vectors, out):
        out[...] = self._oracle.apply(vectors)
        return out


def _host_davidson(operator, diagonal, number_of_roots, *, mpi=None, global_dimension=None,
                   global_row_indices=None, timers=None, **kwargs):
    return odav.davidson(operator, diagonal, number_of_roots, **kwargs)


def _matmul_f(left, right):
    out = np.empty((left.shape[0], right.shape[1]), dtype=np.complex128, order="F")
    np.matmul(left, right, out=out)
    return out


def make_scf_module(packed_cls, factorized_cls):
    """A module holding the SCF's imported names, bound to host versions."""
    scf = types.ModuleType("scf_standin")
    scf.PackedProjectorTerm = packed_cls
    scf.FactorizedProjectorTerm = factorized_cls
    scf.ProjectorTerm = object
    scf.PlaneWaveHamiltonian = HostHamiltonian
    scf._lowest_generalized_eigh = odav.lowest_generalized_eigh
    scf._matmul_f = _matmul_f
    scf.davidson = _host_davidson
    scf.conjugate_gradient = scf.parallel_orbital = scf.rmm_diis = None
    return scf


def rotate_starting_subspace(scf, operator, trials, number_of_bands, mpi):
    """``pw/scf.py:923-953``."""
    atomic_basis = np.asarray(trials, dtype=complex)
    applied = np.empty_like(atomic_basis, order="F")
    operator.apply_into(atomic_basis, applied)
    projected_local = atomic_basis.conj().T @ applied
    overlap_local = atomic_basis.conj().T @ atomic_basis
    split = projected_local.size
    reduced = mpi.sum_array(np.concatenate((projected_local.ravel(), overlap_local.ravel())))
    projected = reduced[:split].reshape(projected_local.shape)
    overlap = reduced[split:].reshape(overlap_local.shape)
    projected = 0.5 * (projected + projected.conj().T)
    overlap = 0.5 * (overlap + overlap.conj().T)
    values, rotation = scf._lowest_generalized_eigh(projected, overlap, number_of_bands)
    return (scf._matmul_f(atomic_basis, rotation), values,
            scf._matmul_f(applied, rotation))


def solve_kpoint(scf, system, terms, nbnd, starting, tolerance, *, ndim=4, v_ion_shift=0.05,
                 timers=None, max_attempts=6):
    """One k-point of the SCF's Davidson branch; returns the last solution,
    the attempt count and the summed iterations / H applications."""
    mpi = SerialMPI()
    shape = system.shape
    workspace = ReferenceWorkspace(system.basis, shape)
    v_c = np.asarray(system.real_potential, dtype=float)
    v_native = np.ascontiguousarray(np.transpose(v_c, (2, 0, 1)))   # (n3, n1, n2)
    average = float(np.sum(v_c)) / np.prod(shape)
    v_eff_g = np.asarray([average], dtype=complex)                  # pw/scf.py:3288
    operator = scf.PlaneWaveHamiltonian(
        system.basis, v_eff_g, terms, local_workspace=workspace, timers=timers,
        real_potential=v_native, native_potential_layout=True,
        potential_average=average + v_ion_shift)
    trial_vectors, trial_eigenvalues, trial_applied = rotate_starting_subspace(
        scf, operator, starting, nbnd, mpi)
    iterations = applications = attempt = 0
    retry_vectors, retry_eigenvalues, retry_applied = (trial_vectors, trial_eigenvalues,
                                                       trial_applied)
    while True:
        solution = scf.davidson(
            operator.apply, operator.diagonal, nbnd, initial_vectors=retry_vectors,
            tolerance=tolerance, max_iterations=20, subspace_multiplier=ndim,
            residual_factor=None, residual_energy_scale=None, occupied_roots=nbnd,
            initial_is_ritz=True, initial_is_orthonormal=True,
            initial_eigenvalues=retry_eigenvalues, initial_applied=retry_applied, mpi=mpi,
            global_dimension=len(system.basis.indices), global_row_indices=operator.local_rows,
            timers=timers, operator_into=operator.apply_into)
        attempt += 1
        iterations += solution.iterations
        applications += solution.hamiltonian_applications
        if solution.converged or solution.number_unconverged == 0 or attempt >= max_attempts:
            break
        retry_vectors = solution.eigenvectors
        retry_eigenvalues = solution.eigenvalues
        retry_applied = None
    return operator, solution, attempt, iterations, applications
