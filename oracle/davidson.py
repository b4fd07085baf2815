"""Oracle block Davidson: serial restatement of the reference eigensolver.

Follows, as text, ``qepy_pw/diagonalization.py``:

* ``davidson``               -- 527-880 (QE ``cegterg``): argument checks,
  per-root tolerances (occupied: tol, others max(5 tol, 5e-6)), canonical
  orthonormalisation of the trials, storage of ``maximum_subspace`` columns,
  ``update_projected`` (658-697), ``ritz_pairs`` (699-726), the convergence
  tests (811-828) and the restart (843-868).
* ``orthonormalize``         -- ``_orthonormalize`` serial branch (450-472).
* ``qe_precondition``        -- ``_qe_precondition`` (475-497) with the
  ``_native_fft.pyx:852-899`` formula x=2(h-e), d=(1+x+sqrt(1+(x-1)^2))/2,
  out = 2 r / d.
* ``normalize_columns``      -- ``_normalize_columns`` (500-524) and
  ``normalize_selected_columns`` (``_native_fft.pyx:1495-1530``).
* ``lowest_generalized_eigh``-- ``generalized_eigh`` (``_native_fft.pyx:667-691``):
  LAPACK zhegvd(itype=1, 'V', 'L'), lowest ``roots`` pairs.

Pinned by (tests/test_oracle_davidson.py): dense ``eigh`` of the brute-force
Hamiltonian of the small workload (eigenvalues to 1e-11 Ha), residual norms
recomputed independently, the reference's own ParO-style test pattern
(``tests/unit/core/test_diagonalization.py:8-31``: random 18x18 symmetric
plus diag(1..18), lowest 4 vs eigvalsh), the linear-dependence and argument
``ValueError`` paths, and non-convergence at ``max_iterations=1``.
Iteration counts / H-application counts are not pinned against the
reference's own run (its extension cannot be built here): parity unpinned
for those two integers.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla


@dataclass(frozen=True)
class OracleResult:
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray
    residual_norms: np.ndarray
    iterations: float
    converged: bool
    number_unconverged: int = 0
    hamiltonian_applications: int = 0


def lowest_generalized_eigh(h, s, roots):
    values, vectors = sla.eigh(np.asarray(h, complex), np.asarray(s, complex),
                               lower=True, type=1, driver="gvd",
                               overwrite_a=False, overwrite_b=False,
                               check_finite=False)
    return values[:roots].copy(), np.asfortranarray(vectors[:, :roots])


def orthonormalize(vectors, threshold=1.0e-12):
    block = np.asarray(vectors, dtype=complex).copy()
    if block.ndim == 1:
        block = block[:, None]
    if block.shape[1] == 0:
        return np.empty((block.shape[0], 0), dtype=complex)
    if block.shape[1] == 1:
        norm = float(np.linalg.norm(block[:, 0]))
        if norm == 0.0:
            return np.empty((block.shape[0], 0), dtype=complex)
        return block / norm
    gram = block.conj().T @ block
    gram = 0.5 * (gram + gram.conj().T)
    squared, rotation = np.linalg.eigh(gram)
    order = np.argsort(squared)[::-1]
    squared = np.maximum(np.real(squared[order]), 0.0)
    if not len(squared) or squared[0] == 0.0:
        return np.empty((block.shape[0], 0), dtype=complex)
    rank = int(np.count_nonzero(squared > threshold ** 2 * squared[0]))
    transform = rotation[:, order[:rank]]
    transform /= np.sqrt(squared[:rank])[None, :]
    return block @ transform


def qe_precondition(residuals, eigenvalues, diagonal):
    x = 2.0 * (np.asarray(diagonal)[:, None] - np.asarray(eigenvalues)[None, :])
    d = 0.5 * (1.0 + x + np.sqrt(1.0 + (x - 1.0) ** 2))
    return np.asfortranarray(2.0 * residuals / d)


def column_squared_norms(block):
    return np.einsum("gb,gb->b", block.real, block.real) + np.einsum(
        "gb,gb->b", block.imag, block.imag)


def normalize_columns(vectors, threshold=1.0e-14):
    block = np.asarray(vectors, dtype=complex)
    if block.ndim == 1:
        block = block[:, None]
    sq = column_squared_norms(block)
    if not len(sq):
        return np.empty((block.shape[0], 0), dtype=complex)
    scale = max(float(np.max(sq)), 1.0e-300)
    keep = sq > threshold ** 2 * scale
    if not np.any(keep):
        return np.empty((block.shape[0], 0), dtype=complex)
    sel = np.flatnonzero(keep)
    return np.asfortranarray(block[:, sel] / np.sqrt(sq[sel])[None, :])


def davidson(operator, diagonal, number_of_roots, *, initial_vectors=None,
             tolerance=1.0e-8, max_iterations=40, subspace_multiplier=4,
             residual_factor=10.0, residual_energy_scale=None, occupied_roots=None,
             initial_is_ritz=False, initial_is_orthonormal=False,
             initial_eigenvalues=None, initial_applied=None,
             operator_into=None, trace=None):
    diagonal = np.asarray(diagonal, dtype=float)
    dim = len(diagonal)
    if not 1 <= number_of_roots <= dim:
        raise ValueError("number_of_roots must be between 1 and the matrix size")
    if tolerance <= 0.0:
        raise ValueError("Davidson tolerance must be positive")
    if max_iterations < 1:
        raise ValueError("Davidson max_iterations must be at least 1")
    if subspace_multiplier < 2:
        raise ValueError("diago_david_ndim must be at least 2")
    if residual_factor is not None and residual_factor <= 0.0:
        raise ValueError("Davidson residual_factor must be positive or None")
    if residual_energy_scale is not None and residual_energy_scale <= 0.0:
        raise ValueError("Davidson residual_energy_scale must be positive or None")
    if occupied_roots is None:
        occupied_roots = number_of_roots
    if not 0 <= occupied_roots <= number_of_roots:
        raise ValueError("occupied_roots must be between zero and number_of_roots")
    tol = np.full(number_of_roots, max(5.0 * tolerance, 5.0e-6))
    tol[:occupied_roots] = tolerance

    if initial_vectors is None:
        trial = np.zeros((dim, number_of_roots), dtype=complex)
        for c in range(number_of_roots):
            trial[c, c] = 1.0
    else:
        trial = np.asarray(initial_vectors, dtype=complex)
        if trial.ndim != 2 or trial.shape[0] != dim or trial.shape[1] < number_of_roots:
            raise ValueError("initial_vectors must have shape "
                             "(matrix size, at least number_of_roots)")
    if initial_is_ritz or initial_is_orthonormal:
        basis = np.asarray(trial[:, :number_of_roots], dtype=complex).copy()
    else:
        basis = orthonormalize(trial)
    if basis.shape[1] < number_of_roots:
        raise ValueError("initial Davidson vectors are linearly dependent")
    maxsub = min(dim, max(basis.shape[1], 2 * number_of_roots,
                          subspace_multiplier * number_of_roots))
    active = basis.shape[1]
    V = np.empty((dim, maxsub), dtype=complex, order="F")
    HV = np.empty_like(V)
    V[:, :active] = basis

    def apply_into(src, dst):
        if operator_into is not None:
            operator_into(src, dst)
        else:
            dst[...] = operator(src)

    if initial_applied is None:
        apply_into(V[:, :active], HV[:, :active])
        happ = active
    else:
        sup = np.asarray(initial_applied, dtype=np.complex128)
        if sup.shape != (dim, active):
            raise ValueError("initial_applied must match the initial Ritz block")
        HV[:, :active] = sup
        happ = 0
    Hs = np.zeros((maxsub, maxsub), dtype=complex)
    Ss = np.zeros_like(Hs)

    def update_projected(first):
        nb = V[:, first:active]
        h_rows = nb.conj().T @ HV[:, :active]
        s_rows = nb.conj().T @ V[:, :active]
        Hs[first:active, :active] = h_rows
        Ss[first:active, :active] = s_rows
        if first:
            Hs[:first, first:active] = h_rows[:, :first].conj().T
            Ss[:first, first:active] = s_rows[:, :first].conj().T
        sl = slice(first, active)
        Hs[sl, sl] = 0.5 * (Hs[sl, sl] + Hs[sl, sl].conj().T)
        Ss[sl, sl] = 0.5 * (Ss[sl, sl] + Ss[sl, sl].conj().T)

    update_projected(0)

    def ritz_pairs():
        roots, coef = lowest_generalized_eigh(Hs[:active, :active], Ss[:active, :active],
                                              number_of_roots)
        X = np.asfortranarray(V[:, :active] @ coef)
        R = np.asfortranarray(HV[:, :active] @ coef)
        R -= X * roots[None, :]
        return roots, X, R

    if initial_is_ritz:
        if initial_eigenvalues is None:
            values = np.real(np.einsum("gb,gb->b", V[:, :number_of_roots].conj(),
                                       HV[:, :number_of_roots]))
        else:
            values = np.asarray(initial_eigenvalues[:number_of_roots], float)
        X = V[:, :number_of_roots].copy(order="F")
        R = np.asfortranarray(HV[:, :number_of_roots] - X * values[None, :])
    else:
        values, X, R = ritz_pairs()
    track = residual_factor is not None or residual_energy_scale is not None

    def norms(block):
        return np.sqrt(np.real(column_squared_norms(block)))

    rnorm = norms(R) if track else np.empty(number_of_roots)
    done_iter = 0
    nunconv = number_of_roots
    unconv = np.ones(number_of_roots, dtype=bool)
    for it in range(1, max_iterations + 1):
        done_iter = it
        prev = values
        corr = qe_precondition(R[:, unconv], values[unconv], diagonal)
        corr = normalize_columns(corr)
        if corr.shape[1] == 0:
            break
        old = active
        active += corr.shape[1]
        V[:, old:active] = corr
        apply_into(V[:, old:active], HV[:, old:active])
        happ += corr.shape[1]
        update_projected(old)
        values, X, R = ritz_pairs()
        if track:
            rnorm = norms(R)
        conv = np.abs(values - prev) < tol
        if residual_factor is not None:
            conv &= rnorm < residual_factor * tol
        if residual_energy_scale is not None:
            conv &= rnorm ** 2 < residual_energy_scale * tol
        unconv = ~conv
        nunconv = int(np.count_nonzero(unconv))
        if trace is not None:
            trace.append({"iteration": it, "active": active, "nunconv": nunconv,
                          "values": values.copy()})
        if nunconv == 0:
            if not track:
                rnorm = norms(R)
            return OracleResult(values, X, rnorm, it, True, 0, happ)
        if active + nunconv > maxsub:
            active = number_of_roots
            V[:, :active] = X
            HV[:, :active] = X * values[None, :] + R
            Hs[:active, :active] = 0.0
            Hs[np.arange(active), np.arange(active)] = values
            Ss[:active, :active] = np.eye(active, dtype=complex)
    if not track:
        rnorm = norms(R)
    return OracleResult(values, X, rnorm, done_iter, False, nunconv, happ)
