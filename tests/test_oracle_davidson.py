"""Pinning the oracle Davidson (CPU).

* small workload: eigenvalues against dense ``eigvalsh`` of the brute-force
  Hamiltonian (golden fixture), residuals recomputed with the dense matrix;
* the reference's own solver-test pattern
  (``tests/unit/core/test_diagonalization.py:8-31``): random 18x18 symmetric
  matrix (seed 7321) + diag(linspace(1, 18)), lowest 4 vs eigvalsh at 1e-9,
  residual < 3e-6;
* argument / linear-dependence errors (``diagonalization.py:558-622``) and
  non-convergence at max_iterations=1 (not an exception);
* a planted mistake (wrong preconditioner sign) measurably slows or breaks
  convergence, so the convergence test has teeth.
"""

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import davidson as odav
from oracle.hpsi import OracleHamiltonian
from qe_python_amd.synthetic import make_system

GOLDEN = Path(__file__).parent / "golden"


def reference_matrix():
    rng = np.random.default_rng(7321)
    a = rng.standard_normal((18, 18))
    return 0.5 * (a + a.T) + np.diag(np.linspace(1.0, 18.0, 18))


def test_reference_pattern_matrix():
    h = reference_matrix()
    trial = np.eye(18, 6, dtype=complex) + 0.01
    res = odav.davidson(lambda x: h @ x, np.diag(h), 4, initial_vectors=trial,
                        tolerance=1e-11, max_iterations=60)
    assert res.converged
    np.testing.assert_allclose(res.eigenvalues, np.linalg.eigvalsh(h)[:4], atol=1e-9)
    resid = h @ res.eigenvectors - res.eigenvectors * res.eigenvalues[None, :]
    assert np.max(np.linalg.norm(resid, axis=0)) < 3e-6


@pytest.fixture(scope="module")
def small_ham():
    s = make_system("small")
    beta, coup, diag = s.packed_term()
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                            (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    return s, ham


def test_small_workload_matches_dense_eigenvalues(small_ham):
    s, ham = small_ham
    exact = np.load(GOLDEN / "small_hpsi.npz")["dense_eigenvalues"]
    res = odav.davidson(ham.apply, ham.diagonal, 7, initial_vectors=s.trial,
                        tolerance=0.5e-10, operator_into=ham.apply_into)
    assert res.converged
    np.testing.assert_allclose(res.eigenvalues, exact[:7], atol=1e-11, rtol=0)
    r = ham.apply(res.eigenvectors) - res.eigenvectors * res.eigenvalues[None, :]
    np.testing.assert_allclose(np.linalg.norm(r, axis=0), res.residual_norms, rtol=1e-6,
                               atol=1e-12)
    assert np.all(res.residual_norms < 10 * 0.5e-10)
    # orthonormal Ritz vectors
    gram = res.eigenvectors.conj().T @ res.eigenvectors
    np.testing.assert_allclose(gram, np.eye(7), atol=1e-10)


def test_sixteen_roots_without_residual_test_reach_dense_values(small_ham):
    # 8 roots would cut the near-degenerate cluster 0.7448/0.7470/0.7472 Ry,
    # where this algorithm stalls (40 iterations leave a 2e-5 Ha error); 16
    # roots end above the cluster.
    s, ham = small_ham
    exact = np.load(GOLDEN / "small_hpsi.npz")["dense_eigenvalues"]
    res = odav.davidson(ham.apply, ham.diagonal, 16, initial_vectors=s.trial,
                        tolerance=0.5e-10, residual_factor=None, max_iterations=40)
    assert res.converged
    np.testing.assert_allclose(res.eigenvalues, exact[:16], atol=1e-10, rtol=0)


def test_planted_preconditioner_sign_error_breaks_convergence(small_ham, monkeypatch):
    s, ham = small_ham
    good = odav.qe_precondition
    monkeypatch.setattr(odav, "qe_precondition",
                        lambda r, e, d: good(r, -np.asarray(e), d))
    res = odav.davidson(ham.apply, ham.diagonal, 7, initial_vectors=s.trial,
                        tolerance=0.5e-10, max_iterations=12)
    assert not res.converged


def test_argument_errors():
    h = reference_matrix()
    op = lambda x: h @ x  # noqa: E731
    d = np.diag(h)
    with pytest.raises(ValueError):
        odav.davidson(op, d, 0)
    with pytest.raises(ValueError):
        odav.davidson(op, d, 2, tolerance=0.0)
    with pytest.raises(ValueError):
        odav.davidson(op, d, 2, subspace_multiplier=1)
    with pytest.raises(ValueError):
        odav.davidson(op, d, 2, initial_vectors=np.ones((18, 1)))
    # rank 1: exact duplicates would not do, the Gram eigensolve's rounding
    # (~1e-16 relative) clears the 1e-24 relative rank threshold
    dep = np.zeros((18, 3), dtype=complex)
    dep[:, 0] = 1.0
    with pytest.raises(ValueError, match="linearly dependent"):
        odav.davidson(op, d, 2, initial_vectors=dep)
    # 14 initial vectors for 4 roots fill 14 of the 16-vector subspace: the
    # first 4-column correction block has no room (the reference's
    # basis_storage assignment, qepy_pw/diagonalization.py:791, raises)
    start = np.eye(18, 14, dtype=complex) + 0.01
    with pytest.raises(ValueError):
        odav.davidson(op, d, 4, initial_vectors=start)
    assert odav.davidson(op, d, 4, initial_vectors=start[:, :12]).iterations >= 1


def test_non_convergence_is_a_result_not_an_error():
    h = reference_matrix()
    res = odav.davidson(lambda x: h @ x, np.diag(h), 4, max_iterations=1,
                        initial_vectors=np.eye(18, 4, dtype=complex) + 0.1)
    assert not res.converged and res.number_unconverged > 0 and res.iterations == 1
