"""Pinning the oracle CG, ParO and RMM-DIIS (CPU).

* the reference's own ParO tests (``tests/unit/core/test_diagonalization.py:
  8-50``): 18x18 symmetric (seed 7321) + diag(linspace(1, 18)), lowest 4 vs
  ``eigvalsh`` at 1e-9 with residual < 3e-6; more than 2 x roots H
  applications for the diagonal 12x12 case (bpcg runs an inner trajectory);
* converged CG and ParO on the small plane-wave workload against the dense
  eigenvalues of its brute-force Hamiltonian (golden fixture);
* exact eigenvectors are a fixed point of every solver;
* RMM-DIIS called repeatedly from its own output, as the SCF does, converges
  to the dense eigenvalues;
* a planted mistake (the Kresse-Furthmueller sign flipped) stops that
  convergence, so the test has teeth;
* the reference's argument errors.
"""

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import iterative as oit
from oracle.hpsi import OracleHamiltonian
from qe_python_amd.synthetic import make_system

GOLDEN = Path(__file__).parent / "golden"


def reference_matrix():
    rng = np.random.default_rng(7321)
    m = rng.normal(size=(18, 18))
    return rng, 0.5 * (m + m.T) + np.diag(np.linspace(1.0, 18.0, 18))


def test_paro_reference_matrix():
    rng, h = reference_matrix()
    initial = rng.normal(size=(18, 4))
    res = oit.parallel_orbital(lambda x: h @ x, np.diag(h), 4, initial_vectors=initial,
                               tolerance=1.0e-10, max_iterations=20)
    assert res.converged and res.number_unconverged == 0
    np.testing.assert_allclose(res.eigenvalues, np.linalg.eigvalsh(h)[:4], atol=1e-9)
    assert np.max(res.residual_norms) < 3.0e-6


def test_paro_reference_application_count():
    rng = np.random.default_rng(91)
    d = np.arange(1.0, 13.0)
    res = oit.parallel_orbital(lambda x: d[:, None] * x, d, 3,
                               initial_vectors=rng.normal(size=(12, 3)), tolerance=1e-8,
                               max_iterations=2)
    assert res.hamiltonian_applications > 2 * 3


@pytest.fixture(scope="module")
def small():
    s = make_system("small")
    beta, coup, diag = s.packed_term()
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                            (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    exact = np.load(GOLDEN / "small_hpsi.npz")["dense_eigenvalues"]
    return s, ham, exact


def test_cg_and_paro_small_workload(small):
    s, ham, exact = small
    res = oit.conjugate_gradient(ham.apply, ham.diagonal, 7, initial_vectors=s.trial[:, :7],
                                 tolerance=1e-12, max_iterations=200)
    assert res.converged
    np.testing.assert_allclose(res.eigenvalues, exact[:7], rtol=0, atol=1e-9)
    res = oit.parallel_orbital(ham.apply, ham.diagonal, 7, initial_vectors=s.trial[:, :7],
                               tolerance=1e-12, max_iterations=100)
    assert res.converged
    np.testing.assert_allclose(res.eigenvalues, exact[:7], rtol=0, atol=1e-9)
    # orthonormal output (ParO rotates by the generalized Ritz vectors)
    v = res.eigenvectors
    np.testing.assert_allclose(v.conj().T @ v, np.eye(7), atol=1e-10)


def exact_pairs(n=30, k=5, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    h = 0.5 * (a + a.conj().T) + np.diag(np.arange(n, dtype=float) * 2.0)
    w, u = np.linalg.eigh(h)
    return h, w[:k], u[:, :k]


@pytest.mark.parametrize("solver", ["cg", "paro", "rmm"])
def test_exact_eigenvectors_are_a_fixed_point(solver):
    h, w, u = exact_pairs()
    op = lambda x: h @ x
    if solver == "cg":
        res = oit.conjugate_gradient(op, np.real(np.diag(h)), 5, initial_vectors=u,
                                     tolerance=1e-10)
    elif solver == "paro":
        res = oit.parallel_orbital(op, np.real(np.diag(h)), 5, initial_vectors=u,
                                   tolerance=1e-10)
    else:
        res = oit.rmm_diis(op, np.real(np.diag(h)), 5, initial_vectors=u, tolerance=1e-10)
    np.testing.assert_allclose(res.eigenvalues, w, atol=1e-11)
    assert np.max(res.residual_norms) < 1e-9


def repeated_rmm(h, k, sweeps):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(h.shape[0], k)) + 1j * rng.normal(size=(h.shape[0], k))
    diag = np.real(np.diag(h)) * 0.5          # kinetic in Ha, so T_ry = diag(h)
    for _ in range(sweeps):
        res = oit.rmm_diis(lambda x: h @ x, diag, k, initial_vectors=v, tolerance=1e-13)
        v = res.eigenvectors
    return res


def test_repeated_rmm_converges_and_a_planted_sign_error_does_not(monkeypatch):
    rng = np.random.default_rng(11)
    n = 60
    a = 0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    h = 0.5 * (a + a.conj().T) + np.diag(np.linspace(1.0, 40.0, n))
    exact = np.linalg.eigvalsh(h)[:4]
    res = repeated_rmm(h, 4, 12)
    np.testing.assert_allclose(res.eigenvalues, exact, atol=1e-9)
    good = oit.kresse_furthmueller
    monkeypatch.setattr(oit, "kresse_furthmueller", lambda r, t, e: -good(r, t, e))
    planted = repeated_rmm(h, 4, 12)
    assert np.max(np.abs(planted.eigenvalues - exact)) > 1e-6


def test_argument_errors():
    h = np.diag(np.arange(1.0, 9.0))
    op = lambda x: h @ x
    x0 = np.eye(8, 3)
    with pytest.raises(ValueError):
        oit.parallel_orbital(op, np.diag(h), 3, initial_vectors=x0, tolerance=0.0)
    with pytest.raises(ValueError):
        oit.conjugate_gradient(op, np.diag(h), 3, initial_vectors=x0, max_iterations=0)
    with pytest.raises(ValueError):
        oit.rmm_diis(op, np.diag(h), 3, initial_vectors=x0, history_dimension=1)
    with pytest.raises(ValueError):
        oit.rmm_diis(op, np.diag(h), 3, initial_vectors=x0, gram_schmidt_block=0)
    with pytest.raises(ValueError, match="at least"):
        oit.conjugate_gradient(op, np.diag(h), 4, initial_vectors=x0)
    with pytest.raises(ValueError, match="dependent"):
        oit.parallel_orbital(op, np.diag(h), 3, initial_vectors=np.ones((8, 3)))
