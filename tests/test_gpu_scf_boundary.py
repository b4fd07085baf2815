"""The drop-in as the reference SCF uses it.

A stand-in SCF module holds the names ``pw/scf.py:29-40`` imports, bound by
value to host implementations; ``qe_python_amd.integration.bind`` rebinds
them, and one k-point is driven through the SCF's sequence
(``tests/scf_standin.py``): a length-1 ``v_eff_g``, a workspace carrying only
the reference's attributes, the native potential layout, a
``potential_average`` (the ionic G=0 term) that differs from <V>, the
starting-subspace rotation, and Davidson with ``initial_is_ritz=True`` and
its retry loop.  The device run must reach the host run's outcome:
same convergence, eigenvalues within the solver tolerance of each other and
of the dense eigenvalues.
"""

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

import scf_standin as standin
from qe_python_amd.synthetic import make_system, phase_matrix

GOLDEN = Path(__file__).parent / "golden"
TOL_HA = 0.5e-10


def _terms(s, kind, mod):
    if kind == "packed":
        beta, coup, diag = s.packed_term()
        return (mod.PackedProjectorTerm(beta, coup, diag),)
    # one factorized term per atom: the synthetic couplings differ per atom
    phases = phase_matrix(s.basis.vectors, s.projectors.positions)
    return tuple(mod.FactorizedProjectorTerm(s.projectors.beta_species,
                                             np.asfortranarray(phases[:, a:a + 1]), c)
                 for a, c in enumerate(s.projectors.couplings))


def _host_module():
    from qe_python_amd import diagonalization as dev
    # the reference's term dataclasses have the same fields; the host module
    # binds them by value like every other name
    return standin.make_scf_module(dev.PackedProjectorTerm, dev.FactorizedProjectorTerm)


def test_bind_rebinds_every_scf_name_and_unbind_restores():
    from qe_python_amd import diagonalization as dev
    from qe_python_amd import integration
    scf = _host_module()
    before = {name: getattr(scf, name) for name in integration.SCF_NAMES}
    previous = integration.bind(scf)
    for name in integration.SCF_NAMES:
        assert getattr(scf, name) is getattr(dev, name), name
    integration.unbind(scf, previous)
    for name in integration.SCF_NAMES:
        assert getattr(scf, name) is before[name], name
    previous = integration.bind(scf, precision="mixed")
    assert issubclass(scf.PlaneWaveHamiltonian, dev.PlaneWaveHamiltonian)
    assert scf.davidson is dev.davidson
    integration.unbind(scf, previous)
    with pytest.raises(ValueError):
        integration.bind(scf, precision="fp32")


@pytest.fixture(scope="module")
def small():
    s = make_system("small")
    exact = np.load(GOLDEN / "small_hpsi.npz")["dense_eigenvalues"]
    return s, exact


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["packed", "factorized"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_scf_kpoint_through_the_rebound_names(small, kind, precision):
    from qe_python_amd import diagonalization as dev
    from qe_python_amd import integration
    s, exact = small
    nbnd = 7
    starting = s.trial[:, :16]
    scf = _host_module()
    # diago_david_ndim = 2: the first 20-iteration attempt leaves roots
    # unconverged, so the retry (initial_is_ritz, initial_applied=None) runs
    h_op, h_sol, h_attempts, h_iter, _ = standin.solve_kpoint(
        scf, s, _terms(s, kind, scf), nbnd, starting, TOL_HA, ndim=2)

    timers = SimpleNamespace(labels=[])
    timers.measure = lambda label: _Label(timers.labels, label)
    previous = integration.bind(scf, precision=precision)
    try:
        d_op, d_sol, d_attempts, d_iter, _ = standin.solve_kpoint(
            scf, s, _terms(s, kind, scf), nbnd, starting, TOL_HA, ndim=2, timers=timers)
    finally:
        integration.unbind(scf, previous)
    assert isinstance(d_op, dev.PlaneWaveHamiltonian)
    assert d_op.precision == precision
    # the preconditioner diagonal carries potential_average (v_ion), not <V>
    np.testing.assert_allclose(d_op.diagonal, h_op.diagonal, rtol=0, atol=1e-12)
    # apply_into on host arrays with the native-layout potential
    x = np.asfortranarray(starting[:, :4])
    out_d = np.empty_like(x, order="F")
    d_op.apply_into(x, out_d)
    ref = h_op.apply(x)
    assert np.linalg.norm(out_d - ref) <= 1e-12 * np.linalg.norm(ref)
    print(f"{kind}/{precision}: host {h_attempts} attempts / {h_iter} iterations, "
          f"device {d_attempts} / {d_iter}")
    assert h_attempts >= 2 and d_attempts >= 2
    assert h_sol.converged and d_sol.converged
    assert d_sol.eigenvectors.shape == (s.npw, nbnd)
    np.testing.assert_allclose(d_sol.eigenvalues, h_sol.eigenvalues, atol=2 * TOL_HA, rtol=0)
    np.testing.assert_allclose(d_sol.eigenvalues, exact[:nbnd], atol=TOL_HA, rtol=0)
    assert {"cdiaghg", "g_psi", "vloc_psi"} <= set(timers.labels)


class _Label:
    def __init__(self, sink, label):
        self.sink, self.label = sink, label

    def __enter__(self):
        self.sink.append(self.label)

    def __exit__(self, *exc):
        return False
