"""Pinning the oracle gradients (CPU) by finite differences and identities.

* tau: central differences of the restated nonlocal energy
  (``pw/scf.py:1521-1631``), h = 1e-5 bohr;
* V(r): E is linear in V, so a central difference of the full band energy at
  a grid point equals dE/dV there up to rounding;
* psi: E is quadratic in psi, so a central difference along a random
  direction equals Re <grad, d> up to rounding;
* the nonlocal energy of the force routine equals the projector part of the
  oracle H expectation value (independent code path).
"""

import numpy as np
import pytest

from oracle.forces import band_energy_gradients, nonlocal_energy_and_forces, packed_projectors
from oracle.hpsi import OracleHamiltonian
from qe_python_amd.synthetic import make_system


@pytest.fixture(scope="module")
def case():
    s = make_system("small")
    rng = np.random.default_rng(4)
    nb = 6
    psi = s.trial[:, :nb] * 3.0
    f = rng.uniform(0.5, 2.0, nb)
    return s, psi, f


def energy(s, psi, f, vr=None, positions=None):
    positions = s.projectors.positions if positions is None else positions
    beta, coup = packed_projectors(s.basis.vectors, s.projectors.beta_species,
                                   s.projectors.couplings, positions)

    class T:
        pass
    t = T()
    t.beta, t.coupling, t.diagonal = beta, coup, np.zeros(s.npw)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape,
                            s.real_potential if vr is None else vr, (t,))
    return float(np.real(np.einsum("gb,gb,b->", psi.conj(), ham.apply(psi), f)))


def test_nonlocal_energy_matches_projector_expectation(case):
    s, psi, f = case
    e_nl, _ = nonlocal_energy_and_forces(psi, f, s.basis.vectors, s.projectors.beta_species,
                                         s.projectors.couplings, s.projectors.positions)
    beta, coup = packed_projectors(s.basis.vectors, s.projectors.beta_species,
                                   s.projectors.couplings, s.projectors.positions)
    o = beta.conj().T @ psi
    assert e_nl == pytest.approx(float(np.real(np.einsum("pb,pq,qb,b->", o.conj(), coup, o, f))),
                                 rel=1e-12)


def test_forces_match_central_differences(case):
    s, psi, f = case
    _, forces = nonlocal_energy_and_forces(psi, f, s.basis.vectors, s.projectors.beta_species,
                                           s.projectors.couplings, s.projectors.positions)
    h = 1e-5
    for a in range(len(s.projectors.positions)):
        for c in range(3):
            plus = s.projectors.positions.copy()
            minus = s.projectors.positions.copy()
            plus[a, c] += h
            minus[a, c] -= h
            ep, _ = nonlocal_energy_and_forces(psi, f, s.basis.vectors, s.projectors.beta_species,
                                               s.projectors.couplings, plus)
            em, _ = nonlocal_energy_and_forces(psi, f, s.basis.vectors, s.projectors.beta_species,
                                               s.projectors.couplings, minus)
            fd = -(ep - em) / (2 * h)
            assert fd == pytest.approx(forces[a, c], rel=1e-6, abs=1e-9)


def test_band_energy_gradients_against_differences(case):
    s, psi, f = case
    e, dv, dtau, gpsi = band_energy_gradients(
        s.basis.indices, s.basis.kinetic, s.basis.vectors, s.shape, s.real_potential,
        s.projectors.beta_species, s.projectors.couplings, s.projectors.positions, psi, f)
    assert e == pytest.approx(energy(s, psi, f), rel=1e-13)
    # E (~1e4 Ha here) is linear in V, so a large step is exact and keeps the
    # difference well above the energy's rounding
    h = 0.1
    for idx in [(0, 0, 0), (5, 17, 30), (20, 3, 11)]:
        vp = s.real_potential.copy()
        vm = s.real_potential.copy()
        vp[idx] += h
        vm[idx] -= h
        fd = (energy(s, psi, f, vr=vp) - energy(s, psi, f, vr=vm)) / (2 * h)
        assert fd == pytest.approx(dv[idx], rel=1e-8, abs=1e-12)
    rng = np.random.default_rng(5)
    d = rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape)
    hp = 1e-2  # E is quadratic in psi: the central difference is exact
    fd = (energy(s, psi + hp * d, f) - energy(s, psi - hp * d, f)) / (2 * hp)
    assert fd == pytest.approx(float(np.real(np.vdot(gpsi, d))), rel=1e-9)
    # position gradient = -forces, consistent with a step of the full energy
    hx = 1e-4
    plus = s.projectors.positions.copy()
    minus = s.projectors.positions.copy()
    plus[1, 2] += hx
    minus[1, 2] -= hx
    fd = (energy(s, psi, f, positions=plus) - energy(s, psi, f, positions=minus)) / (2 * hx)
    assert fd == pytest.approx(dtau[1, 2], rel=1e-5, abs=1e-8)
