"""Pinning the oracle h_psi (CPU): every check is independent of the FFT
sandwich it tests.

* known answer of the reference (``tests/unit/core/test_basis.py:76-89``);
* dense brute force H = V(G-G') + T + B D B^H on the small workload;
* Hermiticity <x|Hy> = <Hx|y>*;
* identity potential: V = 1 => H psi = (T + 1) psi (the round trip check of
  ``tools/benchmark_distributed_fft.py:58-83``);
* constant V, no projectors: plane waves are eigenvectors with T + V0;
* factorized and packed projector terms give the same H;
* golden fixture of the small workload (inputs hash, h_psi columns).
"""

from types import SimpleNamespace

import numpy as np
import pytest

from oracle.dense import dense_hamiltonian, potential_matrix
from oracle.hpsi import OracleHamiltonian, apply_local, linear_slots, prepare_potential
from qe_python_amd.synthetic import make_system, phase_matrix

GOLDEN = __import__("pathlib").Path(__file__).parent / "golden"


def kat_inputs():
    shape = (6, 6, 6)
    grid = np.indices(shape).sum(axis=0)
    vr = 0.3 + 0.2 * np.cos(2.0 * np.pi * grid / shape[0])
    vg = np.fft.fftn(vr) / np.prod(shape)
    idx = np.asarray([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0]], dtype=np.int32)
    c = np.asarray([1.0 + 0.2j, -0.3j, 0.4, -0.1 + 0.5j])
    return shape, vg, idx, c


def test_reference_known_answer_local_potential():
    shape, vg, idx, c = kat_inputs()
    expected = potential_matrix(vg, idx) @ c
    vr = prepare_potential(vg)
    actual = apply_local(vr, linear_slots(idx, shape), shape, c)
    np.testing.assert_allclose(actual, expected, atol=2.0e-14)


def test_planted_sign_error_is_caught_by_brute_force(monkeypatch, small):
    # The known answer's V(r) is even, so it cannot see swapped transform
    # directions; the dense brute force on the asymmetric small workload can.
    import oracle.hpsi as oh
    s, _, dense, term = small
    fftn, ifftn = oh.sfft.fftn, oh.sfft.ifftn

    class Swapped:
        @staticmethod
        def ifftn(x, axes=None, norm=None, workers=None):
            return fftn(x, axes=axes, workers=workers)

        @staticmethod
        def fftn(x, axes=None, norm=None, workers=None):
            return ifftn(x, axes=axes, norm="forward", workers=workers)

    monkeypatch.setattr(oh, "sfft", Swapped)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential, (term,))
    y = ham.apply(s.trial[:, :2])
    ref = dense @ s.trial[:, :2]
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) > 1e-3


@pytest.fixture(scope="module")
def small():
    s = make_system("small")
    beta, coup, diag = s.packed_term()
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                            (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    dense = dense_hamiltonian(s.basis.indices, s.basis.kinetic, s.real_potential, (term,))
    return s, ham, dense, term


def test_hpsi_matches_dense_brute_force(small):
    s, ham, dense, _ = small
    y = ham.apply(s.trial)
    ref = dense @ s.trial
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13


def test_hpsi_hermitian(small):
    s, ham, _, _ = small
    x, y = s.trial[:, :3], s.trial[:, 3:6]
    lhs = x.conj().T @ ham.apply(y)
    rhs = (ham.apply(x).conj().T @ y)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_identity_potential_round_trip(small):
    s, _, _, _ = small
    ones = np.ones(s.shape)
    out = apply_local(ones, linear_slots(s.basis.indices, s.shape), s.shape, s.trial[:, :4],
                      s.basis.kinetic)
    np.testing.assert_allclose(out, (s.basis.kinetic + 1.0)[:, None] * s.trial[:, :4],
                               atol=1e-13)


def test_constant_potential_plane_wave_eigenstates(small):
    s, _, _, _ = small
    v0 = -0.37
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, np.full(s.shape, v0))
    e = np.zeros((s.npw, 3), dtype=complex)
    e[[0, 7, 100], [0, 1, 2]] = 1.0
    out = ham.apply(e)
    expect = e * (s.basis.kinetic[[0, 7, 100]] + v0)[None, :]
    np.testing.assert_allclose(out, expect, atol=1e-13)


def test_factorized_packed_and_materialized_terms_agree(small):
    s, _, _, _ = small
    pos = s.projectors.positions
    coupling = s.projectors.couplings[0]
    beta_s = s.projectors.beta_species
    fact = SimpleNamespace(beta=beta_s, phases=phase_matrix(s.basis.vectors, pos),
                           coupling=coupling)
    mats = tuple((beta_s * np.exp(-1j * (s.basis.vectors @ p))[:, None], coupling) for p in pos)
    x = s.trial[:, :4]
    h_f = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential, (fact,))
    h_m = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential, mats)
    np.testing.assert_allclose(h_f.apply(x), h_m.apply(x), atol=1e-13)
    np.testing.assert_allclose(h_f.diagonal, h_m.diagonal, atol=1e-13)


def test_packed_diagonal_matches_dense_diagonal(small):
    s, ham, dense, _ = small
    avg = float(np.real(s.potential_g.flat[0]))
    # H_GG = T + V(0) + (B D B^H)_GG exactly; the packed diagonal must add that term
    np.testing.assert_allclose(ham.diagonal, np.real(np.diag(dense)) - np.real(
        potential_matrix(np.fft.fftn(s.real_potential) / s.real_potential.size,
                         s.basis.indices).diagonal()) + avg, atol=1e-12)


def test_golden_small_workload(small):
    s, ham, _, _ = small
    data = np.load(GOLDEN / "small_hpsi.npz")
    np.testing.assert_array_equal(data["indices"], s.basis.indices)
    np.testing.assert_allclose(data["real_potential"], s.real_potential, rtol=0, atol=1e-15)
    y = ham.apply(s.trial[:, :4])
    np.testing.assert_allclose(y, data["hpsi"], rtol=0, atol=1e-13)
