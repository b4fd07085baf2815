"""GPU parity of the per-k projector assembly (QE ``init_us_2``; reference
``qepy_pw/upf.py:483-541`` with ``upf.py:722-772, 880-912``, packed as in
``pw/scf.py:657-686``): ``qepw_projector_basis`` against the oracle for
every l <= 3 channel, against the synthetic systems' analytic Gaussian
projectors, the atom-packed term against the host-packed one, a Hamiltonian
built from the device term against one built from host arrays, and the
argument errors."""

import numpy as np
import pytest

from oracle import projectors as op
from qe_python_amd.synthetic import kb_beta_species, make_system, PROJECTOR_SIGMA

pytestmark = pytest.mark.gpu

ALL_CHANNELS = [(p, l, m) for p, l in enumerate(range(4)) for m in range(2 * l + 1)]


def gaussian_table(volume, dq, nq, sigma=PROJECTOR_SIGMA):
    q = np.arange(nq) * dq
    pref = 4.0 * np.pi / np.sqrt(volume) * np.sqrt(0.5 * np.pi)
    g = np.exp(-0.5 * q * q * sigma * sigma)
    return np.column_stack([pref * sigma ** 3 * g, pref * sigma ** 5 * q * g])


def test_projector_basis_matches_oracle_all_l():
    from qe_python_amd.projectors import projector_basis
    rng = np.random.default_rng(17)
    gk = rng.uniform(-4.0, 4.0, (3001, 3))
    gk[0] = 0.0                       # q = 0 takes the +z direction
    gk[1] = [0.0, 0.0, -2.5]          # on the -z axis
    gk[2] = [1.0e-3, 0.0, 0.0]        # inside the first table interval
    dq = 0.01
    nq = int(np.linalg.norm(gk, axis=1).max() / dq) + 5
    table = rng.standard_normal((nq, 4)).cumsum(axis=0) * 1e-2
    ref = op.projector_basis(gk, table, dq, ALL_CHANNELS)
    got = projector_basis(gk, table, dq, ALL_CHANNELS).cpu().numpy()
    assert got.shape == ref.shape == (3001, 16)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    # a channel subset in another order picks the same columns
    sub = [ALL_CHANNELS[i] for i in (12, 0, 5, 3)]
    got = projector_basis(gk, table, dq, sub).cpu().numpy()
    np.testing.assert_allclose(got, ref[:, [12, 0, 5, 3]], rtol=0, atol=1e-13 * np.abs(ref).max())


def test_projector_basis_reproduces_the_synthetic_projectors():
    from qe_python_amd.projectors import projector_basis
    s = make_system("medium")
    gk = s.basis.vectors
    dq = 0.01
    table = gaussian_table(s.volume, dq, int(np.floor(np.linalg.norm(gk, axis=1).max() / dq + 4.5)))
    got = projector_basis(gk, table, dq, [(0, 0, 0), (1, 1, 0), (1, 1, 1), (1, 1, 2)])
    ref = kb_beta_species(gk, s.volume)
    # four-point interpolation at dq = 0.01: ~5e-9 relative
    np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=0, atol=1e-8 * np.abs(ref).max())


def test_packed_term_matches_host_packing():
    import torch

    from qe_python_amd.projectors import packed_projector_term
    s = make_system("small")
    beta_h, coup_h, diag_h = s.packed_term()
    pr = s.projectors
    term = packed_projector_term(torch.from_numpy(np.ascontiguousarray(pr.beta_species)),
                                 pr.couplings, pr.positions, s.basis.vectors)
    assert isinstance(term.beta, torch.Tensor) and term.beta.shape == beta_h.shape
    np.testing.assert_allclose(term.beta.cpu().numpy(), beta_h, rtol=0,
                               atol=1e-14 * np.abs(beta_h).max())
    np.testing.assert_array_equal(term.coupling, coup_h)
    np.testing.assert_allclose(term.diagonal, diag_h, rtol=0, atol=1e-13 * np.abs(diag_h).max())


def test_hamiltonian_from_device_projectors():
    from qe_python_amd.diagonalization import PackedProjectorTerm, PlaneWaveHamiltonian
    from qe_python_amd.projectors import packed_projector_term, projector_basis
    s = make_system("small")
    gk = s.basis.vectors
    dq = 0.01
    table = gaussian_table(s.volume, dq, int(np.floor(np.linalg.norm(gk, axis=1).max() / dq + 4.5)))
    chans = [(0, 0, 0), (1, 1, 0), (1, 1, 1), (1, 1, 2)]
    pr = s.projectors
    beta_d = projector_basis(gk, table, dq, chans)
    term_d = packed_projector_term(beta_d, pr.couplings, pr.positions, gk)
    # the same term packed on the host from the oracle's species block
    beta_s = op.projector_basis(gk, table, dq, chans)
    phases = np.exp(-1j * (gk @ pr.positions.T))
    beta_h = np.concatenate([beta_s * phases[:, a, None] for a in range(pr.natom)], axis=1)
    diag_h = sum(np.real(np.einsum("gi,ij,gj->g", beta_s, c, beta_s.conj()))
                 for c in pr.couplings)
    term_h = PackedProjectorTerm(np.asfortranarray(beta_h), term_d.coupling, diag_h)
    kw = dict(real_potential=s.real_potential)
    ham_d = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_d,), **kw)
    ham_h = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_h,), **kw)
    x = s.trial[:, :12]
    ref = ham_h.apply(x)
    got = ham_d.apply(x)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    np.testing.assert_allclose(ham_d.diagonal, ham_h.diagonal, rtol=0, atol=1e-12)


def test_projector_basis_errors():
    from qe_python_amd.projectors import packed_projector_term, projector_basis
    gk = np.asarray([[0.5, 0.0, 0.0], [0.0, 1.0, 0.0]])
    dq = 0.1
    ok = np.ones((14, 1))
    projector_basis(gk, ok, dq, [(0, 0, 0)])
    with pytest.raises(ValueError, match="does not cover"):
        projector_basis(gk, np.ones((13, 1)), dq, [(0, 0, 0)])
    with pytest.raises(ValueError, match="channels"):
        projector_basis(gk, ok, dq, [(0, 4, 0)])
    with pytest.raises(ValueError, match="channels"):
        projector_basis(gk, ok, dq, [(1, 0, 0)])
    with pytest.raises(ValueError, match="channels"):
        projector_basis(gk, ok, dq, [(0, 1, 3)])
    with pytest.raises(ValueError, match="gk"):
        projector_basis(gk[:, :2], ok, dq, [(0, 0, 0)])
    beta = projector_basis(gk, ok, dq, [(0, 0, 0)])
    with pytest.raises(ValueError, match="coupling"):
        packed_projector_term(beta, [np.eye(2)], np.zeros((1, 3)), gk)


def test_factorized_term_on_the_device():
    import torch

    from qe_python_amd.diagonalization import FactorizedProjectorTerm, PlaneWaveHamiltonian
    from qe_python_amd.projectors import factorized_projector_term, packed_projector_term
    s = make_system("small")
    pr = s.projectors
    gk = s.basis.vectors
    d = pr.couplings[0]                      # one species: one shared coupling
    beta_s = torch.from_numpy(np.ascontiguousarray(pr.beta_species))
    term_d = factorized_projector_term(beta_s, d, pr.positions, gk)
    phases = np.exp(-1j * (gk @ pr.positions.T))
    np.testing.assert_allclose(term_d.phases.cpu().numpy(), phases, rtol=0, atol=1e-14)
    term_h = FactorizedProjectorTerm(np.asfortranarray(pr.beta_species), np.asfortranarray(phases),
                                     d)
    term_p = packed_projector_term(beta_s, [d] * pr.natom, pr.positions, gk)
    kw = dict(real_potential=s.real_potential)
    ham_d = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_d,), **kw)
    ham_h = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_h,), **kw)
    ham_p = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_p,), **kw)
    x = s.trial[:, :12]
    ref = ham_h.apply(x)
    got = ham_d.apply(x)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    # the device factorized and device packed forms build the same blocks
    np.testing.assert_array_equal(got, ham_p.apply(x))
    np.testing.assert_allclose(ham_d.diagonal, ham_h.diagonal, rtol=0, atol=1e-12)
    with pytest.raises(ValueError, match="coupling"):
        factorized_projector_term(beta_s, np.eye(3), pr.positions, gk)


def test_medium_factorized_term_at_a_kpoint():
    """The medium cell's 64 atoms x 4 channels (256 > 64 packed channels: the
    reference SCF's factorized form) at a general k-point: projectors
    assembled on the device from the Gaussian radial tables, factorized, in
    a Hamiltonian applied to 8 bands, against the same Hamiltonian with the
    host-packed analytic projectors (the interpolation's ~5e-9 relative
    error bounds the difference)."""
    import torch

    from qe_python_amd.diagonalization import PackedProjectorTerm, PlaneWaveHamiltonian
    from qe_python_amd.projectors import factorized_projector_term, projector_basis
    s = make_system("medium", kpoint=np.array([0.21, 0.07, 0.33]))
    pr = s.projectors
    gk = s.basis.vectors
    dq = 0.01
    table = gaussian_table(s.volume, dq, int(np.floor(np.linalg.norm(gk, axis=1).max() / dq + 4.5)))
    d = pr.couplings[0]
    beta = projector_basis(gk, table, dq, [(0, 0, 0), (1, 1, 0), (1, 1, 1), (1, 1, 2)])
    term_d = factorized_projector_term(beta, d, pr.positions, gk)
    phases = np.exp(-1j * (gk @ pr.positions.T))
    beta_h = np.concatenate([pr.beta_species * phases[:, a, None] for a in range(pr.natom)],
                            axis=1)
    coup_h = np.kron(np.eye(pr.natom), d)
    diag_h = pr.natom * np.real(np.einsum("gi,ij,gj->g", pr.beta_species, d,
                                          pr.beta_species.conj()))
    term_h = PackedProjectorTerm(np.asfortranarray(beta_h), coup_h, diag_h)
    kw = dict(real_potential=s.real_potential)
    ham_d = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_d,), **kw)
    ham_h = PlaneWaveHamiltonian(s.basis, s.potential_g, (term_h,), **kw)
    x = s.trial[:, :8]
    ref = ham_h.apply(x)
    got = ham_d.apply(x)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"medium factorized at k: relative difference {rel:.2e}", flush=True)
    assert rel <= 1e-8
    # the diagonal is quadratic in beta: twice the interpolation's error
    np.testing.assert_allclose(ham_d.diagonal, ham_h.diagonal, rtol=0,
                               atol=2e-8 * np.abs(ham_h.diagonal).max())
    del torch
