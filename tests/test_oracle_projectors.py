"""Pins of the oracle's projector assembly (CPU): the reference's known
answer for the cubic interpolation, orthonormal harmonics, the reference's
low-l closed forms, and the synthetic systems' analytic Gaussian projectors."""

import numpy as np

from oracle import projectors as op
from qe_python_amd.synthetic import kb_beta_species, PROJECTOR_SIGMA


def test_cubic_interpolation_known_answer():
    # tests/unit/core/test_upf.py:41-48 of the reference
    dq = 0.2
    grid = np.arange(20) * dq
    table = grid ** 3 - 2.0 * grid ** 2 + 0.5 * grid - 1.0
    q = np.asarray([0.27, 0.93, 2.41])
    np.testing.assert_allclose(op.qe_cubic_interpolate(table, q, dq),
                               q ** 3 - 2.0 * q ** 2 + 0.5 * q - 1.0, atol=3e-15)


def test_harmonics_orthonormal_by_quadrature():
    # Gauss-Legendre in cos(theta) times a uniform phi grid integrates
    # products of harmonics of degree <= 6 exactly
    x, w = np.polynomial.legendre.leggauss(12)
    phi = np.arange(24) * (2.0 * np.pi / 24)
    ct, ph = np.meshgrid(x, phi, indexing="ij")
    st = np.sqrt(1.0 - ct ** 2)
    dirs = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=-1).reshape(-1, 3)
    weights = np.repeat(w, len(phi)) * (2.0 * np.pi / len(phi))
    ys = np.column_stack([op.qe_real_harmonics(l, 1.7 * dirs) for l in range(4)])
    gram = ys.T @ (weights[:, None] * ys)
    np.testing.assert_allclose(gram, np.eye(16), atol=1e-13)


def test_low_l_closed_forms_of_the_reference():
    rng = np.random.default_rng(3)
    v = rng.standard_normal((50, 3))
    x, y, z = (v / np.linalg.norm(v, axis=1)[:, None]).T
    pi = np.pi
    np.testing.assert_allclose(op.qe_real_harmonics(0, v)[:, 0], 1 / np.sqrt(4 * pi))
    np.testing.assert_allclose(op.qe_real_harmonics(1, v),
                               np.sqrt(3 / (4 * pi)) * np.column_stack([z, -x, -y]), atol=1e-15)
    l2 = np.column_stack([np.sqrt(5 / (16 * pi)) * (3 * z * z - 1),
                          -np.sqrt(15 / (4 * pi)) * x * z, -np.sqrt(15 / (4 * pi)) * y * z,
                          np.sqrt(15 / (16 * pi)) * (x * x - y * y),
                          np.sqrt(15 / (4 * pi)) * x * y])
    np.testing.assert_allclose(op.qe_real_harmonics(2, v), l2, atol=1e-15)
    # q = 0 takes the +z direction
    np.testing.assert_allclose(op.qe_real_harmonics(1, np.zeros((1, 3))),
                               [[np.sqrt(3 / (4 * pi)), 0.0, 0.0]], atol=1e-16)


def gaussian_table(volume, dq, nq, sigma=PROJECTOR_SIGMA):
    q = np.arange(nq) * dq
    pref = 4.0 * np.pi / np.sqrt(volume) * np.sqrt(0.5 * np.pi)
    g = np.exp(-0.5 * q * q * sigma * sigma)
    return np.column_stack([pref * sigma ** 3 * g, pref * sigma ** 5 * q * g])


def test_projector_basis_reproduces_analytic_gaussians():
    rng = np.random.default_rng(5)
    gk = rng.uniform(-3.0, 3.0, (400, 3))
    volume = 1234.5
    dq = 0.01
    table = gaussian_table(volume, dq, int(np.linalg.norm(gk, axis=1).max() / dq) + 5)
    beta = op.projector_basis(gk, table, dq, [(0, 0, 0), (1, 1, 0), (1, 1, 1), (1, 1, 2)])
    ref = kb_beta_species(gk, volume)
    # the four-point interpolant at dq = 0.01 is good to ~5e-9 relative; a
    # wrong sign, index, harmonic or prefactor would be an O(1) error
    np.testing.assert_allclose(beta, ref, rtol=0, atol=1e-8 * np.abs(ref).max())


def test_expanded_coupling_of_the_synthetic_species():
    # two radial projectors (s, p) with D in Ry: the per-atom (4, 4) couplings
    # of the synthetic systems, in Ha, are diag(d_s, d_p, d_p, d_p) / 2
    from qe_python_amd.projectors import expanded_coupling
    d_ry = np.asarray([[1.5, 0.25], [0.25, -0.75]])
    chans = [(0, 0, 0), (1, 1, 0), (1, 1, 1), (1, 1, 2)]
    np.testing.assert_array_equal(expanded_coupling(chans, d_ry),
                                  np.diag([0.75, -0.375, -0.375, -0.375]))
    # two p projectors couple channel by channel through dij
    chans = [(0, 1, 0), (0, 1, 1), (1, 1, 0), (1, 1, 1)]
    got = expanded_coupling(chans, d_ry)
    assert got[0, 2] == got[2, 0] == got[1, 3] == 0.125
    assert got[0, 1] == got[0, 3] == 0.0
