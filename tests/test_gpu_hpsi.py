"""GPU parity of h_psi (``PlaneWaveHamiltonian.apply``) against the oracle.

Bars: FP64 h_psi relative error <= 1e-12 against the oracle (Frobenius norm
over the block); FP32 (complex64 FFT / V psi / projections) relative error
<= 2e-6.  Also: the reference's V_loc known answer (2e-14), Hermiticity,
identity-potential round trip, term-type equivalence, device
prepare_potential, layouts and the reference's error behaviour.
"""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle.dense import potential_matrix
from oracle.hpsi import OracleHamiltonian, apply_local, linear_slots
from qe_python_amd.synthetic import make_system, phase_matrix

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def device_ham(s, terms=None, **kw):
    from qe_python_amd.diagonalization import PackedProjectorTerm, PlaneWaveHamiltonian
    if terms is None:
        beta, coup, diag = s.packed_term()
        terms = (PackedProjectorTerm(beta, coup, diag),)
    return PlaneWaveHamiltonian(s.basis, s.potential_g, terms, real_potential=s.real_potential,
                                **kw)


def oracle_ham(s):
    beta, coup, diag = s.packed_term()
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    return OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                             (term,), potential_average=float(np.real(s.potential_g.flat[0])))


@pytest.fixture(scope="module")
def small():
    return make_system("small")


@pytest.fixture(scope="module")
def medium():
    return make_system("medium")


def test_known_answer_on_device():
    """tests/unit/core/test_basis.py:76-89 of the reference, through the
    device prepare_potential (inverse_real) and V_loc pipeline."""
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    shape = (6, 6, 6)
    grid = np.indices(shape).sum(axis=0)
    vr = 0.3 + 0.2 * np.cos(2.0 * np.pi * grid / shape[0])
    vg = np.fft.fftn(vr) / np.prod(shape)
    idx = np.asarray([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0]], dtype=np.int32)
    c = np.asarray([1.0 + 0.2j, -0.3j, 0.4, -0.1 + 0.5j])
    basis = SimpleNamespace(indices=idx, kinetic=np.zeros(4))
    ham = PlaneWaveHamiltonian(basis, vg)
    np.testing.assert_allclose(ham.apply(c), potential_matrix(vg, idx) @ c, atol=2.0e-14)


def test_device_prepare_potential_matches_numpy(small):
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    ham = PlaneWaveHamiltonian(small.basis, small.potential_g)
    expect = np.real(np.fft.ifftn(small.potential_g)) * small.potential_g.size
    np.testing.assert_allclose(ham.real_potential, expect, atol=1e-14, rtol=0)


def test_small_hpsi_fp64_parity(small):
    ham = device_ham(small)
    got = ham.apply(small.trial)
    ref = oracle_ham(small).apply(small.trial)
    assert rel(got, ref) <= 1e-12


def test_small_hpsi_against_golden(small):
    from pathlib import Path
    data = np.load(Path(__file__).parent / "golden" / "small_hpsi.npz")
    got = device_ham(small).apply(data["trial4"])
    assert rel(got, data["hpsi"]) <= 1e-12


def test_medium_hpsi_fp64_parity(medium):
    ham = device_ham(medium)
    x = medium.trial[:, :24]
    got = ham.apply(x)
    ref = oracle_ham(medium).apply(x)
    assert rel(got, ref) <= 1e-12


def test_medium_hpsi_fp32_parity(medium):
    ham = device_ham(medium, precision="fp32")
    x = medium.trial[:, :16]
    got = ham.apply(x)
    ref = oracle_ham(medium).apply(x)
    assert rel(got, ref) <= 2e-6


def test_cells_54_64_hpsi_parity():
    for name in ("cell54", "cell64"):
        s = make_system(name, kpoint=np.array([0.21, 0.07, 0.33]))
        x = s.trial[:, :6]
        assert rel(device_ham(s).apply(x), oracle_ham(s).apply(x)) <= 1e-12, name


def test_non_square_plane_generic_kernel():
    """An orthorhombic cell (n1 != n2) runs the generic plane kernel."""
    from qe_python_amd.basis import fft_shape, make_bases
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    rng = np.random.default_rng(17)
    reciprocal = 2.0 * np.pi * np.diag([1 / 9.0, 1 / 11.0, 1 / 7.5])
    basis = make_bases(reciprocal, np.array([0.1, 0.2, 0.05]), 20.0)[0]
    shape = fft_shape([basis], reciprocal, 80.0)
    assert shape[0] != shape[1]
    vr = rng.uniform(-1.0, 0.5, shape)
    x = rng.standard_normal((len(basis), 5)) + 1j * rng.standard_normal((len(basis), 5))
    ham = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr)
    ref = OracleHamiltonian(basis.indices, basis.kinetic, shape, vr).apply(x)
    assert rel(ham.apply(x), ref) <= 1e-12
    ham32 = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr,
                                 precision="fp32")
    assert rel(ham32.apply(x), ref) <= 2e-6


def test_team_kernel_sizes_against_oracle():
    """Square planes of several team-FFT lengths (factorisations 4x6 ... 9x10),
    in FP64 and FP32 (each its own kernel instantiation)."""
    from qe_python_amd.basis import make_bases
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    rng = np.random.default_rng(23)
    for n, n3, ecut in ((24, 24, 6.0), (25, 25, 6.0), (30, 30, 10.0), (48, 45, 20.0),
                        (75, 40, 40.0), (90, 80, 60.0), (100, 40, 48.0), (108, 40, 48.0),
                        (120, 40, 48.0)):
        reciprocal = 2.0 * np.pi / 12.0 * np.eye(3)
        basis = make_bases(reciprocal, np.array([0.3, 0.1, 0.2]), ecut)[0]
        shape = (n, n, n3)
        vr = rng.uniform(-1.0, 0.5, shape)
        x = rng.standard_normal((len(basis), 3)) + 1j * rng.standard_normal((len(basis), 3))
        ham = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr)
        ref = OracleHamiltonian(basis.indices, basis.kinetic, shape, vr).apply(x)
        assert rel(ham.apply(x), ref) <= 1e-12, n
        ham32 = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr,
                                     precision="fp32")
        assert rel(ham32.apply(x), ref) <= 2e-6, n


def test_square_planes_without_team_kernel_use_the_generic_kernel():
    """12 x 12 and 15 x 15 planes (no team FFT) run the generic plane kernel."""
    from qe_python_amd.basis import make_bases
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    rng = np.random.default_rng(29)
    for n, n3, ecut in ((12, 15, 1.0), (15, 12, 1.0)):
        reciprocal = 2.0 * np.pi / 12.0 * np.eye(3)
        basis = make_bases(reciprocal, np.array([0.2, 0.1, 0.0]), ecut)[0]
        shape = (n, n, n3)
        vr = rng.uniform(-1.0, 0.5, shape)
        x = rng.standard_normal((len(basis), 2)) + 1j * rng.standard_normal((len(basis), 2))
        ham = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr)
        ref = OracleHamiltonian(basis.indices, basis.kinetic, shape, vr).apply(x)
        assert rel(ham.apply(x), ref) <= 1e-12, n


@pytest.mark.parametrize("shape", [(128, 128, 20), (144, 150, 16)])
def test_large_planes_split_stage(shape):
    """Planes no workgroup can hold in FP64 (128 x 128, and a non-square
    144 x 150) take the split plane stage: h_psi in FP64 and FP32 against the
    oracle, and the device V(G) -> V(r) transform against numpy."""
    from qe_python_amd.basis import make_bases
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    rng = np.random.default_rng(31)
    reciprocal = 2.0 * np.pi / 12.0 * np.eye(3)
    basis = make_bases(reciprocal, np.array([0.1, 0.3, 0.2]), 8.0)[0]
    vr = rng.uniform(-1.0, 0.5, shape)
    x = rng.standard_normal((len(basis), 3)) + 1j * rng.standard_normal((len(basis), 3))
    ref = OracleHamiltonian(basis.indices, basis.kinetic, shape, vr).apply(x)
    ham = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr)
    assert rel(ham.apply(x), ref) <= 1e-12
    ham32 = PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (), real_potential=vr,
                                 precision="fp32")
    assert rel(ham32.apply(x), ref) <= 2e-6
    vg = np.fft.fftn(vr) / vr.size
    ham_g = PlaneWaveHamiltonian(basis, vg)
    np.testing.assert_allclose(ham_g.real_potential, vr, atol=1e-13, rtol=0)


def test_unsupported_grid_fails_at_construction():
    """FFT lines longer than 256 points are rejected when the Hamiltonian is
    built, with a message naming the limit (every grid up to 256 per axis
    runs: team, generic or split plane stage)."""
    from qe_python_amd.basis import make_bases
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    reciprocal = 2.0 * np.pi / 12.0 * np.eye(3)
    basis = make_bases(reciprocal, np.zeros(3), 6.0)[0]
    shape = (270, 270, 24)
    with pytest.raises(RuntimeError, match="<= 256"):
        PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (),
                             real_potential=np.zeros(shape))


def test_hermiticity_on_device(small):
    ham = device_ham(small)
    x, y = small.trial[:, :4], small.trial[:, 4:9]
    np.testing.assert_allclose(x.conj().T @ ham.apply(y), ham.apply(x).conj().T @ y,
                               rtol=1e-12, atol=1e-12)


def test_identity_potential_round_trip(small):
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    ham = PlaneWaveHamiltonian(small.basis, small.potential_g, (),
                               real_potential=np.ones(small.shape))
    x = small.trial[:, :5]
    np.testing.assert_allclose(ham.apply(x), (small.basis.kinetic + 1.0)[:, None] * x,
                               atol=1e-13, rtol=0)


def test_term_types_agree_and_diagonal(small):
    from qe_python_amd.diagonalization import FactorizedProjectorTerm
    pos = small.projectors.positions
    coupling = small.projectors.couplings[0]
    beta_s = small.projectors.beta_species
    fact = FactorizedProjectorTerm(beta_s, phase_matrix(small.basis.vectors, pos), coupling)
    mats = tuple((beta_s * np.exp(-1j * (small.basis.vectors @ p))[:, None], coupling)
                 for p in pos)
    x = small.trial[:, :6]
    hf, hm = device_ham(small, (fact,)), device_ham(small, mats)
    of = OracleHamiltonian(small.basis.indices, small.basis.kinetic, small.shape,
                           small.real_potential, (fact,))
    assert rel(hf.apply(x), of.apply(x)) <= 1e-12
    assert rel(hm.apply(x), of.apply(x)) <= 1e-12
    np.testing.assert_allclose(hf.diagonal - hf.potential_average,
                               of.diagonal - of.potential_average, atol=1e-12)
    np.testing.assert_allclose(device_ham(small).diagonal, oracle_ham(small).diagonal,
                               atol=1e-12)


def test_native_layout_and_apply_into_and_torch_input(small):
    from qe_python_amd.diagonalization import PackedProjectorTerm, PlaneWaveHamiltonian
    beta, coup, diag = small.packed_term()
    native = np.ascontiguousarray(np.transpose(small.real_potential, (2, 0, 1)))
    ham = PlaneWaveHamiltonian(small.basis, small.potential_g,
                               (PackedProjectorTerm(beta, coup, diag),),
                               real_potential=native, native_potential_layout=True)
    x = np.asfortranarray(small.trial[:, :3])
    ref = oracle_ham(small).apply(x)
    out = np.empty_like(x, order="F")
    assert ham.apply_into(x, out) is out
    assert rel(out, ref) <= 1e-12
    assert rel(ham.apply(x[:, 1]), ref[:, 1]) <= 1e-12
    xt = torch.from_numpy(x).to("cuda")
    assert rel(ham.apply(xt).cpu().numpy(), ref) <= 1e-12


def test_shape_errors_are_value_errors(small):
    ham = device_ham(small)
    with pytest.raises(ValueError):
        ham.apply(np.zeros((small.npw + 1, 2), dtype=complex))
    with pytest.raises(ValueError):
        ham.apply_into(small.trial[:, :2], np.zeros((small.npw, 3), dtype=complex))


def test_grouped_apply_equals_single_apply():
    """``PlaneWaveHamiltonian.device_apply_group`` (one grouped launch per
    FFT stage over several problems of different grids, npw and band counts)
    gives bit for bit what ``device_apply`` gives per problem, in both
    precisions; problems whose grid cannot join the group (generic and split
    plane stages) run on their own inside the same call."""
    from qe_python_amd.basis import make_bases
    from qe_python_amd.diagonalization import PlaneWaveHamiltonian
    rng = np.random.default_rng(41)
    hams = [device_ham(make_system("small"))]
    for name, k in (("cell54", [0.21, 0.07, 0.33]), ("cell64", [0.0, 0.5, 0.25]),
                    ("cell54", [0.4, 0.1, 0.0])):
        hams.append(device_ham(make_system(name, kpoint=np.array(k))))
    reciprocal = 2.0 * np.pi / 12.0 * np.eye(3)
    for shape, ecut in (((30, 30, 30), 10.0), ((15, 12, 12), 1.0), ((128, 128, 20), 8.0)):
        basis = make_bases(reciprocal, np.array([0.1, 0.3, 0.2]), ecut)[0]
        hams.append(PlaneWaveHamiltonian(basis, np.zeros(shape, complex), (),
                                         real_potential=rng.uniform(-1.0, 0.5, shape)))
    bands = [5, 64, 17, 1, 33, 7, 3, 2]
    for dtype in (torch.complex128, torch.complex64):
        requests, singles = [], []
        for ham, nb in zip(hams, bands):
            x = rng.standard_normal((ham.npw, nb)) + 1j * rng.standard_normal((ham.npw, nb))
            src = ham.to_device(x, dtype)
            one = torch.empty_like(src)
            # NaN bytes in the scratch: a kernel reading what no stage wrote
            # would show up in the result
            ham.workspace.get(ham.plan.workspace_bytes(dtype, nb)).fill_(255)
            ham.device_apply(src, one)
            singles.append(one)
            requests.append((ham, src, torch.full_like(src, float("nan"))))
            ham.workspace.get(ham.plan.workspace_bytes(dtype, nb)).fill_(255)
        before = [h.applications for h in hams]
        PlaneWaveHamiltonian.device_apply_group(requests)
        for (ham, src, got), one, b0, nb in zip(requests, singles, before, bands):
            assert bool(torch.isfinite(torch.view_as_real(one)).all()), (ham.plan.shape, dtype)
            assert torch.equal(got, one), (ham.plan.shape, dtype)
            assert ham.applications == b0 + nb
    PlaneWaveHamiltonian.device_apply_group([])


def test_timed_apply_matches_and_times_the_phase(medium):
    """qepw_apply_local_timed: the same result as qepw_apply_local, and its
    events span the FFT phase's kernels (positive, below the span of a
    Python-level event pair around the same call)."""
    ham = device_ham(medium)
    x = ham.to_device(medium.trial[:, :64])
    res = ham._resources(x.dtype)
    ref = torch.empty_like(x)
    ham.plan.apply_local(x, ref, res.potential, res.kinetic, ham.workspace)
    got = torch.full_like(x, float("nan"))
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    stop.record()
    outer0 = torch.cuda.Event(enable_timing=True)
    outer1 = torch.cuda.Event(enable_timing=True)
    outer0.record()
    ham.plan.apply_local_timed(x, got, res.potential, res.kinetic, ham.workspace, start, stop)
    outer1.record()
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    inner = start.elapsed_time(stop)
    assert 0.0 < inner <= outer0.elapsed_time(outer1) + 1e-3
