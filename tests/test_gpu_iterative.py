"""GPU parity of CG, ParO and RMM-DIIS on the device h_psi (reference
``qepy_pw/diagonalization.py:993-1756``) against the oracle restatements
(``oracle/iterative.py``) on the same inputs, the reference's own ParO tests
through a host operator, converged eigenvalues against the dense ones of the
small workload, and the new per-column kernels against their formulas."""

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import iterative as oit
from oracle.hpsi import OracleHamiltonian
from qe_python_amd.synthetic import make_system

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"


@pytest.fixture(scope="module")
def small():
    from qe_python_amd.diagonalization import PackedProjectorTerm, PlaneWaveHamiltonian
    s = make_system("small")
    beta, coup, diag = s.packed_term()
    ham = PlaneWaveHamiltonian(s.basis, s.potential_g, (PackedProjectorTerm(beta, coup, diag),),
                               real_potential=s.real_potential)
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ref = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential, (term,),
                            potential_average=float(np.real(s.potential_g.flat[0])))
    exact = np.load(GOLDEN / "small_hpsi.npz")["dense_eigenvalues"]
    return s, ham, ref, exact


def test_per_column_kernels():
    import torch

    from qe_python_amd import device as dv
    rng = np.random.default_rng(2)
    npw, nb = 1000, 5
    x = rng.normal(size=(nb, npw)) + 1j * rng.normal(size=(nb, npw))
    y = rng.normal(size=(nb, npw)) + 1j * rng.normal(size=(nb, npw))
    a, b = rng.normal(size=nb), rng.normal(size=nb)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    dv.column_axpby(xd, torch.from_numpy(a).cuda(), yd, torch.from_numpy(b).cuda(), npw)
    np.testing.assert_allclose(yd.cpu().numpy(), a[:, None] * x + b[:, None] * y, atol=1e-14)
    # b = 0 does not read y (NaN there stays out of the result)
    yd.fill_(float("nan"))
    dv.column_axpby(xd, None, yd, torch.zeros(nb, dtype=torch.float64).cuda(), npw)
    np.testing.assert_array_equal(yd.cpu().numpy(), x)
    t = rng.uniform(0.0, 30.0, npw)
    td = torch.from_numpy(t).cuda()
    got = dv.column_diagonal_expectations(xd, td, npw).cpu().numpy()
    np.testing.assert_allclose(got, (np.abs(x) ** 2) @ t, rtol=1e-13)
    e = rng.uniform(0.5, 3.0, 3)
    cols = torch.tensor([4, 0, 2], dtype=torch.int32).cuda()
    out = torch.empty((3, npw), dtype=torch.complex128).cuda()
    dv.rmm_precondition(xd, cols, td, torch.from_numpy(e).cuda(), out, npw)
    ref = oit.kresse_furthmueller(x[[4, 0, 2]].T, t, e).T
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-13, atol=0)
    hist = rng.normal(size=(3, nb, npw)) + 1j * rng.normal(size=(3, nb, npw))
    coef = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    out = torch.empty((2, npw), dtype=torch.complex128).cuda()
    dv.history_combine(torch.from_numpy(hist).cuda(), 2, torch.tensor([3, 1], dtype=torch.int32).cuda(),
                       torch.from_numpy(coef).cuda(), out, npw)
    ref = np.stack([coef[0, 0] * hist[0, 3] + coef[0, 1] * hist[1, 3],
                    coef[1, 0] * hist[0, 1] + coef[1, 1] * hist[1, 1]])
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-14)
    # ldr = 0: every column of l against one vector (l^H r)
    from qe_python_amd import _native as nat
    ip = torch.empty(2 * nb, dtype=torch.float64).cuda()
    nat.check(nat.load().qepw_column_inner_products(
        nat.C128, xd.data_ptr(), npw, yd[1].data_ptr(), 0, npw, nb, ip.data_ptr(),
        dv.stream_handle(xd.device)), "qepw_column_inner_products")
    np.testing.assert_allclose(ip.cpu().numpy().view(np.complex128), x.conj() @ x[1],
                               rtol=1e-13)


def test_paro_reference_tests_through_a_host_operator():
    from qe_python_amd.diagonalization import parallel_orbital
    rng = np.random.default_rng(7321)
    m = rng.normal(size=(18, 18))
    h = 0.5 * (m + m.T) + np.diag(np.linspace(1.0, 18.0, 18))
    res = parallel_orbital(lambda x: h @ x, np.diag(h), 4,
                           initial_vectors=rng.normal(size=(18, 4)), tolerance=1e-10,
                           max_iterations=20)
    assert res.converged and res.number_unconverged == 0
    np.testing.assert_allclose(res.eigenvalues, np.linalg.eigvalsh(h)[:4], atol=1e-9)
    assert np.max(res.residual_norms) < 3.0e-6
    rng = np.random.default_rng(91)
    d = np.arange(1.0, 13.0)
    res = parallel_orbital(lambda x: d[:, None] * x, d, 3,
                           initial_vectors=rng.normal(size=(12, 3)), tolerance=1e-8,
                           max_iterations=2)
    assert res.hamiltonian_applications > 2 * 3


def check_against_oracle(got, ref, exact=None, atol=1e-9):
    print(f"happ device {got.hamiltonian_applications} oracle {ref.hamiltonian_applications}; "
          f"max |de| {np.max(np.abs(got.eigenvalues - ref.eigenvalues)):.2e}", flush=True)
    np.testing.assert_allclose(got.eigenvalues, ref.eigenvalues, rtol=0, atol=atol)
    assert got.converged == ref.converged
    assert got.hamiltonian_applications == ref.hamiltonian_applications
    np.testing.assert_allclose(got.residual_norms, ref.residual_norms, rtol=0, atol=1e-7)
    if exact is not None:
        np.testing.assert_allclose(got.eigenvalues, exact[:len(got.eigenvalues)], rtol=0,
                                   atol=1e-8)


def test_cg_small_matches_oracle_and_dense(small):
    from qe_python_amd.diagonalization import conjugate_gradient
    s, ham, ref_ham, exact = small
    kw = dict(initial_vectors=s.trial[:, :7], tolerance=1e-11, max_iterations=100)
    got = conjugate_gradient(ham.apply, ham.diagonal, 7, operator_into=ham.apply_into, **kw)
    ref = oit.conjugate_gradient(ref_ham.apply, ref_ham.diagonal, 7, **kw)
    check_against_oracle(got, ref, exact)
    v = got.eigenvectors
    np.testing.assert_allclose(v.conj().T @ v, np.eye(7), atol=1e-9)


def test_paro_small_matches_oracle_and_dense(small):
    from qe_python_amd.diagonalization import parallel_orbital
    s, ham, ref_ham, exact = small
    kw = dict(initial_vectors=s.trial[:, :7], tolerance=1e-11, max_iterations=60)
    got = parallel_orbital(ham.apply, ham.diagonal, 7, operator_into=ham.apply_into, **kw)
    ref = oit.parallel_orbital(ref_ham.apply, ref_ham.diagonal, 7, **kw)
    check_against_oracle(got, ref, exact)


def test_rmm_small_matches_oracle_and_converges_when_repeated(small):
    from qe_python_amd.diagonalization import rmm_diis
    s, ham, ref_ham, exact = small
    kw = dict(tolerance=1e-11)
    got = rmm_diis(ham.apply, ham.diagonal, 7, initial_vectors=s.trial[:, :7],
                   operator_into=ham.apply_into, **kw)
    ref = oit.rmm_diis(ref_ham.apply, ref_ham.diagonal, 7, initial_vectors=s.trial[:, :7], **kw)
    check_against_oracle(got, ref, atol=1e-9)
    # repeated as the SCF repeats it: device and oracle stay together and the
    # lower six bands become the lowest eigenpairs of H.  RMM-DIIS is a local
    # method: here the top band drifts towards an excited state, slowly
    # enough that ten sweeps amplify rounding differences to ~1e-8 in it.
    v, vr = got.eigenvectors, ref.eigenvectors
    for _ in range(10):
        got = rmm_diis(ham.apply, ham.diagonal, 7, initial_vectors=v, initial_is_orthonormal=True,
                       operator_into=ham.apply_into, **kw)
        ref = oit.rmm_diis(ref_ham.apply, ref_ham.diagonal, 7, initial_vectors=vr,
                           initial_is_orthonormal=True, **kw)
        v, vr = got.eigenvectors, ref.eigenvectors
    np.testing.assert_allclose(got.eigenvalues[:6], ref.eigenvalues[:6], rtol=0, atol=1e-9)
    np.testing.assert_allclose(got.eigenvalues, ref.eigenvalues, rtol=0, atol=1e-7)
    assert np.max(got.residual_norms[:6]) < 1e-5
    np.testing.assert_allclose(got.eigenvalues[:6], exact[:6], rtol=0, atol=1e-8)


def test_ritz_input_and_device_output(small):
    import torch

    from qe_python_amd.diagonalization import conjugate_gradient, davidson, rmm_diis
    s, ham, _, exact = small
    d = davidson(ham.apply, ham.diagonal, 7, initial_vectors=s.trial, tolerance=1e-9,
                 residual_factor=None, operator_into=ham.apply_into)
    hv = ham.apply(d.eigenvectors)
    for solver in (conjugate_gradient, rmm_diis):
        res = solver(ham.apply, ham.diagonal, 7, initial_vectors=d.eigenvectors,
                     initial_is_orthonormal=True, initial_is_ritz=True, initial_applied=hv,
                     initial_eigenvalues=d.eigenvalues, operator_into=ham.apply_into,
                     tolerance=1e-11, device_output=True)
        assert isinstance(res.eigenvectors, torch.Tensor)
        np.testing.assert_allclose(res.eigenvalues, exact[:7], rtol=0, atol=1e-8)


def test_argument_errors(small):
    from qe_python_amd.diagonalization import conjugate_gradient, parallel_orbital, rmm_diis
    s, ham, _, _ = small
    x0 = s.trial[:, :3]
    with pytest.raises(ValueError):
        parallel_orbital(ham.apply, ham.diagonal, 3, initial_vectors=x0, tolerance=0.0)
    with pytest.raises(ValueError):
        conjugate_gradient(ham.apply, ham.diagonal, 3, initial_vectors=x0, max_iterations=0)
    with pytest.raises(ValueError):
        rmm_diis(ham.apply, ham.diagonal, 3, initial_vectors=x0, history_dimension=1)
    with pytest.raises(ValueError, match="at least"):
        rmm_diis(ham.apply, ham.diagonal, 4, initial_vectors=x0)
    with pytest.raises(ValueError, match="dependent"):
        parallel_orbital(ham.apply, ham.diagonal, 3,
                         initial_vectors=np.repeat(s.trial[:, :1], 3, axis=1))
    with pytest.raises(ValueError, match="Ritz"):
        conjugate_gradient(ham.apply, ham.diagonal, 3, initial_vectors=x0, initial_is_ritz=True)


def test_rmm_medium_matches_oracle():
    """One RMM-DIIS call on the medium workload's 128 bands (the SCF's
    per-iteration use): same H-application count and eigenvalues as the
    oracle's call on the host."""
    from qe_python_amd.diagonalization import PackedProjectorTerm, PlaneWaveHamiltonian, rmm_diis
    s = make_system("medium")
    beta, coup, diag = s.packed_term()
    ham = PlaneWaveHamiltonian(s.basis, s.potential_g, (PackedProjectorTerm(beta, coup, diag),),
                               real_potential=s.real_potential)
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ref_ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                                (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    x0 = s.trial[:, :128]
    got = rmm_diis(ham.apply, ham.diagonal, 128, initial_vectors=x0, operator_into=ham.apply_into,
                   tolerance=1e-8)
    ref = oit.rmm_diis(ref_ham.apply, ref_ham.diagonal, 128, initial_vectors=x0, tolerance=1e-8)
    check_against_oracle(got, ref, atol=1e-10)


@pytest.mark.parametrize("solver", ["cg", "paro", "rmm"])
def test_solvers_at_a_general_kpoint(solver):
    """Away from Gamma the Hamiltonian has no real form: the three solvers on
    the small cell at a seeded k-point against the oracle on the same inputs
    (same H-application counts) and, converged, against dense eigenvalues."""
    from oracle.dense import dense_hamiltonian
    from qe_python_amd.diagonalization import (PackedProjectorTerm, PlaneWaveHamiltonian,
                                               conjugate_gradient, parallel_orbital, rmm_diis)
    from qe_python_amd.synthetic import batch_kpoints
    k = batch_kpoints("small", 2, seed=3)[1]
    s = make_system("small", kpoint=k)
    beta, coup, diag = s.packed_term()
    ham = PlaneWaveHamiltonian(s.basis, s.potential_g, (PackedProjectorTerm(beta, coup, diag),),
                               real_potential=s.real_potential)
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ref_ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                                (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    x0 = s.trial[:, :6]
    dev_fn, ora_fn, kw = {
        "cg": (conjugate_gradient, oit.conjugate_gradient,
               dict(tolerance=1e-11, max_iterations=100)),
        "paro": (parallel_orbital, oit.parallel_orbital, dict(tolerance=1e-11, max_iterations=60)),
        "rmm": (rmm_diis, oit.rmm_diis, dict(tolerance=1e-11)),
    }[solver]
    got = dev_fn(ham.apply, ham.diagonal, 6, initial_vectors=x0, operator_into=ham.apply_into,
                 **kw)
    ref = ora_fn(ref_ham.apply, ref_ham.diagonal, 6, initial_vectors=x0, **kw)
    if solver == "rmm":
        check_against_oracle(got, ref, atol=1e-9)
        return
    exact = np.linalg.eigvalsh(dense_hamiltonian(s.basis.indices, s.basis.kinetic,
                                                 s.real_potential, (term,)))
    check_against_oracle(got, ref, exact)
