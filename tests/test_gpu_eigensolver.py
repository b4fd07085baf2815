"""GPU Rayleigh-Ritz (``qe_python_amd.eigensolver``) against host eigh.

A random Hermitian operator K on a 900-dimensional space and a growing,
non-orthogonal basis V (blocks appended as Davidson does, with a restart to
the Ritz vectors) give pencils (V^H K V, V^H V); at every step the device
solve (spectral division, with its host fallback) must give the lowest
nroots eigenvalues of the pencil within 1e-11 (relative to the spectrum
width) and S-orthonormal coefficients whose Ritz vectors have residuals
below 1e-9.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev_block(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV)


@pytest.mark.parametrize("native", [True, False], ids=["native", "python"])
@pytest.mark.parametrize("nroots,blocks", [(24, [40, 24, 24, 24]), (64, [80, 64, 64, 64])])
def test_spectral_rayleigh_ritz_matches_host(nroots, blocks, native, monkeypatch):
    """Both forms of the standardisation and spectral solve: the native calls
    (qepw_rr_standardize / qepw_rr_spectral) and the Python steps."""
    from qe_python_amd import device as dv
    from qe_python_amd import eigensolver
    from qe_python_amd.eigensolver import DeviceRayleighRitz
    monkeypatch.setattr(eigensolver, "NATIVE_RR", native)
    rng = np.random.default_rng(nroots)
    npw = 900
    q, _ = np.linalg.qr(rng.standard_normal((npw, npw)) + 1j * rng.standard_normal((npw, npw)))
    spectrum = np.sort(np.concatenate([np.linspace(-1.0, 2.0, npw - 5), [0.3, 0.3000001, 0.31,
                                                                       0.32, 0.5]]))
    K = (q * spectrum) @ q.conj().T
    maxsub = sum(blocks)
    V = torch.zeros((maxsub, npw), dtype=torch.complex128, device=DEV)
    HV = torch.zeros_like(V)
    ws = dv.Workspace(torch.device(DEV))
    rr = DeviceRayleighRitz(maxsub, nroots, torch.device(DEV), ws)
    active = 0
    first = 0
    used_spectral = 0
    for step, nb in enumerate(blocks + blocks[1:]):
        if active + nb > maxsub:
            # restart to the current Ritz vectors, as the Davidson restart does
            X = host_v @ coef_h
            V[:nroots] = dev_block(X)
            HV[:nroots] = dev_block(K @ X)
            rr.restart(values)
            active = nroots
        blk = rng.standard_normal((npw, nb)) + 1j * rng.standard_normal((npw, nb))
        blk /= np.linalg.norm(blk, axis=0)
        first, active = active, active + nb
        V[first:active] = dev_block(blk)
        HV[first:active] = dev_block(K @ blk)
        rr.update(V, HV, first, active, npw)
        values, coef = rr.solve(active)
        host_v = V[:active].cpu().numpy().T
        h = host_v.conj().T @ K @ host_v
        s = host_v.conj().T @ host_v
        import scipy.linalg as sla
        ref = sla.eigh(0.5 * (h + h.conj().T), 0.5 * (s + s.conj().T), eigvals_only=True)[:nroots]
        np.testing.assert_allclose(values, ref, atol=1e-11 * 3.0, rtol=0)
        coef_h = coef.cpu().numpy().T
        gram = coef_h.conj().T @ s @ coef_h
        np.testing.assert_allclose(gram, np.eye(nroots), atol=1e-10)
        X = host_v @ coef_h
        res = np.linalg.norm(K @ X - X * values[None, :], axis=0)
        # Ritz residuals are those of the exact Rayleigh-Ritz of the subspace:
        # compare with the host solve's residuals
        w, vv = sla.eigh(0.5 * (h + h.conj().T), 0.5 * (s + s.conj().T))
        Xr = host_v @ vv[:, :nroots]
        res_ref = np.linalg.norm(K @ Xr - Xr * w[None, :nroots], axis=0)
        np.testing.assert_allclose(res, res_ref, atol=1e-9, rtol=1e-6)
        used_spectral = rr.stats["spectral"]
    assert used_spectral > 0, rr.stats


@pytest.mark.parametrize("n,ld", [(300, 320), (37, 40)])
def test_sign_iteration_entry_point(n, ld):
    """qepw_sign_iteration against the same scaled Newton-Schulz steps in numpy:
    the iterate (exactly Hermitian) and its square, left in y.  300 and 37 are
    not multiples of the 16 x 16 tiles nor of the K steps."""
    import ctypes
    from qe_python_amd import _native as nat
    from qe_python_amd import device as dv
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = a + a.conj().T
    x0 = a / np.abs(a).sum(axis=0).max()
    alphas = [1.6, 1.4, 1.2, 1.0, 1.0]
    ref = x0.copy()
    for al in alphas:
        ref = 0.5 * al * ref @ (3.0 * np.eye(n) - al * al * ref @ ref)
        ref = 0.5 * (ref + ref.conj().T)
    bufs = [torch.zeros((ld, ld), dtype=torch.complex128, device=DEV) for _ in range(3)]
    bufs[0][:n, :n] = torch.from_numpy(np.ascontiguousarray(x0.T)).to(DEV)
    lib = nat.load()
    ws = dv.Workspace(torch.device(DEV)).get(int(lib.qepw_sign_iteration_workspace(n)))
    which = ctypes.c_int(-1)
    coeffs = (ctypes.c_double * len(alphas))(*alphas)
    st = dv.stream_handle(torch.device(DEV))
    nat.check(lib.qepw_sign_iteration(bufs[0].data_ptr(), bufs[1].data_ptr(), bufs[2].data_ptr(),
                                      ld, n, coeffs, len(alphas), ctypes.byref(which),
                                      ws.data_ptr(), ws.numel(), st), "sign")
    assert which.value == 1  # five steps: the result ping-ponged into x2
    got = bufs[which.value][:n, :n].cpu().numpy().T
    np.testing.assert_allclose(got, ref, atol=1e-12)
    np.testing.assert_array_equal(got, got.conj().T)
    sq = bufs[2][:n, :n].cpu().numpy().T
    np.testing.assert_allclose(sq, got @ got, atol=1e-12)
    with pytest.raises(ValueError):
        nat.check(lib.qepw_sign_iteration(bufs[0].data_ptr(), bufs[0].data_ptr(),
                                          bufs[2].data_ptr(), ld, n, coeffs, 1,
                                          ctypes.byref(which), ws.data_ptr(), ws.numel(), st),
                  "sign")


@pytest.mark.parametrize("native", [True, False], ids=["native", "python"])
def test_dependent_basis_raises_like_zhegvd(native, monkeypatch):
    """A basis with a repeated column has a singular overlap: the
    standardisation reports it as the reference's zhegvd failure."""
    from qe_python_amd import device as dv
    from qe_python_amd import eigensolver
    from qe_python_amd.eigensolver import DeviceRayleighRitz
    monkeypatch.setattr(eigensolver, "NATIVE_RR", native)
    rng = np.random.default_rng(5)
    npw, nb = 200, 12
    blk = rng.standard_normal((npw, nb)) + 1j * rng.standard_normal((npw, nb))
    blk[:, 7] = blk[:, 3]
    V = dev_block(blk)
    HV = dev_block(blk * np.linspace(0.0, 1.0, nb)[None, :])
    rr = DeviceRayleighRitz(nb, 4, torch.device(DEV), dv.Workspace(torch.device(DEV)))
    rr.update(V, HV, 0, nb, npw)
    with pytest.raises(np.linalg.LinAlgError, match="zhegvd"):
        rr.solve(nb)



@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-12), (torch.complex64, 2e-5)])
def test_native_cholesky_qr_matches_python_steps(dtype, tol, monkeypatch):
    """qepw_cholesky_qr (one native call) against the Python steps of the
    same Cholesky QR: same orthonormal block, and both decline an
    ill-conditioned block (the caller's cue for the canonical form)."""
    from qe_python_amd import device as dv
    from qe_python_amd import diagonalization as dg
    from qe_python_amd.schedule import blocking
    rng = np.random.default_rng(11)
    npw, nb = 3000, 40
    host = rng.standard_normal((npw, nb)) + 1j * rng.standard_normal((npw, nb))
    block = dev_block(host).to(dtype)
    ws = dv.Workspace(torch.device(DEV))
    native = blocking(dg._cholesky_orthonormalize(block, npw, ws))
    monkeypatch.setattr(dg, "NATIVE_ORTHO", False)
    python = blocking(dg._cholesky_orthonormalize(block, npw, ws))
    q = native.cpu().numpy().T.astype(np.complex128)
    np.testing.assert_allclose(q, python.cpu().numpy().T, atol=tol * np.abs(q).max(), rtol=0)
    np.testing.assert_allclose(q.conj().T @ q, np.eye(nb), atol=50 * tol)
    # same span as the input
    proj = q @ (q.conj().T @ host)
    assert np.linalg.norm(proj - host) <= 50 * tol * np.linalg.norm(host)
    bad = host.copy()
    bad[:, 5] = bad[:, 4] + 1e-9 * bad[:, 6]
    bad_block = dev_block(bad).to(dtype)
    assert blocking(dg._cholesky_orthonormalize(bad_block, npw, ws)) is None
    monkeypatch.setattr(dg, "NATIVE_ORTHO", True)
    assert blocking(dg._cholesky_orthonormalize(bad_block, npw, ws)) is None
