"""GPU parity of the GEMM and the bandwidth-bound Davidson kernels, called
through the C ABI, against NumPy and the oracle's restatements of the
reference formulas (``oracle.davidson``).

Tolerances: complex128 results within 1e-13 relative (GEMM: 1e-12 of the
product norm); complex64 GEMM within 2e-6 + 1e-7 sqrt(k) of the product
norm (FP32 accumulation over k terms).
"""

import numpy as np
import pytest
import torch

from oracle import davidson as odav

pytestmark = pytest.mark.gpu

DEV = "cuda"


def dev_block(a, dtype=torch.complex128):
    """(rows, cols) host -> (cols, rows) device column-major tensor."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(DEV).to(dtype)


def host_block(t):
    return t.cpu().numpy().T


def rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("op", ["N", "C", "NB", "CB"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (64, 64, 16), (65, 130, 37),
                                   (256, 128, 23847), (17, 300, 4000), (2945, 8, 32),
                                   (384, 384, 384), (500, 130, 257), (3, 2, 0),
                                   # one output column: the matrix-vector kernels
                                   (23847, 1, 256), (256, 1, 23847), (300, 1, 5000),
                                   (1, 1, 3000), (70000, 1, 3), (7, 1, 0)])
def test_gemm_matches_numpy(prec, op, m, n, k):
    from qe_python_amd import _native as nat
    from qe_python_amd import device as dv
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    dtype = torch.complex128 if prec == "c128" else torch.complex64
    a = rand(rng, m, k) if op[0] == "N" else rand(rng, k, m)
    b = rand(rng, n, k) if op.endswith("B") else rand(rng, k, n)
    c = rand(rng, m, n)
    alpha, beta = 0.7 - 0.2j, -0.3 + 0.5j
    ta, tb, tc = dev_block(a, dtype), dev_block(b, dtype), dev_block(c, dtype)
    ws = dv.Workspace(torch.device(DEV))
    code = (nat.OP_N if op[0] == "N" else nat.OP_C) | (nat.OP_BC if op.endswith("B") else 0)
    dv.gemm(code, m, n, k, alpha, ta, tb, beta, tc, ws)
    opa = a if op[0] == "N" else a.conj().T
    b = b.conj().T if op.endswith("B") else b
    if prec == "c64":
        a32, b32, c32 = (x.astype(np.complex64).astype(np.complex128) for x in (opa, b, c))
        ref = alpha * (a32 @ b32) + beta * c32
        # FP32 accumulation over k terms: error grows like sqrt(k) * 2^-24
        tol = 2e-6 + 1e-7 * np.sqrt(k)
    else:
        ref = alpha * (opa @ b) + beta * c
        tol = 1e-12
    got = host_block(tc)
    scale = max(np.linalg.norm(ref), 1e-30)
    assert np.linalg.norm(got - ref) / scale < tol


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("op", ["N", "C", "NB", "CB"])
def test_every_gemm_tile_variant_matches_numpy(prec, op):
    """Each tile variant of the GEMM (forced through a tuning rule, with and
    without split K) on shapes that are not multiples of its tiles, against
    NumPy; the rules are cleared and the committed table restored after."""
    import ctypes

    from qe_python_amd import _native as nat
    from qe_python_amd import autotune
    from qe_python_amd import device as dv
    lib = nat.load()
    dtype = torch.complex128 if prec == "c128" else torch.complex64
    code = (nat.OP_N if op[0] == "N" else nat.OP_C) | (nat.OP_BC if op.endswith("B") else 0)
    ws = dv.Workspace(torch.device(DEV))
    try:
        for variant in range(lib.qepw_gemm_variant_count()):
            for m, n, k, splits in [(70, 130, 45, 1), (4100, 97, 33, 1), (200, 61, 3000, 7),
                                    (33, 129, 513, 3)]:
                rng = np.random.default_rng(variant * 100 + m + n + k)
                a = rand(rng, m, k) if op[0] == "N" else rand(rng, k, m)
                b = rand(rng, n, k) if op.endswith("B") else rand(rng, k, n)
                c = rand(rng, m, n)
                alpha, beta = 0.6 + 0.3j, -0.2 + 0.1j
                ta, tb, tc = dev_block(a, dtype), dev_block(b, dtype), dev_block(c, dtype)
                lib.qepw_gemm_clear_rules()
                rng_ = (ctypes.c_int64 * 6)(m, m, n, n, k, k)
                nat.check(lib.qepw_gemm_set_rule(dv.precision_code(dtype), code, rng_,
                                                 variant, splits), "set_rule")
                dv.forget_gemm_plans()
                plan = (ctypes.c_int64 * 4)()
                nat.check(lib.qepw_gemm_plan(dv.precision_code(dtype), code, m, n, k, plan),
                          "plan")
                assert plan[0] == variant
                dv.gemm(code, m, n, k, alpha, ta, tb, beta, tc, ws)
                opa = a if op[0] == "N" else a.conj().T
                bb = b.conj().T if op.endswith("B") else b
                ref = alpha * (opa @ bb) + beta * c
                tol = 1e-12 if prec == "c128" else 2e-6 + 1e-7 * np.sqrt(k)
                err = np.linalg.norm(host_block(tc) - ref) / np.linalg.norm(ref)
                assert err < tol, (variant, m, n, k, splits, err)
    finally:
        lib.qepw_gemm_clear_rules()
        autotune.load_table()


def test_gemm_split_k_is_deterministic():
    from qe_python_amd import _native as nat
    from qe_python_amd import device as dv
    rng = np.random.default_rng(5)
    a, b = rand(rng, 23847, 128), rand(rng, 23847, 160)
    ta, tb = dev_block(a), dev_block(b)
    ws = dv.Workspace(torch.device(DEV))
    outs = []
    for _ in range(2):
        tc = torch.empty((160, 128), dtype=torch.complex128, device=DEV)
        dv.gemm(nat.OP_C, 128, 160, 23847, 1.0, ta, tb, 0.0, tc, ws)
        outs.append(tc.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_gemv_beta_zero_ignores_nan_output_and_is_deterministic():
    from qe_python_amd import _native as nat
    from qe_python_amd import device as dv
    rng = np.random.default_rng(10)
    a, b = rand(rng, 23847, 256), rand(rng, 23847, 1)
    ws = dv.Workspace(torch.device(DEV))
    outs = []
    for _ in range(2):
        tc = torch.full((1, 256), float("nan"), dtype=torch.complex128, device=DEV)
        dv.gemm(nat.OP_C, 256, 1, 23847, 1.0, dev_block(a), dev_block(b), 0.0, tc, ws)
        outs.append(host_block(tc))
    np.testing.assert_allclose(outs[0], a.conj().T @ b, rtol=1e-12, atol=1e-12)
    assert np.array_equal(outs[0], outs[1])


def test_gemm_beta_zero_ignores_nan_output():
    from qe_python_amd import _native as nat
    from qe_python_amd import device as dv
    rng = np.random.default_rng(9)
    a, b = rand(rng, 40, 20), rand(rng, 20, 10)
    tc = torch.full((10, 40), float("nan"), dtype=torch.complex128, device=DEV)
    dv.gemm(nat.OP_N, 40, 10, 20, 1.0, dev_block(a), dev_block(b), 0.0, tc,
            dv.Workspace(torch.device(DEV)))
    np.testing.assert_allclose(host_block(tc), a @ b, rtol=1e-13, atol=1e-13)


@pytest.fixture(scope="module")
def blocks():
    rng = np.random.default_rng(11)
    npw, nb = 3001, 9
    return rng, npw, nb, rand(rng, npw, nb), rand(rng, npw, nb)


def test_column_norms_and_inner_products(blocks):
    from qe_python_amd import device as dv
    rng, npw, nb, x, y = blocks
    sq = dv.column_squared_norms(dev_block(x), npw).cpu().numpy()
    np.testing.assert_allclose(sq, odav.column_squared_norms(x), rtol=1e-13)
    ip = dv.column_inner_products(dev_block(x), dev_block(y), npw).cpu().numpy()
    np.testing.assert_allclose(ip[0::2] + 1j * ip[1::2], np.einsum("gb,gb->b", x.conj(), y),
                               rtol=1e-13)
    sq32 = dv.column_squared_norms(dev_block(x, torch.complex64), npw).cpu().numpy()
    x32 = x.astype(np.complex64).astype(np.complex128)
    np.testing.assert_allclose(sq32, odav.column_squared_norms(x32), rtol=1e-13)


def test_subtract_energies_precondition_normalize_restart(blocks):
    from qe_python_amd import device as dv
    rng, npw, nb, x, r = blocks
    e = rng.uniform(-1, 1, nb)
    diag = rng.uniform(0, 20, npw)
    td = dev_block(r)
    dv.subtract_band_energies(td, dev_block(x), torch.from_numpy(e).to(DEV), npw)
    np.testing.assert_allclose(host_block(td), r - x * e[None, :], rtol=1e-14, atol=1e-14)

    cols = np.array([0, 3, 4, 8], dtype=np.int32)
    out = torch.empty((len(cols), npw), dtype=torch.complex128, device=DEV)
    dv.qe_precondition(dev_block(r), torch.from_numpy(cols).to(DEV),
                       torch.from_numpy(e[cols]).to(DEV), torch.from_numpy(diag).to(DEV),
                       out, npw)
    np.testing.assert_allclose(host_block(out), odav.qe_precondition(r[:, cols], e[cols], diag),
                               rtol=1e-14)

    sq = odav.column_squared_norms(r)
    sel = np.array([1, 2, 7], dtype=np.int32)
    dst = torch.empty((len(sel), npw), dtype=torch.complex128, device=DEV)
    dv.normalize_selected_columns(dev_block(r), torch.from_numpy(sq).to(DEV),
                                  torch.from_numpy(sel).to(DEV), dst, npw)
    np.testing.assert_allclose(host_block(dst), r[:, sel] / np.sqrt(sq[sel])[None, :],
                               rtol=1e-14)

    hv = torch.empty((nb, npw), dtype=torch.complex128, device=DEV)
    dv.restart_applied(dev_block(x), dev_block(r), torch.from_numpy(e).to(DEV), hv, npw)
    np.testing.assert_allclose(host_block(hv), x * e[None, :] + r, rtol=1e-14, atol=1e-14)


def test_permute_rows_round_trip_and_cast(blocks):
    from qe_python_amd import device as dv
    rng, npw, nb, x, _ = blocks
    perm = rng.permutation(npw).astype(np.int32)
    tp = torch.from_numpy(perm).to(DEV)
    g = torch.empty((nb, npw), dtype=torch.complex128, device=DEV)
    dv.permute_rows(0, tp, dev_block(x), g, npw)
    assert np.array_equal(host_block(g), x[perm])
    back = torch.empty_like(g)
    dv.permute_rows(1, tp, g, back, npw)
    assert np.array_equal(host_block(back), x)
    g32 = torch.empty((nb, npw), dtype=torch.complex64, device=DEV)
    dv.permute_rows(0, tp, dev_block(x), g32, npw)
    assert np.array_equal(host_block(g32), x[perm].astype(np.complex64))


def test_projector_diagonal_accumulation():
    from qe_python_amd import device as dv
    rng = np.random.default_rng(2)
    npw, nch = 777, 5
    beta, coupled = rand(rng, npw, nch), rand(rng, npw, nch)
    acc = torch.ones(npw, dtype=torch.float64, device=DEV)
    dv.accumulate_projector_diagonal(dev_block(coupled), dev_block(beta), npw, 3.0, acc)
    expect = 1.0 + 3.0 * np.real(np.einsum("gi,gi->g", coupled, beta.conj()))
    np.testing.assert_allclose(acc.cpu().numpy(), expect, rtol=1e-13)
