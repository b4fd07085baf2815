"""The entry point that allocates nothing (``qepw_gemm_static``) against
``qepw_gemm`` with a caller workspace, through the C ABI.

Both run the same kernels, split plan and fixed-order split-K reduction, so
the outputs must be bitwise equal.  Two calls in a row on the shared static
scratch, the second smaller (its partial tiles overwrite only part of the
first call's), for both precisions; plus a one-column product (the
matrix-vector path, whose partials also go to the scratch) and a product
without a split.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pair(x):
    return (ctypes.c_double * 2)(float(np.real(x)), float(np.imag(x)))


def _matrix(rng, cols, rows, dtype):
    real = torch.float64 if dtype == torch.complex128 else torch.float32
    re = torch.from_numpy(rng.standard_normal((cols, rows))).to(real)
    im = torch.from_numpy(rng.standard_normal((cols, rows))).to(real)
    return torch.complex(re, im).to("cuda")


def _both(lib, prec, op, m, n, k, a, b, c0, alpha, beta):
    """(qepw_gemm result, qepw_gemm_static result) from the same C start."""
    from qe_python_amd import _native as nat
    stream = torch.cuda.current_stream().cuda_stream
    c_ref, c_static = c0.clone(), c0.clone()
    need = int(lib.qepw_gemm_workspace(prec, op, m, n, k))
    ws = torch.empty(max(need, 256), dtype=torch.uint8, device="cuda")
    rc = lib.qepw_gemm(prec, op, m, n, k, _pair(alpha), a.data_ptr(), a.shape[1], b.data_ptr(),
                       b.shape[1], _pair(beta), c_ref.data_ptr(), c_ref.shape[1], ws.data_ptr(),
                       ws.numel(), stream)
    assert rc == nat.OK, nat.last_error()
    rc = lib.qepw_gemm_static(prec, op, m, n, k, _pair(alpha), a.data_ptr(), a.shape[1],
                              b.data_ptr(), b.shape[1], _pair(beta), c_static.data_ptr(),
                              c_static.shape[1], stream)
    assert rc == nat.OK, nat.last_error()
    torch.cuda.synchronize()
    return need, c_ref, c_static


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_static_scratch_gemm_equals_workspace_gemm(dtype):
    from qe_python_amd import _native as nat
    lib = nat.load()
    prec = nat.C128 if dtype == torch.complex128 else nat.C64
    limit = int(lib.qepw_gemm_static_scratch_bytes())
    assert limit == 32 << 20
    rng = np.random.default_rng(2026)
    # (op, m, n, k): update_projected's A^H B over npw at the medium size (the
    # largest split-K scratch of a solve), then a smaller A^H B, an A B
    # without a split, and a one-column A^H b
    cases = [(nat.OP_C, 128, 512, 23847), (nat.OP_C, 40, 72, 10118), (nat.OP_N, 3000, 64, 96),
             (nat.OP_C, 256, 1, 23847)]
    needs = []
    for op, m, n, k in cases:
        if op & nat.OP_C:
            a = _matrix(rng, m, k, dtype)          # k x m column-major
        else:
            a = _matrix(rng, k, m, dtype)          # m x k column-major
        b = _matrix(rng, n, k, dtype)              # k x n column-major
        c0 = _matrix(rng, n, m, dtype)             # m x n column-major
        need, ref, got = _both(lib, prec, op, m, n, k, a, b, c0, 0.5 - 0.25j, 1.0 + 0.5j)
        needs.append(need)
        assert torch.equal(got, ref), (op, m, n, k)
        # and the product itself is right (FP64 host reference)
        ah = a.cpu().numpy().astype(np.complex128).T      # column-major -> (rows, cols)
        bh = b.cpu().numpy().astype(np.complex128).T
        opa = ah.conj().T if op & nat.OP_C else ah
        want = (0.5 - 0.25j) * (opa @ bh) + (1.0 + 0.5j) * c0.cpu().numpy().astype(
            np.complex128).T
        tol = 1e-12 if dtype == torch.complex128 else 2e-5
        err = np.linalg.norm(got.cpu().numpy().T - want) / np.linalg.norm(want)
        assert err < tol, (op, m, n, k, err)
    # the first call used a split-K scratch no smaller than the second's
    assert needs[0] >= needs[1] > 0 and needs[2] == 0 and needs[3] > 0
    assert max(needs) <= limit


def test_static_scratch_gemm_rejects_a_product_beyond_the_scratch():
    """A^H B with 1024 x 1040 outputs over k = 4096 splits K in two: 34 MB
    of partial tiles in FP64, beyond the 32 MiB static array -- rejected
    before any launch, with the sizes in the message."""
    from qe_python_amd import _native as nat
    lib = nat.load()
    m, n, k = 1024, 1040, 4096
    need = int(lib.qepw_gemm_workspace(nat.C128, nat.OP_C, m, n, k))
    assert need > int(lib.qepw_gemm_static_scratch_bytes())
    a = torch.zeros((m, k), dtype=torch.complex128, device="cuda")
    b = torch.zeros((n, k), dtype=torch.complex128, device="cuda")
    c = torch.full((n, m), 3.0 + 0j, dtype=torch.complex128, device="cuda")
    rc = lib.qepw_gemm_static(nat.C128, nat.OP_C, m, n, k, _pair(1.0), a.data_ptr(), k,
                              b.data_ptr(), k, _pair(0.0), c.data_ptr(), m,
                              torch.cuda.current_stream().cuda_stream)
    assert rc == nat.ERR_WORKSPACE
    assert "static" in nat.last_error() and str(need) in nat.last_error()
    torch.cuda.synchronize()
    assert torch.all(c == 3.0)
