"""Device buffers and thin wrappers over the C ABI.

PyTorch-ROCm is used only as the allocator and stream provider: device
matrices are torch tensors whose raw pointers go to ``libqepw_hip.so``; no
torch arithmetic runs on the hot path.

Matrix convention: a column-major (nrows x ncols) block with leading
dimension ld is a torch tensor of shape (ncols, ld) with unit stride along
dim 1, so ``t[a:b]`` is the column block a:b (as in the reference's
F-ordered ``basis_storage[:, a:b]``, ``qepy_pw/diagonalization.py:632-637``).
"""

from __future__ import annotations

import ctypes
import threading
from contextlib import contextmanager

import numpy as np
import torch

from . import _native as nat

C128 = torch.complex128
C64 = torch.complex64


def precision_code(dtype: torch.dtype) -> int:
    if dtype == C128:
        return nat.C128
    if dtype == C64:
        return nat.C64
    raise ValueError(f"unsupported complex dtype {dtype}")


def real_dtype(dtype: torch.dtype) -> torch.dtype:
    return torch.float64 if dtype == C128 else torch.float32


_TLS = threading.local()


def stream_handle(device: torch.device) -> int:
    """HIP stream handle of the current torch stream (the one fixed by an
    enclosing ``fixed_stream`` scope of this thread, if any)."""
    fixed = getattr(_TLS, "stream", None)
    if fixed is not None:
        return fixed
    return torch.cuda.current_stream(device).cuda_stream


@contextmanager
def fixed_stream(device: torch.device):
    """Within one solve the current stream does not change: look it up once
    instead of per launch (thread-local, so concurrent solves on their own
    streams are unaffected)."""
    prev = getattr(_TLS, "stream", None)
    _TLS.stream = torch.cuda.current_stream(device).cuda_stream
    try:
        yield
    finally:
        _TLS.stream = prev


def upload(array, device: torch.device) -> torch.Tensor:
    """Small host array -> device tensor without stalling the host: staged in
    pinned memory (torch's caching host allocator keeps the staging block
    until the copy has run) and copied asynchronously on the current stream;
    a pageable copy would wait for all queued device work first."""
    src = torch.from_numpy(np.ascontiguousarray(array))
    pinned = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
    pinned.copy_(src)
    return pinned.to(device, non_blocking=True)


def require_device(device=None) -> torch.device:
    """Resolve the HIP device; fail loudly when there is none."""
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the qe_python_amd device path "
                           "has no CPU fallback")
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    nat.load()
    return dev


_NOT_MATRIX = "device matrix must be a (ncols, ld) tensor with unit inner stride"


def _matrix(t: torch.Tensor):
    """(ld, ncols, nrows) of a (ncols, ld) device matrix view, one pass over
    its metadata (the GEMM wrapper runs ~1000 times per Davidson solve)."""
    shape, stride = t.shape, t.stride()
    if len(shape) != 2:
        raise ValueError(_NOT_MATRIX)
    s0, s1 = shape
    st0, st1 = stride
    if st1 != 1 and s1 > 1 and s0 > 0:
        # the kernels read each column as s1 contiguous elements, one column
        # or many
        raise ValueError(_NOT_MATRIX)
    if t.is_conj() or t.is_neg():
        # a lazy conjugate / negative view: its memory holds the plain values
        raise ValueError("device matrix is a lazy conj/neg view; resolve it first")
    if s0 > 1:
        return (st0 if st0 > 1 else 1), s0, s1
    return max(1, s1, st0), s0, s1


def ld_of(t: torch.Tensor) -> int:
    return _matrix(t)[0]


class Workspace:
    """Grow-only device scratch (the device analogue of ``FFTScratchPool``,
    ``qepy_pw/basis.py:407-454``)."""

    def __init__(self, device: torch.device):
        self.device = device
        self.buffer = torch.empty(0, dtype=torch.uint8, device=device)

    def get(self, nbytes: int) -> torch.Tensor:
        nbytes = max(int(nbytes), 256)
        if self.buffer.numel() < nbytes:
            self.buffer = torch.empty(int(nbytes * 1.25) + 4096, dtype=torch.uint8,
                                      device=self.device)
        return self.buffer


# workspace bytes of a GEMM shape (a pure function of the shape)
_GEMM_WORKSPACE: dict = {}


def forget_gemm_plans() -> None:
    """Drop the cached workspace sizes: a change of the library's tile rules
    (qe_python_amd.autotune) changes the split plan of a shape."""
    _GEMM_WORKSPACE.clear()


# set to a list to record (class, dtype, m, n, k, start, stop) for every GEMM,
# with HIP events on the launch stream around it; the class is the label of
# the enclosing ``gemm_class`` scope of this thread ("other" outside one)
GEMM_TRACE = None


@contextmanager
def gemm_class(label: str):
    """Label the GEMMs launched by this thread within the scope (for
    ``GEMM_TRACE``)."""
    prev = getattr(_TLS, "gemm_class", None)
    _TLS.gemm_class = label
    try:
        yield
    finally:
        _TLS.gemm_class = prev


def gemm(op_a: int, m: int, n: int, k: int, alpha, a: torch.Tensor, b: torch.Tensor,
         beta, c: torch.Tensor, ws: Workspace) -> None:
    """c = alpha op(a) b + beta c on the device (column-major views).

    Every operand view must hold the matrix the sizes name (a view of
    ``ncols`` columns of ``nrows`` elements is a ``(>= ncols, >= nrows)``
    tensor), so that no size slip reaches the device as an access past an
    allocation."""
    lib = nat.load()
    dt = c.dtype
    if a.dtype != dt or b.dtype != dt:
        raise ValueError("gemm operands must share one complex dtype")
    prec = precision_code(dt)
    lda, a0, a1 = _matrix(a)
    ldb, b0, b1 = _matrix(b)
    ldc, c0, c1 = _matrix(c)
    if m > 0 and n > 0 and k > 0:
        ac, ar = (m, k) if op_a & nat.OP_C else (k, m)
        bc, br = (k, n) if op_a & nat.OP_BC else (n, k)
        if a0 < ac or a1 < ar or b0 < bc or b1 < br or c0 < n or c1 < m:
            for name, t, cols, rows in (("a", a, ac, ar), ("b", b, bc, br), ("c", c, n, m)):
                if t.shape[0] < cols or t.shape[1] < rows:
                    raise ValueError(f"gemm operand {name} of shape {tuple(t.shape)} is "
                                     f"smaller than the {rows} x {cols} matrix the sizes name")
    key = (prec, op_a, m, n, k)
    need = _GEMM_WORKSPACE.get(key)
    if need is None:
        need = int(lib.qepw_gemm_workspace(prec, op_a, m, n, k))
        if len(_GEMM_WORKSPACE) > 8192:
            _GEMM_WORKSPACE.clear()
        _GEMM_WORKSPACE[key] = need
    buf = ws.get(need)
    trace = GEMM_TRACE
    if trace is not None:
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(torch.cuda.current_stream(c.device))
    rc = lib.qepw_gemm(prec, op_a, m, n, k, nat.complex_pair(alpha), a.data_ptr(), lda,
                       b.data_ptr(), ldb, nat.complex_pair(beta), c.data_ptr(), ldc,
                       buf.data_ptr(), buf.numel(), stream_handle(c.device))
    if rc:
        nat.check(rc, "qepw_gemm")
    if trace is not None:
        stop.record(torch.cuda.current_stream(c.device))
        trace.append((getattr(_TLS, "gemm_class", None) or "other", dt, m, n, k, start, stop))


def column_squared_norms(v: torch.Tensor, nrows: int, out: torch.Tensor | None = None) -> torch.Tensor:
    nb = v.shape[0]
    if out is None:
        out = torch.empty(nb, dtype=torch.float64, device=v.device)
    rc = nat.load().qepw_column_squared_norms(precision_code(v.dtype), v.data_ptr(), ld_of(v),
                                              nrows, nb, out.data_ptr(),
                                              stream_handle(v.device))
    nat.check(rc, "qepw_column_squared_norms")
    return out


def column_inner_products(l: torch.Tensor, r: torch.Tensor, nrows: int) -> torch.Tensor:
    nb = l.shape[0]
    out = torch.empty(2 * nb, dtype=torch.float64, device=l.device)
    rc = nat.load().qepw_column_inner_products(precision_code(l.dtype), l.data_ptr(), ld_of(l),
                                               r.data_ptr(), ld_of(r), nrows, nb,
                                               out.data_ptr(), stream_handle(l.device))
    nat.check(rc, "qepw_column_inner_products")
    return out


def subtract_band_energies(r: torch.Tensor, x: torch.Tensor, e: torch.Tensor, nrows: int) -> None:
    rc = nat.load().qepw_subtract_band_energies(precision_code(r.dtype), r.data_ptr(), ld_of(r),
                                                x.data_ptr(), ld_of(x), e.data_ptr(), nrows,
                                                r.shape[0], stream_handle(r.device))
    nat.check(rc, "qepw_subtract_band_energies")


def qe_precondition(r: torch.Tensor, cols: torch.Tensor | None, e: torch.Tensor,
                    diag: torch.Tensor, out: torch.Tensor, nrows: int) -> None:
    rc = nat.load().qepw_qe_precondition(precision_code(r.dtype), r.data_ptr(), ld_of(r),
                                         None if cols is None else cols.data_ptr(),
                                         e.data_ptr(), diag.data_ptr(), out.data_ptr(),
                                         ld_of(out), nrows, out.shape[0],
                                         stream_handle(r.device))
    nat.check(rc, "qepw_qe_precondition")


def normalize_selected_columns(src: torch.Tensor, sq: torch.Tensor, sel: torch.Tensor,
                               dst: torch.Tensor, nrows: int) -> None:
    rc = nat.load().qepw_normalize_selected_columns(precision_code(src.dtype), src.data_ptr(),
                                                    ld_of(src), sq.data_ptr(), sel.data_ptr(),
                                                    dst.data_ptr(), ld_of(dst), nrows,
                                                    dst.shape[0], stream_handle(src.device))
    nat.check(rc, "qepw_normalize_selected_columns")


def restart_applied(x: torch.Tensor, r: torch.Tensor, e: torch.Tensor, hv: torch.Tensor,
                    nrows: int) -> None:
    rc = nat.load().qepw_restart_applied(precision_code(x.dtype), x.data_ptr(), ld_of(x),
                                         r.data_ptr(), ld_of(r), e.data_ptr(), hv.data_ptr(),
                                         ld_of(hv), nrows, x.shape[0], stream_handle(x.device))
    nat.check(rc, "qepw_restart_applied")


def permute_rows(direction: int, perm: torch.Tensor | None, src: torch.Tensor,
                 dst: torch.Tensor, nrows: int) -> None:
    """direction 0: dst[d] = src[perm[d]]; 1: dst[perm[d]] = src[d]."""
    rc = nat.load().qepw_permute_rows(precision_code(src.dtype), precision_code(dst.dtype),
                                      direction, None if perm is None else perm.data_ptr(),
                                      src.data_ptr(), ld_of(src), dst.data_ptr(), ld_of(dst),
                                      nrows, src.shape[0], stream_handle(src.device))
    nat.check(rc, "qepw_permute_rows")


def copy_columns(src: torch.Tensor, dst: torch.Tensor, nrows: int, ncols: int | None = None,
                 src_cols: torch.Tensor | None = None, dst_cols: torch.Tensor | None = None,
                 scale: torch.Tensor | None = None) -> None:
    """dst[dst_cols[j]] = scale[j] * src[src_cols[j]] for the first ``ncols``
    columns (column-major (ncols, ld) tensors; maps int32 and scale float64 on
    the device, each optional), converting between complex128 and complex64."""
    if ncols is None:
        ncols = int(src_cols.shape[0]) if src_cols is not None else int(src.shape[0])
    rc = nat.load().qepw_copy_columns(
        precision_code(src.dtype), precision_code(dst.dtype), src.data_ptr(), ld_of(src),
        None if src_cols is None else src_cols.data_ptr(), dst.data_ptr(), ld_of(dst),
        None if dst_cols is None else dst_cols.data_ptr(),
        None if scale is None else scale.data_ptr(), nrows, ncols, stream_handle(dst.device))
    nat.check(rc, "qepw_copy_columns")


def set_block(dst: torch.Tensor, nrows: int, ncols: int, diag: float = 0.0,
              d: torch.Tensor | None = None) -> None:
    """dst[:ncols, :nrows] (column-major) = zero with ``diag`` (or the device
    float64 ``d``) on its diagonal."""
    rc = nat.load().qepw_set_block(precision_code(dst.dtype), dst.data_ptr(), ld_of(dst), nrows,
                                   ncols, float(diag), None if d is None else d.data_ptr(),
                                   stream_handle(dst.device))
    nat.check(rc, "qepw_set_block")


def index_tensor(values, device: torch.device) -> torch.Tensor:
    """Host integer indices -> device int32 (staged through pinned memory)."""
    return upload(np.asarray(values, dtype=np.int32), device)


def accumulate_projector_diagonal(coupled: torch.Tensor, beta: torch.Tensor, nrows: int,
                                  multiplicity: float, diagonal: torch.Tensor) -> None:
    rc = nat.load().qepw_accumulate_projector_diagonal(
        coupled.data_ptr(), ld_of(coupled), beta.data_ptr(), ld_of(beta), nrows,
        coupled.shape[0], float(multiplicity), diagonal.data_ptr(),
        stream_handle(coupled.device))
    nat.check(rc, "qepw_accumulate_projector_diagonal")


def column_axpby(x: torch.Tensor, a: torch.Tensor | None, y: torch.Tensor,
                 b: torch.Tensor | None, nrows: int) -> None:
    """y[j] = a[j] x[j] + b[j] y[j] over the rows of (ncols, nrows) blocks (a, b:
    device float64 per row, None for 1)."""
    ncols = int(y.shape[0])
    if x.shape[0] < ncols or (a is not None and a.numel() < ncols) or (
            b is not None and b.numel() < ncols):
        raise ValueError("column_axpby operands are shorter than the updated block")
    rc = nat.load().qepw_column_axpby(precision_code(y.dtype), x.data_ptr(), ld_of(x),
                                      None if a is None else a.data_ptr(), y.data_ptr(), ld_of(y),
                                      None if b is None else b.data_ptr(), nrows, ncols,
                                      stream_handle(y.device))
    nat.check(rc, "qepw_column_axpby")


def column_diagonal_expectations(v: torch.Tensor, d: torch.Tensor, nrows: int) -> torch.Tensor:
    nb = int(v.shape[0])
    out = torch.empty(nb, dtype=torch.float64, device=v.device)
    rc = nat.load().qepw_column_diagonal_expectations(precision_code(v.dtype), v.data_ptr(),
                                                      ld_of(v), d.data_ptr(), nrows, nb,
                                                      out.data_ptr(), stream_handle(v.device))
    nat.check(rc, "qepw_column_diagonal_expectations")
    return out


def rmm_precondition(r: torch.Tensor, cols: torch.Tensor | None, kinetic: torch.Tensor,
                     expectations: torch.Tensor, out: torch.Tensor, nrows: int) -> None:
    ncols = int(out.shape[0])
    if expectations.numel() < ncols or (cols is not None and cols.numel() < ncols):
        raise ValueError("rmm_precondition metadata is shorter than the output block")
    rc = nat.load().qepw_rmm_precondition(precision_code(r.dtype), r.data_ptr(), ld_of(r),
                                          None if cols is None else cols.data_ptr(),
                                          kinetic.data_ptr(), expectations.data_ptr(),
                                          out.data_ptr(), ld_of(out), nrows, ncols,
                                          stream_handle(r.device))
    nat.check(rc, "qepw_rmm_precondition")


def history_combine(hist: torch.Tensor, count: int, cols: torch.Tensor | None,
                    coef: torch.Tensor, out: torch.Tensor, nrows: int) -> None:
    """out[j] = sum_i coef[j, i] hist[i, cols[j]] for a (depth, nb, nrows)
    history and (ncols, count) complex128 coefficients."""
    ncols = int(out.shape[0])
    if count > hist.shape[0] or tuple(coef.shape) != (ncols, count) or (
            cols is not None and cols.numel() < ncols):
        raise ValueError("history_combine operands do not match")
    rc = nat.load().qepw_history_combine(precision_code(hist.dtype), hist.data_ptr(),
                                         hist.stride(1), hist.stride(0), count,
                                         None if cols is None else cols.data_ptr(),
                                         coef.data_ptr(), out.data_ptr(), ld_of(out), nrows,
                                         ncols, stream_handle(out.device))
    nat.check(rc, "qepw_history_combine")


class DeviceFFTPlan:
    """Device plan of the sphere-pruned V_loc FFT for one k-point basis
    (device counterpart of the serial ``LocalPotentialWorkspace`` tables,
    ``qepy_pw/basis.py:460-555``)."""

    def __init__(self, miller: np.ndarray, shape, device: torch.device):
        lib = nat.load()
        self.device = device = require_device(device)
        miller = np.ascontiguousarray(miller, dtype=np.int32)
        if miller.ndim != 2 or miller.shape[1] != 3:
            raise ValueError("indices must have shape (nplane, 3)")
        self.shape = tuple(int(s) for s in shape)
        self.npw = int(miller.shape[0])
        handle = ctypes.c_void_p()
        rc = lib.qepw_plan_create(miller.ctypes.data, self.npw, *self.shape,
                                  device.index, ctypes.byref(handle))
        nat.check(rc, "qepw_plan_create")
        self.handle = handle
        # FP64 needs the most LDS: a grid it supports runs in FP32 too
        nat.check(lib.qepw_plan_supported(handle, nat.C128), "qepw_plan_supported")
        info = (ctypes.c_int64 * 8)()
        nat.check(lib.qepw_plan_query(handle, info), "qepw_plan_query")
        self.sticks = int(info[4])
        self.active_x = int(info[5])
        self.stick_group = int(info[6])
        perm = np.empty(self.npw, dtype=np.int32)
        nat.check(lib.qepw_plan_permutation(handle, perm.ctypes.data), "qepw_plan_permutation")
        self.perm = perm
        self.perm_device = torch.from_numpy(perm).to(device)

    def __del__(self):
        handle = getattr(self, "handle", None)
        lib = getattr(nat, "_LIB", None) if nat is not None else None
        if handle is not None and lib is not None:  # (module globals are gone at exit)
            lib.qepw_plan_destroy(handle)
            self.handle = None

    def workspace_bytes(self, dtype: torch.dtype, nb: int) -> int:
        return int(nat.load().qepw_apply_local_workspace(self.handle, precision_code(dtype), nb))

    def apply_local(self, vectors: torch.Tensor, result: torch.Tensor, potential: torch.Tensor,
                    kinetic: torch.Tensor | None, ws: Workspace, max_bands: int = 256) -> None:
        nb = vectors.shape[0]
        if nb == 0:
            return
        prec = precision_code(vectors.dtype)
        if result.dtype != vectors.dtype or potential.dtype != real_dtype(vectors.dtype):
            raise ValueError("local-potential operands disagree in precision")
        need = self.workspace_bytes(vectors.dtype, min(nb, max_bands))
        buf = ws.get(need)
        rc = nat.load().qepw_apply_local(
            self.handle, prec, vectors.data_ptr(), ld_of(vectors), nb, potential.data_ptr(),
            None if kinetic is None else kinetic.data_ptr(), result.data_ptr(), ld_of(result),
            buf.data_ptr(), min(buf.numel(), need), stream_handle(vectors.device))
        nat.check(rc, "qepw_apply_local")

    def apply_local_timed(self, vectors: torch.Tensor, result: torch.Tensor,
                          potential: torch.Tensor, kinetic: torch.Tensor | None, ws: Workspace,
                          start: "torch.cuda.Event", stop: "torch.cuda.Event") -> None:
        """``apply_local`` with ``start`` / ``stop`` (torch timing events,
        already recorded once so that their HIP events exist) recorded by the
        library right around the phase's kernels."""
        nb = vectors.shape[0]
        if nb == 0:
            return
        prec = precision_code(vectors.dtype)
        if result.dtype != vectors.dtype or potential.dtype != real_dtype(vectors.dtype):
            raise ValueError("local-potential operands disagree in precision")
        need = self.workspace_bytes(vectors.dtype, min(nb, 256))
        buf = ws.get(need)
        rc = nat.load().qepw_apply_local_timed(
            self.handle, prec, vectors.data_ptr(), ld_of(vectors), nb, potential.data_ptr(),
            None if kinetic is None else kinetic.data_ptr(), result.data_ptr(), ld_of(result),
            buf.data_ptr(), min(buf.numel(), need), start.cuda_event, stop.cuda_event,
            stream_handle(vectors.device))
        nat.check(rc, "qepw_apply_local_timed")

    @staticmethod
    def apply_local_group(items) -> None:
        """``apply_local`` of several plans at once, ``items`` a sequence of
        (plan, vectors, result, potential, kinetic, workspace) in one
        precision: grouped launches of the FFT stages (qepw_apply_local_group);
        each plan's workspace holds all its bands."""
        items = [it for it in items if it[1].shape[0] > 0]
        if not items:
            return
        n = len(items)
        dtype = items[0][1].dtype
        prec = precision_code(dtype)
        plans, vecs, ldv, nbs, pots, kins, outs, ldr, bufs, sizes = ([] for _ in range(10))
        for plan, vectors, result, potential, kinetic, ws in items:
            if (vectors.dtype != dtype or result.dtype != dtype
                    or potential.dtype != real_dtype(dtype)):
                raise ValueError("local-potential operands disagree in precision")
            nb = int(vectors.shape[0])
            need = plan.workspace_bytes(dtype, nb)
            buf = ws.get(need)
            plans.append(plan.handle)
            vecs.append(vectors.data_ptr())
            ldv.append(ld_of(vectors))
            nbs.append(nb)
            pots.append(potential.data_ptr())
            kins.append(None if kinetic is None else kinetic.data_ptr())
            outs.append(result.data_ptr())
            ldr.append(ld_of(result))
            bufs.append(buf.data_ptr())
            sizes.append(min(buf.numel(), need))
        P = ctypes.c_void_p * n
        rc = nat.load().qepw_apply_local_group(
            n, P(*plans), prec, P(*vecs), (ctypes.c_int64 * n)(*ldv), (ctypes.c_int * n)(*nbs),
            P(*pots), P(*kins), P(*outs), (ctypes.c_int64 * n)(*ldr), P(*bufs),
            (ctypes.c_size_t * n)(*sizes), stream_handle(items[0][1].device))
        nat.check(rc, "qepw_apply_local_group")

    def inverse_real(self, potential_g: torch.Tensor, ws: Workspace) -> torch.Tensor:
        """(n3, n2, n1) float64 = Re backward FFT of a C-order complex grid."""
        if potential_g.dtype != C128 or tuple(potential_g.shape) != self.shape:
            raise ValueError("potential_g must be a complex128 grid of the plan's shape")
        g = potential_g.contiguous()
        n1, n2, n3 = self.shape
        out = torch.empty((n3, n2, n1), dtype=torch.float64, device=self.device)
        lib = nat.load()
        need = lib.qepw_inverse_real_workspace(self.handle)
        buf = ws.get(need)
        rc = lib.qepw_inverse_real(self.handle, g.data_ptr(), out.data_ptr(), buf.data_ptr(),
                                   buf.numel(), stream_handle(self.device))
        nat.check(rc, "qepw_inverse_real")
        return out
