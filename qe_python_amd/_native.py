"""ctypes binding of ``libqepw_hip.so`` (the C ABI of ``include/qepw.h``).

This is the device counterpart of the reference's ``_load_native_fft()``
(``qepy_pw/basis.py:48-74``): it loads the in-tree gfx950 library once and
raises if it is missing -- there is no CPU fallback.  Error codes returned
by the library become the reference's exception types: ``ValueError`` for
invalid arguments, ``RuntimeError`` for HIP / unsupported-shape failures.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_int64, c_size_t, c_void_p, POINTER
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libqepw_hip.so"

C128 = 0
C64 = 1
OP_N = 0
OP_C = 1
OP_BC = 2

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_UNSUPPORTED = -3
ERR_WORKSPACE = -4

# (name, restype, argtypes) of every exported symbol, in header order.
SIGNATURES = [
    ("qepw_last_error", ctypes.c_char_p, []),
    ("qepw_abi_version", c_int, []),
    ("qepw_plan_create", c_int, [c_void_p, c_int64, c_int, c_int, c_int, c_int, POINTER(c_void_p)]),
    ("qepw_plan_destroy", c_int, [c_void_p]),
    ("qepw_plan_query", c_int, [c_void_p, POINTER(c_int64)]),
    ("qepw_plan_permutation", c_int, [c_void_p, c_void_p]),
    ("qepw_plan_supported", c_int, [c_void_p, c_int]),
    ("qepw_plan_device_permutation", c_void_p, [c_void_p]),
    ("qepw_apply_local_workspace", c_size_t, [c_void_p, c_int, c_int]),
    ("qepw_apply_local", c_int, [c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p, c_void_p,
                                 c_void_p, c_int64, c_void_p, c_size_t, c_void_p]),
    ("qepw_apply_local_timed", c_int, [c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p,
                                       c_void_p, c_void_p, c_int64, c_void_p, c_size_t, c_void_p,
                                       c_void_p, c_void_p]),
    ("qepw_apply_local_group", c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p]),
    ("qepw_inverse_real_workspace", c_size_t, [c_void_p]),
    ("qepw_inverse_real", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    ("qepw_gemm_workspace", c_size_t, [c_int, c_int, c_int64, c_int64, c_int64]),
    ("qepw_gemm", c_int, [c_int, c_int, c_int64, c_int64, c_int64, POINTER(c_double), c_void_p,
                          c_int64, c_void_p, c_int64, POINTER(c_double), c_void_p, c_int64,
                          c_void_p, c_size_t, c_void_p]),
    ("qepw_gemm_variant_count", c_int, []),
    ("qepw_gemm_plan", c_int, [c_int, c_int, c_int64, c_int64, c_int64, POINTER(c_int64)]),
    ("qepw_gemm_set_rule", c_int, [c_int, c_int, POINTER(c_int64), c_int, c_int64]),
    ("qepw_gemm_clear_rules", c_int, []),
    ("qepw_gemm_static_scratch_bytes", c_size_t, []),
    ("qepw_gemm_static", c_int, [c_int, c_int, c_int64, c_int64, c_int64, POINTER(c_double),
                                 c_void_p, c_int64, c_void_p, c_int64, POINTER(c_double),
                                 c_void_p, c_int64, c_void_p]),
    ("qepw_column_squared_norms", c_int, [c_int, c_void_p, c_int64, c_int64, c_int, c_void_p,
                                          c_void_p]),
    ("qepw_column_inner_products", c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64, c_int64,
                                           c_int, c_void_p, c_void_p]),
    ("qepw_subtract_band_energies", c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64,
                                            c_void_p, c_int64, c_int, c_void_p]),
    ("qepw_qe_precondition", c_int, [c_int, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_int64, c_int64, c_int, c_void_p]),
    ("qepw_normalize_selected_columns", c_int, [c_int, c_void_p, c_int64, c_void_p, c_void_p,
                                                c_void_p, c_int64, c_int64, c_int, c_void_p]),
    ("qepw_restart_applied", c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
                                     c_void_p, c_int64, c_int64, c_int, c_void_p]),
    ("qepw_permute_rows", c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p,
                                  c_int64, c_int64, c_int, c_void_p]),
    ("qepw_copy_columns", c_int, [c_int, c_int, c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                  c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    ("qepw_set_block", c_int, [c_int, c_void_p, c_int64, c_int64, c_int, c_double, c_void_p,
                               c_void_p]),
    ("qepw_accumulate_projector_diagonal", c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int64,
                                                   c_int, c_double, c_void_p, c_void_p]),
    ("qepw_accumulate_density_workspace", c_size_t, [c_void_p, c_int, c_int]),
    ("qepw_accumulate_density", c_int, [c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p,
                                        c_double, c_void_p, c_void_p, c_size_t, c_void_p]),
    ("qepw_scale_rows", c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p,
                                c_int64, c_int64, c_int, c_void_p]),
    ("qepw_build_projectors", c_int, [c_int, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_int,
                                      c_void_p, c_int64, c_int64, c_void_p]),
    ("qepw_column_axpby", c_int, [c_int, c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                  c_void_p, c_int64, c_int, c_void_p]),
    ("qepw_column_diagonal_expectations", c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64,
                                                  c_int, c_void_p, c_void_p]),
    ("qepw_rmm_precondition", c_int, [c_int, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_int64, c_int64, c_int, c_void_p]),
    ("qepw_history_combine", c_int, [c_int, c_void_p, c_int64, c_int64, c_int, c_void_p,
                                     c_void_p, c_void_p, c_int64, c_int64, c_int, c_void_p]),
    ("qepw_phase_blocks", c_int, [c_int, c_void_p, c_int64, c_int, c_void_p, c_int64, c_int,
                                  c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    ("qepw_projector_basis", c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int, c_double,
                                     c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int64,
                                     c_void_p]),
    ("qepw_projector_forces", c_int, [c_int, c_void_p, c_int64, c_int64, c_void_p, c_int64,
                                      c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    ("qepw_hermitian_complete", c_int, [c_void_p, c_int64, c_int, c_int, c_void_p]),
    ("qepw_shift_scale", c_int, [c_void_p, c_int64, c_int, c_double, c_void_p, c_int64, c_void_p,
                                 c_void_p]),
    ("qepw_hermitize", c_int, [c_void_p, c_int64, c_int, c_void_p]),
    ("qepw_sign_check", c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int, c_void_p, c_void_p]),
    ("qepw_set_diagonal", c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p]),
    ("qepw_sign_iteration_workspace", c_size_t, [c_int]),
    ("qepw_sign_iteration", c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int,
                                    POINTER(c_double), c_int, POINTER(c_int), c_void_p,
                                    c_size_t, c_void_p]),
    ("qepw_set_host_lapack", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("qepw_rr_workspace", c_size_t, [c_int]),
    ("qepw_rr_standardize", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                    POINTER(c_int), c_void_p, c_size_t, c_void_p]),
    ("qepw_rr_spectral", c_int, [c_void_p, c_void_p, c_void_p]),
    ("qepw_cholesky_qr", c_int, [c_int, c_void_p, c_int64, c_int64, c_int, c_double, c_void_p,
                                 c_int64, POINTER(c_int), c_void_p, c_size_t, c_void_p]),
]


class RRSpectralArgs(ctypes.Structure):
    """``qepw_rr_spectral_args`` of include/qepw.h."""

    _fields_ = [("n", c_int), ("k", c_int), ("maxsub", c_int), ("count_hint", c_int),
                ("host_single", c_int), ("have_prev", c_int), ("sigma", c_double),
                ("gap", c_double), ("bound_prev", c_double), ("A", c_void_p), ("W", c_void_p),
                ("X", c_void_p), ("X2", c_void_p), ("Y", c_void_p), ("u_prev", c_void_p),
                ("gauss", c_void_p), ("coef", c_void_p), ("rb", c_void_p), ("work", c_void_p),
                ("work_bytes", c_size_t), ("stream", c_void_p)]


class RRSpectralResult(ctypes.Structure):
    """``qepw_rr_spectral_result`` of include/qepw.h."""

    _fields_ = [("status", c_int), ("m_below", c_int), ("ns_steps", c_int),
                ("slow_shift", c_int), ("bound", c_double)]


def _register_host_lapack(lib) -> None:
    """Hand the library the LAPACK routines of SciPy's own LAPACK (the
    function pointers ``scipy.linalg.cython_lapack`` exports), used by the
    native Rayleigh-Ritz steps for their small host factorisations."""
    import scipy.linalg.cython_lapack as cl
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    ptrs = []
    for name in ("zpotrf", "ztrtri", "zheevd", "cheevd"):
        capsule = cl.__pyx_capi__[name]
        ptrs.append(get_ptr(capsule, get_name(capsule)))
    check(lib.qepw_set_host_lapack(*ptrs), "qepw_set_host_lapack")

_LIB: ctypes.CDLL | None = None


class NativeLibraryMissing(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load the gfx950 kernel library (once); raise if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -m qe_python_amd.build` "
            "(the device path has no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _LIB = lib
    _register_host_lapack(lib)
    # the committed GEMM tile table (qe_python_amd.autotune); QEPW_GEMM_TABLE
    # names another table, or "none" for the built-in choice
    table = os.environ.get("QEPW_GEMM_TABLE")
    if table != "none":
        from . import autotune
        autotune.load_table(table)
    return lib


def last_error() -> str:
    return load().qepw_last_error().decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc == OK:
        return
    msg = last_error() or what
    if rc == ERR_ARG:
        raise ValueError(msg)
    raise RuntimeError(msg)


_PAIRS: dict = {}


def complex_pair(z) -> ctypes.Array:
    """(re, im) as a ctypes double[2]; the few coefficients in use are
    cached (the arrays are only read by the library)."""
    z = complex(z)
    pair = _PAIRS.get(z)
    if pair is None:
        pair = (c_double * 2)(z.real, z.imag)
        if len(_PAIRS) < 1024:
            _PAIRS[z] = pair
    return pair
