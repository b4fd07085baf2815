"""The C-ABI library loads and exports every symbol ``include/qepw.h``
declares (CPU only: no compute call without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "qepw.h"
LIB = ROOT / "qe_python_amd" / "libqepw_hip.so"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qepw_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_symbols()
    assert "qepw_apply_local" in names and "qepw_gemm" in names
    assert len(names) >= 20


@pytest.mark.skipif(not LIB.exists(), reason="libqepw_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="libqepw_hip.so not built")
def test_python_binding_covers_the_header():
    from qe_python_amd import _native
    bound = {name for name, _, _ in _native.SIGNATURES}
    assert bound == set(declared_symbols())
    assert _native.load().qepw_abi_version() == 1


@pytest.mark.skipif(not LIB.exists(), reason="libqepw_hip.so not built")
def test_argument_errors_need_no_device():
    from qe_python_amd import _native
    lib = _native.load()
    handle = ctypes.c_void_p()
    rc = lib.qepw_plan_create(None, 0, 8, 8, 8, 0, ctypes.byref(handle))
    assert rc == _native.ERR_ARG
    assert "empty basis" in _native.last_error()
    rc = lib.qepw_plan_create(None, 0, 7, 8, 8, 0, None)
    assert rc == _native.ERR_ARG
