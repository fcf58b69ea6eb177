"""CPU checks of the drop-in boundary: libflgpu.so loads without a GPU and
exports every entry point include/flgpu.h declares; the Python binding
declares exactly that set."""
import ctypes
import os
import re

from gnu_femlisp_amd import flgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "flgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\bint\s+(flgpu_\w+)\s*\(", src))


def test_library_exports_header():
    syms = header_symbols()
    assert len(syms) >= 40
    lib = ctypes.CDLL(flgpu.LIB_PATH)
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_matches_header():
    assert header_symbols() == set(flgpu.SIGNATURES)


def test_errors_without_device_are_reported():
    """Argument errors come back as negative status with a message."""
    L = flgpu.lib()
    assert L.flgpu_ctx_create(0, None) < 0
    assert "NULL" in flgpu.last_error()
    assert L.flgpu_vec_create(None, 4, 0, None) < 0
