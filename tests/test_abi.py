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


def _lisp_layout(name):
    """(field, offset, type) triples of a struct layout in lisp/fl-gpu.lisp"""
    src = open(os.path.join(ROOT, "lisp", "fl-gpu.lisp")).read()
    m = re.search(r"\(defparameter \+%s\+\s*'\((.*?)\)\s*\"" % re.escape(name), src, re.S)
    assert m, name
    return [(f.replace("-", "_"), int(o), t)
            for f, o, t in re.findall(r"\(([a-z-]+) (\d+) :(int|double)\)", m.group(1))]


def test_lisp_struct_layouts_match_the_header():
    """the femlisp side packs flgpu_mg_params, flgpu_solve_params and flgpu_solve_report by
    byte offset; the offsets and sizes must be those of the C structs (ctypes lays a
    Structure out as the C compiler does)"""
    for lname, cls, size in (("mg-params-layout", flgpu.MgParams, 56),
                             ("solve-params-layout", flgpu.SolveParams, 48),
                             ("solve-report-layout", flgpu.SolveReport, 56)):
        lay = _lisp_layout(lname)
        assert ctypes.sizeof(cls) == size, (cls, ctypes.sizeof(cls))
        assert [f for f, _, _ in lay] == [f for f, _ in cls._fields_]
        for f, off, typ in lay:
            assert getattr(cls, f).offset == off, (lname, f)
            ct = dict(cls._fields_)[f]
            assert ct is (ctypes.c_int if typ == "int" else ctypes.c_double), (lname, f)


def test_lisp_binding_forms_are_balanced():
    """no Lisp is installed here: check at least that every top-level form of the binding
    closes (strings, comments and #\\ characters skipped) and that every foreign function
    it calls is declared by def-function and exported by the library"""
    src = open(os.path.join(ROOT, "lisp", "fl-gpu.lisp")).read()
    depth, i, forms = 0, 0, 0
    while i < len(src):
        c = src[i]
        if c == ";":
            i = src.index("\n", i)
        elif c == '"':
            i += 1
            while src[i] != '"':
                i += 2 if src[i] == "\\" else 1
        elif src.startswith("#\\", i):
            i += 2
        elif c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
            assert depth >= 0, src[max(0, i - 80):i + 1]
            forms += depth == 0
        i += 1
    assert depth == 0 and forms > 30
    declared = set(re.findall(r'def-function \("(flgpu_\w+)"', src))
    used = set(n.replace("-", "_") for n in re.findall(r"#'(flgpu-[a-z0-9-]+)", src))
    used |= set(n.replace("-", "_") for n in re.findall(r"\((flgpu-[a-z0-9-]+) \(gpu-handle", src))
    assert used <= declared, used - declared
    assert declared <= header_symbols()
