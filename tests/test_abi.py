"""The C ABI library builds for gfx950, loads, and exports every entry point
the header declares, with the ctypes signatures the Python host binds (no
GPU calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "ocsc.h")
LIB = os.path.join(REPO, "ocelot_amd", "libocsc.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ocsc_\w+)\s*\(", txt)))


def test_header_declares_the_api():
    names = declared()
    for must in ("ocsc_create", "ocsc_destroy", "ocsc_apply", "ocsc_el_field", "ocsc_sym_kernel",
                 "ocsc_potential", "ocsc_last_error", "ocsc_drift", "ocsc_track_drift"):
        assert must in names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from ocelot_amd import build
        build.build(verbose=False)
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_symbol(lib):
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ocsc_\w+)", out))
    assert set(declared()) <= exported


def test_python_binding_covers_the_header():
    from ocelot_amd import _lib
    assert set(_lib.SIGNATURES) == set(declared())
    _lib.load(LIB)


def test_library_has_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-S", LIB], capture_output=True, text=True).stdout
    assert ".hip_fatbin" in out
    blob = open(LIB, "rb").read()
    assert b"gfx950" in blob


def test_last_error_is_empty_before_any_call(lib):
    lib.ocsc_last_error.restype = ctypes.c_char_p
    assert lib.ocsc_last_error() == b""


def test_missing_library_fails_loudly(tmp_path):
    from ocelot_amd import _lib
    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved


@pytest.mark.parametrize("compiler,std", [("gcc", "-std=c11"), ("g++", "-std=c++17")])
def test_documented_c_usage_compiles_against_the_header(tmp_path, compiler, std):
    """The C caller example of the integration notes, wrapped in a function,
    compiles warning-free against include/ocsc.h as C and as C++ (argument
    counts and types of every entry point it calls)."""
    txt = open(os.path.join(REPO, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", txt, re.S)
    assert blocks, "no C example found"
    body = "".join(b.replace('#include "ocsc.h"\n', "") for b in blocks)
    src = ('#include <stdio.h>\n#include <stdint.h>\n#include "ocsc.h"\n'
           "void example(double* d_r, double* d_q, int64_t n, void* stream, const double* draws, double* tws) {\n"
           + body + "}\n")
    f = tmp_path / ("ex.c" if compiler == "gcc" else "ex.cpp")
    f.write_text(src)
    r = subprocess.run([compiler, std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(REPO, "include"),
                        str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
