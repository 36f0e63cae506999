"""CPU-side checks of the C ABI: the library loads and exports every entry
point the public header declares, and the Python binding covers them."""
import ctypes
import os
import re

from amrsolver_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(amr_[a-z_0-9]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIBPATH)
    names = declared_entry_points()
    assert len(names) >= 14
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    assert declared_entry_points() == set(capi.SIGNATURES)


def test_struct_layouts_match_header():
    # sizes of the POD structs as laid out by the C compiler (x86-64 SysV)
    assert ctypes.sizeof(capi.Fab) == 64
    assert ctypes.sizeof(capi.Scheme) == 32
    assert ctypes.sizeof(capi.Geom) == 136
    assert ctypes.sizeof(capi.Problem) == 16


def test_stage_coefficients_cpu():
    # pure host arithmetic: no device needed
    n, mode, a, b, c, d = capi.stage_coeffs(3, 3, 0.1, 1)
    assert (n, mode, a, b, c, d) == (3, 0, 0.75, 0.25, 1.0, 0.1)
    n, mode, a, b, c, d = capi.stage_coeffs(3, 4, 0.3, 2)
    assert (n, mode, a, c, d) == (4, 1, 2.0 / 3, 1.0 / 3, 0.3 / 6)
    n, mode, a, b, c, d = capi.stage_coeffs(2, 4, 0.3, 3)
    assert (n, mode, a, b, c, d) == (4, 0, 1.0 / 4, 1.0 / 4, 3.0, 0.3)


def test_errors_are_reported_without_device():
    import pytest
    with pytest.raises(capi.AmrError, match="unsupported order_rk"):
        capi.stage_coeffs(4, 10, 0.1, 0)


def test_user_source_prelude_repeats_the_closures_struct():
    """The run-time compiled user_source sees amr_closures_t through a copy
    of the header's definition: the two texts must agree (layout of the
    struct passed by value to the compiled kernel)."""
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "amrsolver_hip.h")).read()
    src = open(os.path.join(root, "amrsolver_amd", "csrc", "user_source.cpp")).read()
    pat = re.compile(r"typedef struct \{(.*?)\} amr_closures_t;", re.S)
    norm = lambda t: " ".join(t.split())
    assert norm(pat.search(hdr).group(1)) == norm(pat.search(src).group(1))
