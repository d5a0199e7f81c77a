"""The C-ABI library loads and exports every entry point include/uasd_hip.h
declares; host-only entry points report errors through status codes."""
import ctypes
import os
import re

import numpy as np

from uppasd_amd import _lib as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "uasd_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(uasd_[a-z0-9_]+)\s*\(", txt)))


def test_exports_match_header():
    L = B.lib()
    names = declared()
    assert names, "no declarations found"
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(B.EXPORTS)


def test_version_and_errors_without_device():
    L = B.lib()
    assert L.uasd_version() == 1
    h = ctypes.c_void_p()
    assert L.uasd_ctx_create(ctypes.byref(h), 0, 1, 0) == -1       # bad natom
    assert b"natom" in L.uasd_last_error()
    assert L.uasd_ctx_create(ctypes.byref(h), 10, 1, 7) == -1       # bad precision
    assert L.uasd_set_exchange(None, 0, 0, None, None, None, None) == -1
    assert L.uasd_ctx_destroy(None) == 0


def test_builder_rejects_bad_symmetry():
    import uppasd_amd as U
    lat = U.bcc((2, 2, 2))
    sh = U.Shells.from_lines(lat, [(1, 1, 1, 1, 0.5, 0.5, 0.5, 1.0)])
    try:
        U.build_pair_list(lat, sh, 7, np.ones((2, 1)))
    except B.UasdError as e:
        assert e.code == -4
    else:
        raise AssertionError("sym 7 accepted")


def test_fortran_binding_names_match_header():
    """Every bind(C) name in the Fortran module is declared in the header,
    and the module and the test driver compile and link with amdflang."""
    src = open(os.path.join(ROOT, "uppasd_amd", "fortran", "uasd_hip.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(uasd_[a-z0-9_]+)"\)', src))
    assert bound and bound <= set(declared())
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "fortran"))
    import build_driver
    binary = build_driver.build()
    if binary is None:
        import pytest
        pytest.skip("amdflang not installed")
    assert os.path.exists(binary)
