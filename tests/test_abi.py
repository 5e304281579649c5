"""CPU checks of the C ABI: the library is built, loads, and exports every
entry point include/symcode.h declares; argument validation happens before
any GPU work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "symcode.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sc_\w+)\s*\(", text, re.M)))


def test_header_declares_core_entry_points():
    names = declared()
    for must in ("sc_create", "sc_convectiondiffusion", "sc_wray3", "sc_rk4", "sc_to_phys",
                 "sc_to_spec", "sc_les_rhs", "sc_cutoff", "sc_closure_dynsmag"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from symcode import _lib

    L = _lib.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert set(declared()) - {"sc_last_error", "sc_cnn_last_error", "sc_pmlp_last_error"} <= set(_lib.SIGNATURES)


def test_create_validates_arguments_without_gpu():
    from symcode import _lib

    L = _lib.lib()
    ctx = ctypes.c_void_p()
    assert L.sc_create(3, 7, 1.0, 64, 0, ctypes.byref(ctx)) == 1
    assert b"even" in L.sc_last_error(None)
    assert L.sc_create(4, 8, 1.0, 64, 0, ctypes.byref(ctx)) == 1
    assert L.sc_create(3, 8, 1.0, 16, 0, ctypes.byref(ctx)) == 1
    assert not ctx.value
    # NULL context is rejected, not dereferenced
    assert L.sc_project(None, None) == 1
    assert L.sc_set_sync(None, 1) == 1


def test_graft_entry_importable():
    import __graft_entry__ as ge

    assert callable(ge.build) and callable(ge.smoke)
