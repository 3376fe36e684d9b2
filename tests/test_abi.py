"""The C-ABI library loads and exports every symbol include/momfbd_hip.h declares."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "momfbd_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mfbd_\w+)\s*\(", txt, re.M)))


def test_header_and_binding_agree():
    from redux_amd import _lib
    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_symbol():
    from redux_amd import _lib
    try:
        L = _lib.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_errors_without_gpu_work():
    """Argument validation returns codes without touching a device."""
    import ctypes as C
    from redux_amd import _lib
    try:
        L = _lib.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    g = _lib.Geometry(64, 32, 20, 1, 1, 8, 74, 74, 5, 5)
    prm = _lib.Params(1e-3, 1e-10, 1e-2, 0, 5, 500, 3, 5, 5, -1, 0, 0, 1, 0)
    h = C.c_void_p()
    assert L.mfbd_ctx_create(C.byref(g), C.byref(prm), 0, C.byref(h)) == 0
    # solve before structure/basis/constraints/patches -> ESTATE
    assert L.mfbd_solve(h, None, None, None, None) == _lib.MFBD_ESTATE
    assert L.mfbd_set_constraints(h, 10, 5, 1, None, None, None) == _lib.MFBD_EINVAL
    bad = _lib.Geometry(64, 32, 20, 1, 1, 8, 60, 60, 5, 5)     # block smaller than patch + start
    h2 = C.c_void_p()
    assert L.mfbd_ctx_create(C.byref(bad), C.byref(prm), 0, C.byref(h2)) == _lib.MFBD_EINVAL
    assert L.mfbd_ctx_destroy(h) == 0
