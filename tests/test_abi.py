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
    neg = _lib.Params(1e-3, 1e-10, 1e-2, 0, 5, 500, 3, 5, 5, -1, 0, 0, 1, 0, 2, 1, -1)   # lineRounds < 0
    assert L.mfbd_ctx_create(C.byref(g), C.byref(neg), 0, C.byref(h2)) == _lib.MFBD_EINVAL
    assert L.mfbd_ctx_destroy(h) == 0


def test_header_constants_match_binding():
    """Every status, precision and gradient constant the header defines has the same
    value in the ctypes binding."""
    import re
    from redux_amd import _lib as L
    text = open(os.path.join(os.path.dirname(__file__), "..", "include", "momfbd_hip.h")).read()
    defs = dict(re.findall(r"#define (MFBD_\w+) (\d+)", text))
    pairs = {"MFBD_OK": L.MFBD_OK, "MFBD_EINVAL": L.MFBD_EINVAL, "MFBD_ENOMEM": L.MFBD_ENOMEM,
             "MFBD_EDEVICE": L.MFBD_EDEVICE, "MFBD_ENONFINITE": L.MFBD_ENONFINITE, "MFBD_ESTATE": L.MFBD_ESTATE,
             "MFBD_FP64": L.FP64, "MFBD_FP32": L.FP32, "MFBD_MIXED": L.MIXED, "MFBD_GRAD_FD": L.GRAD_FD,
             "MFBD_GRAD_VOGEL": L.GRAD_VOGEL, "MFBD_GRAD_ADJOINT": L.GRAD_ADJOINT,
             "MFBD_GRAD_SCHEDULE_REFERENCE": L.GRAD_SCHEDULE_REFERENCE, "MFBD_BASIS_ZERNIKE": L.BASIS_ZERNIKE,
             "MFBD_BASIS_KL": L.BASIS_KL, "MFBD_GETSTEP_STEEPEST": L.GETSTEP_STEEPEST,
             "MFBD_GETSTEP_CONJUGATE": L.GETSTEP_CONJUGATE, "MFBD_GETSTEP_BFGS": L.GETSTEP_BFGS,
             "MFBD_GETSTEP_BFGS_INV": L.GETSTEP_BFGS_INV, "MFBD_SAVE_COBJ": L.SAVE_COBJ,
             "MFBD_SAVE_PSF": L.SAVE_PSF, "MFBD_SAVE_PSF_AVG": L.SAVE_PSF_AVG, "MFBD_SAVE_RESIDUAL": L.SAVE_RESIDUAL}
    for name, val in pairs.items():
        assert int(defs[name]) == val, name
