"""Loads the in-tree libmomfbd_hip.so and declares its C-ABI (include/momfbd_hip.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmomfbd_hip.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)

MFBD_OK, MFBD_EINVAL, MFBD_ENOMEM, MFBD_EDEVICE, MFBD_ENONFINITE, MFBD_ESTATE, MFBD_ESCRATCH = 0, 1, 2, 3, 4, 5, 6
FP64, FP32, MIXED = 0, 1, 2
PATCH_CONVERGED, PATCH_MAXITER, PATCH_FAILED, PATCH_NO_PROGRESS, PATCH_NONFINITE = 0, 1, 2, 3, 4
GRAD_FD, GRAD_VOGEL, GRAD_ADJOINT = 1, 2, 3
GRAD_SCHEDULE_REFERENCE = 0
GETSTEP_STEEPEST, GETSTEP_CONJUGATE, GETSTEP_BFGS, GETSTEP_BFGS_INV = 1, 2, 3, 4
SAVE_COBJ, SAVE_PSF, SAVE_PSF_AVG, SAVE_RESIDUAL = 2, 32, 64, 128
BASIS_ZERNIKE, BASIS_KL = 0, 1


class Geometry(C.Structure):
    _fields_ = [("s", C.c_int), ("P", C.c_int), ("K", C.c_int), ("nObj", C.c_int), ("nCh", C.c_int),
                ("nImg", C.c_int), ("dataY", C.c_int), ("dataX", C.c_int), ("startY", C.c_int),
                ("startX", C.c_int)]


class Params(C.Structure):
    _fields_ = [("FTOL", C.c_double), ("EPS", C.c_double), ("graddiffStep", C.c_double),
                ("regAlpha", C.c_double), ("minIter", C.c_int), ("maxIter", C.c_int), ("nDoneIter", C.c_int),
                ("modeStart", C.c_int), ("modeStep", C.c_int), ("apodSize", C.c_int), ("noFilter", C.c_int),
                ("precision", C.c_int), ("maxBatch", C.c_int), ("gradient", C.c_int),
                ("getstep", C.c_int), ("linePhases", C.c_int), ("lineRounds", C.c_int),
                ("lineGraph", C.c_int)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("nF", C.c_int), ("nDF", C.c_int),
                ("failCount", C.c_int), ("initialMetric", C.c_double), ("finalMetric", C.c_double)]


class Momfbd(C.Structure):
    _fields_ = [("version", C.c_char_p), ("modified", C.c_char_p), ("time", C.c_char_p), ("date", C.c_char_p),
                ("nCh", C.c_int), ("clip", C.POINTER(C.c_int16)), ("nPH", C.c_int), ("nModes", C.c_int),
                ("pix2cf", C.c_float), ("cf2pix", C.c_float), ("pupil", fp), ("modes", fp),
                ("nPatchesY", C.c_int), ("nPatchesX", C.c_int), ("nPoints", C.c_int),
                ("region", ip), ("nim", ip), ("dx", ip), ("dy", ip), ("img", fp),
                ("npsf", C.c_int), ("psf", fp), ("nobj", C.c_int), ("cobj", fp), ("nres", C.c_int), ("res", fp),
                ("nalpha", C.c_int), ("nm", C.c_int), ("alpha", fp), ("ndiv", C.c_int), ("div", fp),
                ("nFileNames", C.c_int), ("fileNames", C.POINTER(C.c_char_p))]


_lib = None

# every exported symbol and its (restype, argtypes); include/momfbd_hip.h declares the same set
SIGNATURES = {
    "mfbd_ctx_create": (C.c_int, [C.POINTER(Geometry), C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]),
    "mfbd_ctx_destroy": (C.c_int, [C.c_void_p]),
    "mfbd_last_error": (C.c_char_p, [C.c_void_p]),
    "mfbd_set_structure": (C.c_int, [C.c_void_p, ip, ip, dp, dp, dp]),
    "mfbd_set_basis": (C.c_int, [C.c_void_p, dp, dp, ip, dp]),
    "mfbd_set_constraints": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, ip, ip, dp]),
    "mfbd_load_patches": (C.c_int, [C.c_void_p, C.c_int, fp, dp, dp, dp]),
    "mfbd_solve": (C.c_int, [C.c_void_p, dp, C.POINTER(Result), dp, fp]),
    "mfbd_init_patches": (C.c_int, [C.c_void_p]),
    "mfbd_eval_metric": (C.c_int, [C.c_void_p, C.c_int, ip, dp, dp]),
    "mfbd_eval_grad": (C.c_int, [C.c_void_p, C.c_int, ip, dp, C.c_int, C.c_int, dp]),
    "mfbd_restore": (C.c_int, [C.c_void_p, C.c_int, ip, dp, fp]),
    "mfbd_restore_outputs": (C.c_int, [C.c_void_p, C.c_int, ip, dp, C.c_int, fp, fp, fp]),
    "mfbd_debug_get": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, dp, ip]),
    "mfbd_kernel_stats": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, dp, ip, C.POINTER(C.c_longlong)]),
    "mfbd_set_patch_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_ubyte), C.c_int, C.c_int, ip, ip, dp]),
    "mfbd_solve_group": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(dp), C.POINTER(C.c_void_p),
                                   C.POINTER(dp), C.POINTER(fp)]),
    "mfbd_write_ana": (C.c_int, [C.c_char_p, C.c_int, C.c_int, ip, C.c_void_p, C.c_char_p]),
    "mfbd_write_fits": (C.c_int, [C.c_char_p, C.c_int, C.c_int, ip, C.c_void_p]),
    "mfbd_write_ana_compressed": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_char_p, C.c_int]),
    "mfbd_file_last_error": (C.c_char_p, []),
    "mfbd_write_momfbd": (C.c_int, [C.c_char_p, C.POINTER(Momfbd)]),
    "mfbd_mosaic_size": (C.c_int, [C.c_int, C.c_int, ip, ip, ip, ip]),
    "mfbd_mosaic": (C.c_int, [C.c_int, C.c_int, fp, ip, ip, C.c_int, C.c_int, fp, ip, ip, ip, ip]),
    "mfbd_mosaic_noalloc": (C.c_int, [C.c_int, C.c_int, fp, ip, ip, C.c_int, C.c_int, fp, ip, ip, ip, ip]),
    "mfbd_setup_last_error": (C.c_char_p, []),
    "mfbd_pupil_size": (C.c_int, [C.c_double, C.c_int, C.c_double, C.c_double, dp, dp, ip]),
    "mfbd_make_pupil": (C.c_int, [C.c_int, C.c_double, C.c_double, dp, dp]),
    "mfbd_pupil_support": (C.c_int, [C.c_int, dp, ip, ip, ip, ip, ip, dp]),
    "mfbd_noll_covariance": (C.c_int, [C.c_int, C.c_int, dp]),
    "mfbd_make_modes": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                  C.c_double, dp, dp, ip]),
    "mfbd_make_modes_device": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, ip, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.c_int, dp, dp, ip]),
    "mfbd_basis_last_error": (C.c_char_p, []),
    "mfbd_normalize_modes": (C.c_int, [C.c_int, C.c_int, dp, dp, C.c_double, C.c_double]),
    "mfbd_measure_jacobian": (C.c_int, [C.c_int, C.c_int, dp, dp, C.c_double, ip, dp, dp]),
    "mfbd_make_constraints": (C.c_int, [C.c_int, ip, ip, ip, C.c_int, ip, ip, ip, ip, ip, dp, C.c_int]),
}


def lib():
    """The loaded library.  Raises if the HIP extension has not been built: there is
    no CPU fallback for the solver."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; run redux_amd.build.build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, t=dp):
    return a.ctypes.data_as(t)
