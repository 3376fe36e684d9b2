"""Per-job setup through the C-ABI (pupil, modes, Jacobian, nullspace).

Function names and return shapes match what redux_amd.synth expects from a setup
provider: calculate_pupil_size, generate_pupil, mode_set, normalize_modes,
measure_jacobian, pix2cf, noll_cov, constraints.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import BASIS_KL, BASIS_ZERNIKE, ip, lib, ptr


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().mfbd_setup_last_error().decode()}")


def calculate_pupil_size(wavelength, nPixels, telescopeD, arcSecsPerPixel):
    fc, r = np.zeros(1), np.zeros(1)
    P = np.zeros(1, dtype=np.int32)
    _check(lib().mfbd_pupil_size(wavelength, nPixels, telescopeD, arcSecsPerPixel, ptr(fc), ptr(r), ptr(P, ip)),
           "mfbd_pupil_size")
    return float(fc[0]), float(r[0]), int(P[0])


def generate_pupil(P, radius, co_radius=0.0):
    pup = np.zeros((P, P))
    area = np.zeros(1)
    _check(lib().mfbd_make_pupil(P, radius, co_radius, ptr(pup), ptr(area)), "mfbd_make_pupil")
    nPS, nOS = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
    ps, pio = np.zeros(P * P, dtype=np.int32), np.zeros(P * P, dtype=np.int32)
    os_ = np.zeros(4 * P * P, dtype=np.int32)
    a2 = np.zeros(1)
    _check(lib().mfbd_pupil_support(P, ptr(pup), ptr(nPS, ip), ptr(ps, ip), ptr(pio, ip), ptr(nOS, ip),
                                    ptr(os_, ip), ptr(a2)), "mfbd_pupil_support")
    n, m = int(nPS[0]), int(nOS[0])
    return dict(pupil=pup, ps=ps[:n].copy(), pio=pio[:n].copy(), os=os_[:m].copy(), area=float(a2[0]))


def mode_set(P, radius, modes, basis="zernike", kl_min=2, kl_max=2000, kl_cutoff=1e-3, angle=0.0):
    ids = np.ascontiguousarray(modes, dtype=np.int32)
    B = np.zeros((len(ids), P, P))
    rms = np.zeros(len(ids))
    tilt = np.zeros(2, dtype=np.int32)
    b = BASIS_KL if basis == "kl" else BASIS_ZERNIKE
    _check(lib().mfbd_make_modes(P, radius, angle, len(ids), ptr(ids, ip), b, kl_min, kl_max, kl_cutoff, ptr(B),
                                 ptr(rms), ptr(tilt, ip)), "mfbd_make_modes")
    return B, rms, (int(tilt[0]), int(tilt[1]))


def mode_set_device(P, radius, modes, basis="zernike", kl_min=2, kl_max=2000, kl_cutoff=1e-3, angle=0.0, device=0):
    """mode_set with the Zernike evaluations, normalisation and KL sums on the GPU (FP64)."""
    ids = np.ascontiguousarray(modes, dtype=np.int32)
    B = np.zeros((len(ids), P, P))
    rms = np.zeros(len(ids))
    tilt = np.zeros(2, dtype=np.int32)
    b = BASIS_KL if basis == "kl" else BASIS_ZERNIKE
    rc = lib().mfbd_make_modes_device(P, radius, angle, len(ids), ptr(ids, ip), b, kl_min, kl_max, kl_cutoff, device,
                                      ptr(B), ptr(rms), ptr(tilt, ip))
    if rc != 0:
        raise RuntimeError(f"mfbd_make_modes_device failed ({rc}): {lib().mfbd_basis_last_error().decode()}")
    return B, rms, (int(tilt[0]), int(tilt[1]))


def normalize_modes(B, pupil, area, scale=1.0):
    Bc = np.ascontiguousarray(B, dtype=np.float64).copy()
    K, P = Bc.shape[0], Bc.shape[1]
    pup = np.ascontiguousarray(pupil, dtype=np.float64)
    _check(lib().mfbd_normalize_modes(P, K, ptr(Bc), ptr(pup), area, scale), "mfbd_normalize_modes")
    return Bc


def measure_jacobian(B, pupil, scale):
    Bc = np.ascontiguousarray(B, dtype=np.float64)
    pup = np.ascontiguousarray(pupil, dtype=np.float64)
    tilt = np.zeros(2, dtype=np.int32)
    J, Ji = np.zeros(4), np.zeros(4)
    _check(lib().mfbd_measure_jacobian(Bc.shape[1], Bc.shape[0], ptr(Bc), ptr(pup), scale, ptr(tilt, ip), ptr(J),
                                       ptr(Ji)), "mfbd_measure_jacobian")
    return (int(tilt[0]), int(tilt[1])), J.reshape(2, 2), Ji.reshape(2, 2)


def pix2cf(arcsecperpix, telescope_d):
    """util::pix2cf (src/lib/momfbd/util.cpp:126-129)."""
    return (math.pi * math.pi) / 648000.0 * (arcsecperpix * telescope_d)


def noll_cov(i, j):
    v = np.zeros(1)
    _check(lib().mfbd_noll_covariance(int(i), int(j), ptr(v)), "mfbd_noll_covariance")
    return float(v[0])


def constraints(n_images_per_obj_ch, wavefronts, K):
    nObj = len(n_images_per_obj_ch)
    nch = np.array([len(c) for c in n_images_per_obj_ch], dtype=np.int32)
    nfr = np.array([f for c in n_images_per_obj_ch for f in c], dtype=np.int32)
    wfs = np.array([w for obj in wavefronts for ch in obj for w in ch], dtype=np.int32)
    nImg = int(nfr.sum())
    cap = nImg * K * (nImg + 2)
    rows, cols = np.zeros(cap, dtype=np.int32), np.zeros(cap, dtype=np.int32)
    vals = np.zeros(cap)
    nPar, nFree, nnz = (np.zeros(1, dtype=np.int32) for _ in range(3))
    _check(lib().mfbd_make_constraints(nObj, ptr(nch, ip), ptr(nfr, ip), ptr(wfs, ip), K, ptr(nPar, ip),
                                       ptr(nFree, ip), ptr(nnz, ip), ptr(rows, ip), ptr(cols, ip), ptr(vals), cap),
           "mfbd_make_constraints")
    n = int(nnz[0])
    return int(nPar[0]), int(nFree[0]), rows[:n].copy(), cols[:n].copy(), vals[:n].copy()
