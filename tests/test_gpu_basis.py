"""Mode basis built on the device (mfbd_make_modes_device) against the host builder
(mfbd_make_modes, itself pinned against the numpy setup oracle and the Zernike / KL
invariants in the CPU tests).  The host evaluates the radial polynomials in long double,
the device in FP64, so they agree to rounding: within 1e-12 of the unit RMS the modes
are normalised to (measured: 2e-14 at most, up to the Zernikes of radial order ~60 that a
2000-mode KL expansion mixes in).  The device modes are also checked directly: unit RMS
over the pupil and near-orthogonality of distinct Zernikes."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(P, radius, modes, basis, kl_max):
    from redux_amd import setup as S
    t0 = time.perf_counter()
    Bh, rh, th = S.mode_set(P, radius, modes, basis=basis, kl_max=kl_max)
    t1 = time.perf_counter()
    Bd, rd, td = S.mode_set_device(P, radius, modes, basis=basis, kl_max=kl_max)
    t2 = time.perf_counter()
    print(f"P={P} {basis} K={len(modes)} kl_max={kl_max}: host {t1 - t0:.3f}s device {t2 - t1:.3f}s "
          f"max|diff| {np.abs(Bh - Bd).max():.3e}", flush=True)
    return (Bh, rh, th), (Bd, rd, td)


@pytest.mark.parametrize("P,radius,basis,modes,kl_max", [
    (64, 32.0, "zernike", list(range(2, 52)), 2000),
    (64, 32.0, "kl", list(range(2, 102)), 2000),
    (32, 15.5, "kl", list(range(2, 38)), 200),
    (128, 64.0, "kl", list(range(2, 152)), 2000),     # the basis of the 128-pixel pupil, 150 KL modes
])
def test_device_basis_matches_the_host_basis(P, radius, basis, modes, kl_max):
    (Bh, rh, th), (Bd, rd, td) = _both(P, radius, modes, basis, kl_max)
    assert th == td
    np.testing.assert_array_equal(rh, rd)
    assert np.abs(Bh - Bd).max() <= 1e-12


def test_device_zernikes_are_orthonormal_over_the_pupil():
    from redux_amd import setup as S
    P, radius = 64, 32.0
    pup = S.generate_pupil(P, radius)
    Bd, _, _ = S.mode_set_device(P, radius, list(range(2, 38)), basis="zernike")
    w = pup["pupil"].ravel()
    G = (Bd.reshape(len(Bd), -1) * w) @ Bd.reshape(len(Bd), -1).T / pup["area"]
    np.testing.assert_allclose(np.diag(G), 1.0, atol=1e-12)
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() < 0.05          # a discrete, anti-aliased pupil: nearly orthogonal


def test_bad_arguments_are_refused():
    from redux_amd import setup as S
    with pytest.raises(RuntimeError, match="KL index"):
        S.mode_set_device(32, 15.5, [2, 3, 4000], basis="kl", kl_max=200)
