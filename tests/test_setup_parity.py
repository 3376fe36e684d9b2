"""Product setup (C++ behind the C-ABI, host only) against the oracle's numpy setup."""
import numpy as np
import pytest

from oracle import setup_np as O

S = pytest.importorskip("redux_amd.setup")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from redux_amd import _lib
    try:
        _lib.lib()
    except RuntimeError as e:
        pytest.skip(str(e))


def test_pupil_size():
    for s, arc in [(64, 0.067), (88, 0.059), (128, 0.04)]:
        assert O.calculate_pupil_size(6.302e-7, s, 0.97, arc) == S.calculate_pupil_size(6.302e-7, s, 0.97, arc)


@pytest.mark.parametrize("P,r", [(32, 16.0), (64, 32.0), (64, 29.3)])
def test_pupil_and_supports(P, r):
    a, b = O.generate_pupil(P, r), S.generate_pupil(P, r)
    assert np.abs(a["pupil"] - b["pupil"]).max() == 0
    assert abs(a["area"] - b["area"]) < 1e-9
    for k in ("ps", "pio", "os"):
        assert np.array_equal(a[k], b[k])


def test_zernike_modes():
    B1, r1, t1 = O.mode_set(32, 16.0, list(range(2, 22)))
    B2, r2, t2 = S.mode_set(32, 16.0, list(range(2, 22)))
    assert t1 == t2 == (0, 1)
    assert np.abs(B1 - B2).max() < 1e-12
    assert np.abs(r1 - r2).max() < 1e-15


def test_kl_modes():
    B1, r1, _ = O.mode_set(64, 32.0, list(range(2, 30)), basis="kl", kl_max=40)
    B2, r2, _ = S.mode_set(64, 32.0, list(range(2, 30)), basis="kl", kl_max=40)
    assert np.abs(r1 - r2).max() < 1e-14
    assert np.abs(B1 - B2).max() < 1e-11


def test_normalisation_and_jacobian():
    p = O.generate_pupil(32, 16.0)
    B = O.mode_set(32, 16.0, list(range(2, 22)))[0]
    B1 = O.normalize_modes(B, p["pupil"], p["area"], 1 / 6.3e-7)
    B2 = S.normalize_modes(B, p["pupil"], p["area"], 1 / 6.3e-7)
    assert np.abs(B1 - B2).max() / np.abs(B1).max() < 1e-14
    t1, J1, Ji1 = O.measure_jacobian(B1, p["pupil"], 1e-3)
    t2, J2, Ji2 = S.measure_jacobian(B1, p["pupil"], 1e-3)
    assert t1 == t2
    assert np.abs(J1 - J2).max() < 1e-9 * np.abs(J1).max()
    assert np.abs(Ji1 - Ji2).max() < 1e-9 * np.abs(Ji1).max()


def test_covariance():
    for i, j in [(2, 2), (4, 4), (4, 11), (7, 17), (2, 8), (5, 6)]:
        assert abs(O.noll_cov(i, j) - S.noll_cov(i, j)) < 1e-15


@pytest.mark.parametrize("nobj,nfr,K", [(1, 8, 20), (2, 24, 50), (3, 5, 7)])
def test_constraints(nobj, nfr, K):
    a = O.constraints([[nfr]] * nobj, [[list(range(nfr))]] * nobj, K)
    b = S.constraints([[nfr]] * nobj, [[list(range(nfr))]] * nobj, K)
    assert a[:2] == b[:2]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert np.abs(a[4] - b[4]).max() < 1e-14
