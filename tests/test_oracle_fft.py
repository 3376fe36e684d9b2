"""The oracle's Fourier helpers against the reference's own known-answer tests
(testsuite/image/fourier.cpp) and brute-force DFTs."""
import numpy as np
import pytest

TOL = 1e-7   # fourier.cpp:17


def brute_dft2(a, sign):
    ny, nx = a.shape
    y = np.arange(ny)
    x = np.arange(nx)
    Wy = np.exp(sign * 2j * np.pi * np.outer(y, y) / ny)
    Wx = np.exp(sign * 2j * np.pi * np.outer(x, x) / nx)
    return Wy @ a @ Wx


@pytest.mark.parametrize("n", [8, 16, 12, 88])
def test_fft_matches_brute_force(oracle_lib, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for sign in (-1, +1):
        ref = brute_dft2(a, sign)
        got = oracle_lib.fft2(a, sign)
        assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("n", [16, 64])
def test_constant_and_delta(oracle_lib, n):
    # fourier.cpp:236-285: constant -> delta of n^2 at (0,0); delta -> constant 1
    F = oracle_lib.fft2(np.ones((n, n)), -1)
    exp = np.zeros((n, n), complex)
    exp[0, 0] = n * n
    assert np.abs(F - exp).max() ** 2 < TOL
    d = np.zeros((n, n))
    d[0, 0] = 1
    assert np.abs(oracle_lib.fft2(d, -1) - 1).max() ** 2 < TOL
    # REORDER_FT puts DC at (n/2, n/2)
    Fr = oracle_lib.reorder_into(np.real(F))
    assert Fr[n // 2, n // 2] == n * n


def test_cosine_pair(oracle_lib):
    # fourier.cpp:300-318, nY = nX = 64: normalized, reordered transform
    n = 64
    y = np.arange(n)[:, None]
    x = np.arange(n)[None, :]
    a = np.cos(y * 2 * np.pi / n) + np.cos(x * 4 * np.pi / n)
    F = oracle_lib.fft2(a.astype(complex), -1) / (n * n)
    Fc = oracle_lib.reorder_into(np.real(F)) + 1j * oracle_lib.reorder_into(np.imag(F))
    exp = np.zeros((n, n))
    exp[n // 2 + 1, n // 2] = exp[n // 2 - 1, n // 2] = 0.5
    exp[n // 2, n // 2 + 2] = exp[n // 2, n // 2 - 2] = 0.5
    assert np.abs(Fc - exp).max() < 1e-12


def test_reorder_known_answers(oracle_lib):
    inp = np.arange(1, 101, dtype=float)
    r4 = oracle_lib.reorder_into(inp[:16].reshape(4, 4))
    assert r4.ravel().tolist() == [11, 12, 9, 10, 15, 16, 13, 14, 3, 4, 1, 2, 7, 8, 5, 6]
    assert oracle_lib.reorder_into(r4).ravel().tolist() == inp[:16].tolist()
    r5 = oracle_lib.reorder_into(inp[:25].reshape(5, 5))
    assert r5.ravel().tolist() == [19, 20, 16, 17, 18, 24, 25, 21, 22, 23, 4, 5, 1, 2, 3, 9, 10, 6, 7, 8,
                                   14, 15, 11, 12, 13]
    big = oracle_lib.reorder_into(inp[:16].reshape(4, 4), (5, 6))
    assert big.ravel().tolist() == [11, 12, 0, 0, 9, 10, 15, 16, 0, 0, 13, 14, 0, 0, 0, 0, 0, 0,
                                    3, 4, 0, 0, 1, 2, 7, 8, 0, 0, 5, 6]
    cen = oracle_lib.reorder_into(inp[:16].reshape(4, 4), (5, 6), center=True)
    assert cen.ravel().tolist() == [0, 11, 12, 9, 10, 0, 0, 15, 16, 13, 14, 0, 0, 3, 4, 1, 2, 0,
                                    0, 7, 8, 5, 6, 0, 0, 0, 0, 0, 0, 0]


def test_autocorrelation_known_answer(oracle_lib):
    # fourier.cpp:845-911: 2x2 block of 3s -> 36 / 18 / 9 at the wrap-around lags
    n = 64
    a = np.zeros((n, n))
    a[5, 5] = a[6, 6] = a[6, 5] = a[5, 6] = 3
    exp = np.zeros((n, n))
    exp[0, 0] = 36
    exp[0, 1] = exp[1, 0] = exp[n - 1, 0] = exp[0, n - 1] = 18
    exp[1, 1] = exp[n - 1, n - 1] = exp[n - 1, 1] = exp[1, n - 1] = 9
    assert np.abs(oracle_lib.autocorrelate(a) - exp).max() < TOL
    assert np.abs(oracle_lib.autocorrelate(a, center=True) - np.fft.fftshift(exp)).max() < TOL
    c = oracle_lib.autocorrelate(a.astype(complex), center=False)
    assert np.abs(c - exp).max() < TOL


def test_noise_estimate_white(oracle_lib):
    # white noise of sigma: noise() returns sigma for the mean power outside the mask
    rng = np.random.default_rng(0)
    s = 88
    a = rng.standard_normal((s, s)) * 3.0
    ft = np.fft.fft2(a)
    assert abs(oracle_lib.noise(ft) - 3.0) < 0.1
