"""Pins for the oracle beyond the FFT known answers: closed forms, published
tables, invariants and central differences of the oracle's own metric."""
import math

import numpy as np
import pytest

from oracle import setup_np as S


def test_pupil_area_matches_reference_output():
    # testsuite/image/zernike.cpp:127 records "Area = 13198.1" for n=150, r=0.4321*150.
    # That comment predates the current makePupil (the test that printed it returns
    # at zernike.cpp:25); the current ramp gives 13199.05, within one pixel of it.
    p, area = S.make_pupil(150, 0.4321 * 150)
    assert abs(area - 13198.1) < 1.0
    assert abs(area - math.pi * (0.4321 * 150) ** 2) / area < 1e-4
    assert p.min() >= 0 and p.max() <= 1
    # utils.cpp:68-121 by hand: on the x axis (angle 0, d = 1) a pixel at distance D
    # from the centre with |D - r| <= 1 gets 0.5 - (D - r)/2
    r = 0.4321 * 150
    for D in (64, 65):
        assert abs(p[75, 75 + D] - (0.5 - (D - r) / 2)) < 1e-12
    assert p[75, 75 + 63] == 1.0 and p[75, 75 + 66] == 0.0


def test_pupil_point_symmetric():
    p, _ = S.make_pupil(64, 29.3)
    c = 32
    sub = p[1:, 1:]          # pixels (1..63) around centre 32 pair with their mirror
    assert np.abs(sub - sub[::-1, ::-1]).max() < 1e-12
    assert p[c, c] == 1.0


def test_noll_index_table():
    table = {1: (0, 0), 2: (1, 1), 3: (1, -1), 4: (2, 0), 5: (2, -2), 6: (2, 2), 7: (3, -1), 8: (3, 1),
             9: (3, -3), 10: (3, 3), 11: (4, 0), 22: (6, 0)}
    for j, nm in table.items():
        assert S.noll_to_nm(j) == nm


def test_radial_closed_forms():
    r = np.linspace(0, 1, 11).astype(np.longdouble)
    assert np.allclose(S.radial_poly(2, 0, r), 2 * r ** 2 - 1)
    assert np.allclose(S.radial_poly(3, 1, r), 3 * r ** 3 - 2 * r)
    assert np.allclose(S.radial_poly(4, 0, r), 6 * r ** 4 - 6 * r ** 2 + 1)
    assert np.allclose(S.radial_poly(4, 2, r), 4 * r ** 4 - 3 * r ** 2)


def test_zernike_orthonormal_on_pupil():
    P, rad = 64, 32.0
    pup = S.fetch_pupil(P, rad)
    Z = np.stack([S.zernike(j, P, rad) for j in range(2, 16)])
    G = np.einsum("iyx,jyx,yx->ij", Z, Z, pup["pupil"]) / pup["area"]
    assert np.abs(np.diag(G) - 1).max() < 1e-12           # numeric normalisation
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() < 0.05                        # discretised orthogonality (rim pixels)


def test_noll_covariance_matches_noll_1976():
    # Noll (1976) residual variances: Delta1 = 1.0299, Delta2 = 0.582, Delta3 = 0.134, Delta4 = 0.111
    assert abs(S.noll_cov(2, 2) - (1.0299 - 0.582)) < 2e-3
    assert abs(S.noll_cov(3, 3) - S.noll_cov(2, 2)) < 1e-15
    assert abs(S.noll_cov(4, 4) - (0.134 - 0.111)) < 1e-3
    assert S.noll_cov(4, 11) < 0             # defocus / spherical anticorrelated
    assert S.noll_cov(2, 3) == 0 and S.noll_cov(4, 5) == 0


def test_kl_diagonalises_covariance():
    kl = S.kl_expansion(2, 40)
    ids = list(range(2, 41))
    C = np.array([[S.noll_cov(i, j) for j in ids] for i in ids])
    W = np.zeros((len(ids), len(ids)))
    var = np.zeros(len(ids))
    for k, i in enumerate(ids):
        cov, weights = kl[i]
        var[k] = cov
        for zj, c in weights:
            W[ids.index(zj), k] = c
    assert np.abs(W.T @ W - np.eye(len(ids))).max() < 1e-10
    D = W.T @ C @ W
    assert np.abs(D - np.diag(var)).max() < 1e-12
    for k, i in enumerate(ids):              # sign rule: largest coefficient positive (zernike.cpp:451-460)
        col = W[:, k]
        assert col[np.argmax(np.abs(col))] > 0


def _constraint_matrix(nImgPerObj, K):
    nImg = sum(nImgPerObj)
    rows = []
    for m in range(K):
        if m < 2:
            r = np.zeros(nImg * K)
            r[[i * K + m for i in range(nImgPerObj[0])]] = 1
            rows.append(r)
        for o in range(1, len(nImgPerObj)):
            for f in range(nImgPerObj[o]):
                r = np.zeros(nImg * K)
                r[f * K + m] = 1
                r[(sum(nImgPerObj[:o]) + f) * K + m] = -1
                rows.append(r)
    return np.array(rows)


@pytest.mark.parametrize("nobj,nfr,K", [(1, 8, 20), (2, 6, 12), (2, 24, 10)])
def test_nullspace_identities(nobj, nfr, K):
    nPar, nFree, rows, cols, vals = S.constraints([[nfr]] * nobj, [[list(range(nfr))]] * nobj, K)
    Nmat = np.zeros((nPar, nFree))
    Nmat[rows, cols] = vals
    C = _constraint_matrix([nfr] * nobj, K)
    assert nFree == nPar - C.shape[0]
    assert np.abs(C @ Nmat).max() < 1e-12
    assert np.abs(Nmat.T @ Nmat - np.eye(nFree)).max() < 1e-12


def test_metric_invariants(oracle_lib, cfg1_problem):
    pb = cfg1_problem
    st = oracle_lib.State(pb)
    a = 0.5 * pb["truth"]["alpha"]
    L = st.metric(a)
    assert L > 0
    otf = st.get("otf")
    st0 = oracle_lib.State(pb)
    st0.metric(np.zeros_like(a))
    S0 = st0.get("otf")
    assert abs(np.abs(S0).max() - 1.0) < 1e-12     # otfNormalization (channel.cpp:941-952)
    assert np.abs(otf).max() <= 1.0 + 1e-12
    pb2 = dict(pb)
    pb2["data"] = pb["data"] * 3.0
    L3 = oracle_lib.State(pb2).metric(a)
    assert abs(L3 / L - 9.0) < 1e-9


def test_fd_gradient_matches_central_difference(oracle_lib, cfg1_problem):
    pb = cfg1_problem
    st = oracle_lib.State(pb)
    a = 0.5 * pb["truth"]["alpha"]
    g = st.grad(a, method="fd")
    h = 1e-4 * pb["objects"][0]["wavelength"]
    for i, m in [(0, 0), (1, 1), (3, 4), (5, 7)]:
        ap, am = a.copy(), a.copy()
        ap[i, m] += h
        am[i, m] -= h
        gc = (st.metric(ap) - st.metric(am)) / (2 * h)
        assert abs(g[i, m] - gc) < 0.02 * np.abs(g).max()


def test_vogel_gradient_interior_proportional(oracle_lib, cfg1_problem):
    """The reference's Vogel gradient weights each pupil pixel by an extra pupil
    factor, so it equals c * dL/dphi only where pupil == 1.  Checked per pixel."""
    pb = cfg1_problem
    st = oracle_lib.State(pb)
    a = 0.5 * pb["truth"]["alpha"]
    st.grad(a, method="vogel")
    vog = st.get("vogel")[0].ravel()
    phi = st.get("phi")[0].ravel()
    otf, ft = st.get("otf")[0].ravel(), st.get("imgft")[0].ravel()
    P, Q = st.get("P")[0].ravel(), st.get("Q")[0].ravel()
    pupil = pb["objects"][0]["pupil"]
    ps, pio, os_, area = S.pupil_support(pupil)
    Pn, N = pb["P"], pb["N"]
    G = np.zeros(N * N, complex)
    G[os_] = (P[os_] * Q[os_] * ft[os_] - np.abs(P[os_]) ** 2 * otf[os_]) / Q[os_] ** 2
    PF = pupil.ravel() * np.exp(1j * phi)
    ap = np.zeros((N, N), complex)
    py, px = ps // Pn, ps % Pn
    ap[(py - Pn // 2) % N, (px - Pn // 2) % N] = PF[ps]
    W = np.fft.ifft2(np.fft.fft2(ap) * np.fft.fft2(np.fft.ifftshift(G.reshape(N, N))))
    true = np.imag(np.conj(PF[ps]) * W[(py - Pn // 2) % N, (px - Pn // 2) % N])
    inner = pupil.ravel()[ps] == 1.0
    ratio = vog[ps][inner] / true[inner]
    assert np.abs(ratio / ratio.mean() - 1).max() < 1e-6
    assert abs(ratio.mean() * N * N - 1) < 1e-6
