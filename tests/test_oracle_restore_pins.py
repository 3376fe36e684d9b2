"""Pins for the oracle's metric, restored object and Scharmer filter that do not restate the
oracle's formulas:
- the metric (Solver::metric / Object::calcMetric, solver.cpp:763-817, object.cpp:597-615) is,
  per OTF-support point, the minimum over one complex object value F of the regularised
  least-squares misfit sum_j |D_j - F S_j|^2 + gamma |F|^2, computed here by numpy's lstsq;
- a single frame with gamma = 0 fits exactly, so its metric vanishes;
- the restored object without the filter (Object::restorePatch, object.cpp:278-331) is that
  least-squares F without gamma, and a single frame reproduces F S = D on the support;
- the filter (ScharmerFilter, utils.cpp:334-392) on inputs whose result is known in closed
  form: constant, thresholded, checkerboard (2 x 2 pre-smoothing), left/right halves (the
  u / -u symmetrisation), a ring of low signal (the region connected to DC)."""
import numpy as np
import pytest

from oracle import setup_np as S


def _support(pb, o=0):
    return S.pupil_support(pb["objects"][o]["pupil"])[2]


def _lstsq_residual(Dj, Sj, gamma):
    """min_F sum_j |D_j - F S_j|^2 + gamma |F|^2 at one point, and the minimiser."""
    A = np.concatenate([Sj, [np.sqrt(gamma)]])[:, None]
    b = np.concatenate([Dj, [0.0]])
    F, *_ = np.linalg.lstsq(A, b, rcond=None)
    r = b - A[:, 0] * F[0]
    return float(np.vdot(r, r).real), F[0]


def _one_frame(pb, frame=0, noise_fudge=None):
    """The patch with a single frame (metric and restoration only, so the null space is the
    identity)."""
    q = dict(pb)
    q["data"] = pb["data"][frame:frame + 1]
    q["imgCh"] = pb["imgCh"][frame:frame + 1]
    K = pb["K"]
    q["nullspace"] = dict(nPar=K, nFree=K, rows=np.arange(K), cols=np.arange(K), vals=np.ones(K))
    if noise_fudge is not None:
        q["channels"] = [dict(c, noiseFudge=noise_fudge) for c in pb["channels"]]
    return q


def test_metric_is_least_squares_minimum(oracle_lib, cfg1_problem):
    pb = cfg1_problem
    st = oracle_lib.State(pb)
    a = 0.5 * pb["truth"]["alpha"]
    L = st.metric(a)
    Sj = st.get("otf").reshape(st.nImg, -1)
    Dj = st.get("imgft").reshape(st.nImg, -1)
    gamma = st.scalar("gamma", 0)
    assert gamma > 0
    os_ = _support(pb)
    tot = 0.0
    for u in os_:
        r, _ = _lstsq_residual(Dj[:, u], Sj[:, u], gamma)
        tot += r
    w = pb["objects"][0]["weight"]
    Lref = w * tot / pb["s"] ** 2
    assert abs(L - Lref) <= 1e-10 * abs(Lref)


def test_metric_of_single_frame_without_regularisation_vanishes(oracle_lib, cfg1_problem):
    q = _one_frame(cfg1_problem, frame=2, noise_fudge=0.0)
    st = oracle_lib.State(q)
    assert st.scalar("gamma", 0) == 0.0
    a = 0.3 * cfg1_problem["truth"]["alpha"][2:3]
    L = st.metric(a)
    ftsum = st.get("ftsum")[0].ravel()[_support(cfg1_problem)].sum() / cfg1_problem["s"] ** 2
    assert ftsum > 0
    assert abs(L) <= 1e-12 * ftsum
    # with the regularisation of the noise estimate the single frame no longer fits exactly
    st2 = oracle_lib.State(_one_frame(cfg1_problem, frame=2))
    assert st2.metric(a) > 1e-6 * ftsum


def test_unfiltered_restoration_is_least_squares_object(oracle_lib, cfg1_problem):
    pb = dict(cfg1_problem)
    pb["settings"] = dict(pb["settings"], NO_FILTER=1)
    st = oracle_lib.State(pb)
    a = 0.5 * pb["truth"]["alpha"]
    full = st.restore_full(a, 0)
    Fhat = np.fft.fftshift(np.fft.fft2(full)).ravel()       # centred, DC at (N/2, N/2)
    Sj = st.get("otf").reshape(st.nImg, -1)
    Dj = st.get("imgft").reshape(st.nImg, -1)
    os_ = _support(pb)
    ref = np.array([_lstsq_residual(Dj[:, u], Sj[:, u], 0.0)[1] for u in os_])
    scale = np.abs(ref).max()
    assert np.abs(Fhat[os_] - ref).max() <= 1e-9 * scale
    mask = np.ones(Fhat.size, bool)
    mask[os_] = False
    assert np.abs(Fhat[mask]).max() <= 1e-9 * scale         # zero outside the OTF support


def test_single_frame_restoration_reproduces_data(oracle_lib, cfg1_problem):
    q = _one_frame(cfg1_problem, frame=5)
    q["settings"] = dict(q["settings"], NO_FILTER=1)
    st = oracle_lib.State(q)
    a = 0.7 * cfg1_problem["truth"]["alpha"][5:6]
    Fhat = np.fft.fftshift(np.fft.fft2(st.restore_full(a, 0))).ravel()
    Sj = st.get("otf").reshape(-1)
    Dj = st.get("imgft").reshape(-1)
    os_ = _support(cfg1_problem)
    live = np.abs(Sj[os_]) > 1e-6 * np.abs(Sj).max()
    u = os_[live]
    assert np.abs(Fhat[u] * Sj[u] - Dj[u]).max() <= 1e-8 * np.abs(Dj[u]).max()


def test_filtered_restoration_applies_filter_to_least_squares_object(oracle_lib, cfg1_problem):
    pb = cfg1_problem
    st = oracle_lib.State(pb)
    a = 0.5 * pb["truth"]["alpha"]
    full = st.restore_full(a, 0)
    Sj = st.get("otf").reshape(st.nImg, -1)
    Dj = st.get("imgft").reshape(st.nImg, -1)
    N = st.N
    os_ = _support(pb)
    Fls = np.zeros(N * N, complex)
    Fls[os_] = [_lstsq_residual(Dj[:, u], Sj[:, u], 0.0)[1] for u in os_]
    q2 = (np.abs(Sj) ** 2).sum(0)
    noise_var = np.mean([st.scalar("noise", i) ** 2 for i in range(st.nImg)])
    filt = oracle_lib.scharmer_filter(Fls.reshape(N, N), q2.reshape(N, N), noise_var,
                                      0.9 * pb["objects"][0]["freqCutoff"])
    # the restored object is the real part of the inverse transform (ift into a real array),
    # so its transform is the Hermitian part of the filtered F; the filter itself is not
    # exactly symmetric (its last smoothing window is half-open)
    idx = (N - np.arange(N)) % N
    herm = 0.5 * (filt + np.conj(filt[idx][:, idx]))
    Fhat = np.fft.fftshift(np.fft.fft2(full))
    assert np.abs(Fhat - herm).max() <= 1e-9 * np.abs(Fls).max()
    assert np.count_nonzero(np.abs(filt) > 0) < len(os_)     # the filter removes part of the support


# ---- the Scharmer filter on inputs with closed-form results ----

N = 64
H = N // 2


def _smooth8_of_delta(base, dc):
    """8 x 8 half-open box mean (smooth(rr, 4, 4), utils.cpp:305-328) of a constant `base`
    with the DC pixel set to `dc`: base + (dc - base) / count where the window holds DC."""
    out = np.empty((N, N))
    for y in range(N):
        y0, y1 = max(y - 4, 0), min(y + 4, N)
        for x in range(N):
            x0, x1 = max(x - 4, 0), min(x + 4, N)
            cnt = (y1 - y0) * (x1 - x0)
            inside = y0 <= H < y1 and x0 <= H < x1
            out[y, x] = base + ((dc - base) / cnt if inside else 0.0)
    return out


def _radius2():
    y, x = np.mgrid[0:N, 0:N]
    return (y - H) ** 2 + (x - H) ** 2


@pytest.mark.parametrize("k", [0.3, 0.6])
def test_filter_constant_signal(oracle_lib, k):
    # |data|^2 q2 = c everywhere: noise N^2 / c = k, filter 1 - k, DC set to 1, smoothed
    c, cut = 5.0, 24.0
    noise = k * c / N ** 2
    out = oracle_lib.scharmer_filter(np.ones((N, N)), np.full((N, N), c), noise, cut).real
    exp = _smooth8_of_delta(1.0 - k, 1.0)
    exp[_radius2() > cut ** 2] = 0.0
    assert np.abs(out - exp).max() < 1e-14
    assert out.min() >= 0.0 and out.max() <= 1.0


def test_filter_below_threshold_keeps_only_dc(oracle_lib):
    # 1 - k < 0.2 is cut to 0; DC (set to 1) is then the only connected point
    c, k = 2.0, 0.9
    out = oracle_lib.scharmer_filter(np.ones((N, N)), np.full((N, N), c), k * c / N ** 2, 30.0).real
    exp = _smooth8_of_delta(0.0, 1.0)
    exp[_radius2() > 30.0 ** 2] = 0.0
    assert np.abs(out - exp).max() < 1e-15
    assert abs(out[H, H] - 1.0 / 64) < 1e-15


def test_filter_presmooths_with_2x2_box(oracle_lib):
    # a checkerboard c1 / c2: the 2 x 2 half-open box (smooth(pt, 1, 1)) sees two of each,
    # so away from the border the filter is that of the constant (c1 + c2) / 2; a 3 x 3
    # box would leave an alternating filter
    c1, c2, cut = 2.0, 6.0, 24.0
    y, x = np.mgrid[0:N, 0:N]
    q2 = np.where((y + x) % 2 == 0, c1, c2)
    k = 0.3
    noise = k * 0.5 * (c1 + c2) / N ** 2
    out = oracle_lib.scharmer_filter(np.ones((N, N)), q2, noise, cut).real
    exp = _smooth8_of_delta(1.0 - k, 1.0)
    exp[_radius2() > cut ** 2] = 0.0
    assert np.abs(out - exp).max() < 1e-14


def test_filter_symmetrises_u_and_minus_u(oracle_lib):
    # signal c1 left of DC, c2 right of it: every u pairs with -u on the other side, so the
    # filter is 1 - (k1 + k2)/2 on both sides (away from the column through DC)
    c1, c2, cut = 4.0, 8.0, 26.0
    k1, k2 = 0.4, 0.2          # noise N^2 / c1 and noise N^2 / c2
    noise = k1 * c1 / N ** 2
    assert abs(noise * N ** 2 / c2 - k2) < 1e-15
    y, x = np.mgrid[0:N, 0:N]
    q2 = np.where(x < H, c1, c2).astype(float)
    out = oracle_lib.scharmer_filter(np.ones((N, N)), q2, noise, cut).real
    sel = (np.abs(x - H) >= 6) & (_radius2() <= cut ** 2)
    assert sel.sum() > 500
    assert np.abs(out[sel] - (1.0 - 0.5 * (k1 + k2))).max() < 1e-14


def test_filter_keeps_only_region_connected_to_dc(oracle_lib):
    # high signal everywhere except a ring of radius 24..32, where 1 - k < 0.2: the outer
    # disc is cut off from DC and removed.  Checked clear of the ring's edge and of the DC
    # pixel (the 2 x 2 and 8 x 8 boxes blur both by a few pixels).
    n, h, cut = 128, 64, 58.0
    y, x = np.mgrid[0:n, 0:n]
    r = np.sqrt((y - h) ** 2 + (x - h) ** 2)
    c_hi, c_lo, k_hi = 10.0, 1.0, 0.1
    noise = k_hi * c_hi / n ** 2                       # ring: k = noise n^2 / c_lo = 1.0
    q2 = np.where((r >= 24) & (r <= 32), c_lo, c_hi)
    out = oracle_lib.scharmer_filter(np.ones((n, n)), q2, noise, cut).real
    inner = (r >= 7.0) & (r <= 16.0)
    outer = (r >= 33.0) & (r <= cut)
    assert np.abs(out[inner] - (1.0 - k_hi)).max() < 1e-14
    assert np.abs(out[outer]).max() == 0.0
    # without the ring the outer disc keeps the filter value
    out2 = oracle_lib.scharmer_filter(np.ones((n, n)), np.full((n, n), c_hi), noise, cut).real
    assert np.abs(out2[outer] - (1.0 - k_hi)).max() < 1e-14
