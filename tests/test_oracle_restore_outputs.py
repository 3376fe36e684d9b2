"""The oracle's optional restoration outputs (object.cpp:332-407) pinned by the mathematics.

- PSF of an image: a direct (matrix) inverse DFT of the centred OTF, the image re-centred,
  PSF[m] = N^-2 sum_c OTF[c] exp(+2 pi i (c - N/2).(m - N/2) / N);
- averaged PSF: the mean of the per-image PSFs of the object;
- convolved object: the circular convolution sum_m obj[m] PSF0[(n - m) mod N] of the restored
  object with the PSF before re-centring (brute force at sample pixels);
- residual: convolved object minus the windowed frame ((x - mean) * window + mean of the
  cut-out), computed here from the frames.
A transposed DFT, a wrong sign of the exponent, a missed re-centring, a PSF scaled by N^2 or a
residual of the wrong image fails one of these."""
import numpy as np
import pytest

from tests.golden_problem import load


@pytest.fixture(scope="module")
def golden():
    return load()


def _dft_psf(otf_c):
    N = otf_c.shape[0]
    c = np.arange(N) - N // 2
    E = np.exp(2j * np.pi * np.outer(c, c) / N)          # E[c, m - N/2]
    return (E.T @ otf_c @ E).real / N**2


def test_psf_cobj_residual_against_direct_sums(golden, oracle_lib):
    pb, z = golden
    st = oracle_lib.State(pb)
    a = z["alphas"][2]
    out = st.restore_outputs(a, 2 | 32 | 128)
    otf = st.get("otf")
    N, s = st.N, st.s
    assert N == s
    obj = np.float32(st.restore(a)[0]).astype(np.float64)
    win = st.get("window")
    y0, x0 = pb["start"]
    scale = np.abs(out["cobj"]).max()
    for i in range(st.nImg):
        psf = _dft_psf(otf[i])
        assert np.abs(out["psf"][i] - psf).max() <= 1e-6 * np.abs(psf).max(), i
        psf0 = np.roll(psf, (-(N // 2), -(N // 2)), axis=(0, 1))        # PSF0[n] = PSF[n + N/2]
        for (ny, nx) in [(0, 0), (5, 61), (32, 32), (63, 7), (17, 40)]:
            yy = (ny - np.arange(N))[:, None] % N
            xx = (nx - np.arange(N))[None, :] % N
            ref = np.sum(obj * psf0[yy, xx])
            assert abs(out["cobj"][i, ny, nx] - ref) <= 1e-6 * scale, (i, ny, nx, out["cobj"][i, ny, nx], ref)
        cut = pb["data"][i, y0:y0 + s, x0:x0 + s].astype(np.float64)
        avg = cut.mean()
        w = (cut - avg) * win + avg
        assert np.abs(out["res"][i] - (out["cobj"][i] - w)).max() <= 1e-6 * scale, i


def test_psf_average_is_the_mean_and_flags_resolve_like_the_reference(golden, oracle_lib):
    pb, z = golden
    st = oracle_lib.State(pb)
    a = z["alphas"][1]
    per = st.restore_outputs(a, 32)["psf"]
    avg = st.restore_outputs(a, 64)["psf"]
    both = st.restore_outputs(a, 32 | 64)["psf"]            # GET_PSF_AVG wins (config.cpp:497-500)
    assert avg.shape == (1, st.s, st.s) and both.shape == avg.shape
    assert np.abs(avg[0] - per.mean(axis=0)).max() <= 1e-6 * np.abs(avg).max()
    assert np.array_equal(both, avg)
    # residual alone equals the residual saved with the convolved objects
    r1 = st.restore_outputs(a, 128)["res"]
    r2 = st.restore_outputs(a, 2 | 128)["res"]
    assert np.array_equal(r1, r2)


def test_cropped_outputs_on_a_padded_grid(oracle_lib):
    """s < N: the outputs are the centred s x s of the N x N arrays (object.cpp:393-408)."""
    from oracle import setup_np as S
    from redux_amd import synth
    pb = synth.build_problem(dict(synth.CFG2_PATCH, n_frames=2), seed=4, setup=S)
    st = oracle_lib.State(pb)
    a = np.zeros((st.nImg, st.K))
    out = st.restore_outputs(a, 32 | 2 | 128)
    N, s = st.N, st.s
    off = (N - s) // 2
    assert s < N
    otf = st.get("otf")
    for i in (0, st.nImg - 1):
        psf = _dft_psf(otf[i])[off:off + s, off:off + s]
        assert np.abs(out["psf"][i] - psf).max() <= 1e-6 * np.abs(psf).max(), i
        # windowed frame padded with its mean; the residual's crop is its centre
        y0, x0 = pb["start"]
        cut = pb["data"][i, y0:y0 + s, x0:x0 + s].astype(np.float64)
        avg = cut.mean()
        w = (cut - avg) * st.get("window") + avg
        scale = np.abs(out["cobj"][i]).max()
        assert np.abs(out["res"][i] - (out["cobj"][i] - w)).max() <= 1e-6 * scale, i
