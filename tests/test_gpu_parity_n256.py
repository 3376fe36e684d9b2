"""HIP path against the FP64 oracle on the 256 x 256 OTF grid of configs 3-5: a
128 x 128 patch with a 128-pixel pupil, 36 KL modes, two objects of 3 frames each
(the second with a phase diversity).  At N = 256 neither the FP32 tile (512 KB) nor
the FP64 tile fits in LDS, so this exercises the multi-pass FFT path: field, row and
column passes through HBM, power spectrum, support extraction, and the unfused
Vogel / adjoint gradient.

Tolerances as for the other geometries: FP64 metric 1e-12 relative and gradients
1e-9; FP32 metric 1e-6 and gradients 1e-4; adjoint against the oracle's central
differences (h = 1e-4 lambda) 1e-6 (FP64) and 1e-3 (FP32) of max|g|.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(s=128, n_frames=3, n_objects=2, modes=list(range(2, 38)), basis="kl", d_r0=20.0, mean_counts=2000.0,
           diversity=[0.0, math.pi / math.sqrt(3.0)], margin=5, pupil_diameter=128.0, settings=dict(FTOL=1e-3))


@pytest.fixture(scope="module")
def patch():
    from oracle import setup_np
    from redux_amd import synth
    pb = synth.build_problem(CFG, seed=5, setup=setup_np)
    assert pb["N"] == 256
    return pb


@pytest.fixture(scope="module")
def ref(patch, oracle_lib):
    st = oracle_lib.State(patch)
    at = patch["truth"]["alpha"]
    rng = np.random.default_rng(4)
    alphas = [np.zeros_like(at), 0.5 * at, at * rng.standard_normal(at.shape) * 0.3]
    return dict(state=st, alphas=alphas, metric=[st.metric(a) for a in alphas],
                vogel=[st.grad(a, method="vogel") for a in alphas])


def _solver(pb, prec):
    from redux_amd.momfbd import Solver
    sv = Solver(pb, precision=prec, max_batch=1)
    sv.load([pb])
    sv.init()
    return sv


@pytest.mark.parametrize("prec,tol_m,tol_g", [("fp64", 1e-12, 1e-9), ("fp32", 1e-6, 1e-4)])
def test_metric_and_vogel_gradient(patch, ref, prec, tol_m, tol_g):
    sv = _solver(patch, prec)
    for k, a in enumerate(ref["alphas"]):
        L = sv.metric(a[None])[0]
        assert abs(L - ref["metric"][k]) <= tol_m * abs(ref["metric"][k]), (prec, k, L, ref["metric"][k])
        g = sv.grad(a[None], method="vogel")[0]
        err = np.linalg.norm(g - ref["vogel"][k]) / np.linalg.norm(ref["vogel"][k])
        assert err <= tol_g, (prec, k, err)
    sv.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("fp32", 1e-3)])
def test_adjoint_gradient(patch, ref, prec, tol):
    sv = _solver(patch, prec)
    st = ref["state"]
    a = ref["alphas"][2]
    g = sv.grad(a[None], method="adjoint")[0]
    h = 1e-4 * patch["objects"][0]["wavelength"]
    for i, m in [(0, 0), (1, 1), (2, 9), (3, 4), (4, 30), (5, 35)]:
        ap, am = a.copy(), a.copy()
        ap[i, m] += h
        am[i, m] -= h
        gc = (st.metric(ap) - st.metric(am)) / (2 * h)
        assert abs(g[i, m] - gc) <= tol * np.abs(g).max(), (i, m, g[i, m], gc)
    sv.close()


def test_restore_fp64(patch, ref):
    st = ref["state"]
    sv = _solver(patch, "fp64")
    a = ref["alphas"][1]
    img = sv.restore(a[None])[0]
    want = st.restore(a)
    assert np.linalg.norm(img - want) / np.linalg.norm(want) < 1e-6
    sv.close()


def _with_lines(patch, prec, lines):
    import os
    from redux_amd.momfbd import Solver
    if lines:
        os.environ["MFBD_FFT_LINES"] = lines
    try:
        sv = Solver(patch, precision=prec, max_batch=1)
        sv.load([patch])
    finally:
        os.environ.pop("MFBD_FFT_LINES", None)
    sv.init()
    return sv


@pytest.mark.parametrize("prec,shapes", [("fp64", ["4,4", "4,16", "8,8", "16,4", "16,16"]),
                                         ("fp32", ["8,8", "8,32", "16,16", "32,8", "32,32"])])
def test_transform_tile_shapes_give_identical_results(patch, ref, prec, shapes):
    """The autotuner's candidates (lines per workgroup of the row and column passes) only
    redistribute the same butterflies: metric, Vogel and adjoint gradients and the
    restored objects are bit-identical for every shape, and the autotuned choice is one
    of them."""
    a = ref["alphas"][2]
    sv = _with_lines(patch, prec, None)
    cand = {"fp64": (4, 8, 16), "fp32": (8, 16, 32)}[prec]
    lr, lc = (int(v) for v in sv.debug("fft_lines"))
    assert lr in cand and lc in cand, (lr, lc)
    base = (sv.metric(a[None]), sv.grad(a[None], method="vogel"), sv.grad(a[None], method="adjoint"),
            sv.restore(a[None]))
    sv.close()
    for sh in shapes:
        print("lines", prec, sh, flush=True)
        sv = _with_lines(patch, prec, sh)
        assert [int(v) for v in sv.debug("fft_lines")] == [int(v) for v in sh.split(",")]
        got = (sv.metric(a[None]), sv.grad(a[None], method="vogel"), sv.grad(a[None], method="adjoint"),
               sv.restore(a[None]))
        sv.close()
        for x, y in zip(got, base):
            np.testing.assert_array_equal(x, y, err_msg=sh)


def test_transform_tile_shape_outside_the_candidates_is_refused(patch):
    from redux_amd.momfbd import SolverError
    with pytest.raises(SolverError, match="MFBD_FFT_LINES"):
        _with_lines(patch, "fp64", "12,16")


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("fp32", 1e-5)])
def test_restore_outputs_on_the_256_grid(patch, ref, prec, tol):
    """PSFs, convolved objects and residuals (mfbd_restore_outputs) through the multi-pass
    transforms of this grid, against the oracle; tolerances relative to the largest PSF /
    convolved-object value, as on the smaller grids."""
    sv = _solver(patch, prec)
    a = ref["alphas"][2]
    got = sv.restore_outputs(a, 2 | 32 | 128)
    want = ref["state"].restore_outputs(a, 2 | 32 | 128)
    sc = np.abs(want["cobj"]).max()
    for key, scale in (("psf", np.abs(want["psf"]).max()), ("cobj", sc), ("res", sc)):
        err = np.abs(got[key] - want[key]).max() / scale
        print(prec, key, f"{err:.2e}", flush=True)
        assert err <= tol, (prec, key, err)
    sv.close()
