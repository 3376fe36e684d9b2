"""HIP path against the FP64 oracle on one patch of the benchmark geometry:
s = 88 (not a power of two), N = 128 OTF grid with a 64-pixel pupil, two objects
(the second with a phase diversity) of 24 frames each, 50 KL modes.  This is the
shape the batched benchmark runs, so it exercises the N = 128 fused FP32 kernels
and the direct-DFT frame noise estimate.

Tolerances as for the golden problem: FP64 metric 1e-12 relative, gradients 1e-9;
FP32 metric 1e-6, gradients 1e-4 (relative 2-norm); frame noise and regularisation
gamma 1e-12 relative (FP64 in both precisions).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def patch():
    from redux_amd import synth
    from oracle import setup_np
    return synth.build_problem(synth.CFG2_PATCH, seed=11, setup=setup_np)


@pytest.fixture(scope="module")
def ref(patch, oracle_lib):
    st = oracle_lib.State(patch)
    at = patch["truth"]["alpha"]
    rng = np.random.default_rng(3)
    alphas = [np.zeros_like(at), 0.5 * at, at * rng.standard_normal(at.shape) * 0.3]
    return dict(state=st, alphas=alphas, metric=[st.metric(a) for a in alphas],
                vogel=[st.grad(a, method="vogel") for a in alphas])


def _solver(pb, prec):
    from redux_amd.momfbd import Solver
    sv = Solver(pb, precision=prec, max_batch=1)
    sv.load([pb])
    sv.init()
    return sv


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_frame_noise_and_gamma(patch, ref, prec):
    sv = _solver(patch, prec)
    st = ref["state"]
    nImg = len(patch["imgCh"])
    noise = sv.debug("noise_host")
    want = np.array([st.scalar("noise", i) for i in range(nImg)])
    assert np.abs(noise - want).max() <= 1e-12 * np.abs(want).max(), np.abs(noise - want).max()
    gam = sv.debug("gamma_host")
    for o in range(len(patch["objects"])):
        assert abs(gam[o] - st.scalar("gamma", o)) <= 1e-12 * abs(st.scalar("gamma", o))
    sv.close()


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
    """dL/dalpha from the adjoint kernels against central differences of the
    oracle's FP64 metric (h = 1e-4 lambda)."""
    sv = _solver(patch, prec)
    st = ref["state"]
    a = ref["alphas"][2]
    g = sv.grad(a[None], method="adjoint")[0]
    h = 1e-4 * patch["objects"][0]["wavelength"]
    for i, m in [(0, 0), (5, 1), (23, 9), (24, 4), (40, 30), (47, 49)]:
        ap, am = a.copy(), a.copy()
        ap[i, m] += h
        am[i, m] -= h
        gc = (st.metric(ap) - st.metric(am)) / (2 * h)
        assert abs(g[i, m] - gc) <= tol * np.abs(g).max(), (i, m, g[i, m], gc)
    sv.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("fp32", 1e-3)])
def test_restore(patch, ref, prec, tol):
    st = ref["state"]
    sv = _solver(patch, prec)
    for a in ref["alphas"][1:]:
        img = sv.restore(a[None])[0]
        want = st.restore(a)
        err = np.linalg.norm(img - want) / np.linalg.norm(want)
        assert err < tol, (prec, err)
    sv.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_batched_solve_equals_one_at_a_time(prec):
    """Patches of one batch step through the modes at their own pace, so a batched
    evaluation holds gradient requests with different enabled-mode counts; each must
    get its own (modes beyond its count have zero gradient, Solver::gradient).  The
    batched solve of four benchmark-geometry patches repeats the one-at-a-time solves
    bit for bit."""
    from redux_amd import synth
    from redux_amd.momfbd import Solver
    su, pbs = synth.build_field_batch(synth.CFG2_PATCH, 4, seed=7)
    sv = Solver(pbs[0], precision=prec, max_batch=len(pbs))
    sv.load(pbs)
    batch = sv.run()
    sv.close()
    for k, p in enumerate(pbs):
        sv = Solver(p, precision=prec, max_batch=1)
        sv.load([p])
        solo = sv.run()[0]
        sv.close()
        b = batch[k]
        assert solo["nF"] == b["nF"] and solo["nDF"] == b["nDF"], (prec, k, solo["nF"], b["nF"])
        assert np.array_equal(solo["alpha"], b["alpha"]), (prec, k)
        assert np.array_equal(solo["img"], b["img"]), (prec, k)


def test_solve_is_deterministic_across_runs_and_lane_counts():
    """The same context solved again, and with 1, 2 or 3 evaluation lanes (contiguous
    shares of the batch on streams of their own, overlapping on the device), gives the
    same results bit for bit: every patch sees its own sequence of requests, and the
    set-up on the synchronous lane completes before any lane evaluates."""
    import os
    from redux_amd import synth
    from redux_amd.momfbd import Solver
    print("determinism across runs and lanes", flush=True)
    _, pbs = synth.build_field_batch(synth.CFG2_PATCH, 48, seed=77)
    sv = Solver(pbs[0], precision="fp64", max_batch=len(pbs))
    sv.load(pbs)
    runs = []
    try:
        for lanes in ("2", "2", "1", "3"):
            os.environ["MFBD_EVAL_LANES"] = lanes
            runs.append(sv.run())
    finally:
        os.environ.pop("MFBD_EVAL_LANES", None)
        sv.close()
    for k, res in enumerate(runs[1:], 1):
        for a, b in zip(res, runs[0]):
            assert a["initialMetric"] == b["initialMetric"] and a["finalMetric"] == b["finalMetric"], k
            np.testing.assert_array_equal(a["alpha"], b["alpha"])
            np.testing.assert_array_equal(a["img"], b["img"])
