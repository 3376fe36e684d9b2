"""Config-5 geometry at full size per patch: 128 x 128 patch on the 256 x 256 grid,
128-pixel pupil, 150 KL modes, 2 objects x 100 frames.

- The FP64 metric and the reference's analytic (Vogel) gradient against the oracle
  (1e-12 and 1e-9 relative, as at the other geometries).
- The hand-written adjoint dL/dalpha in the FP32 form at 100 random alpha points against
  central finite differences (h = 1e-4 lambda) of the FP64 metric, three random
  components per point, and against the FP64 adjoint at every component: within 1e-3 of
  max|g| (the FP32 form's transforms).  The FP64 adjoint within 5e-6 of max|g|: the
  differences themselves carry ~1.3e-6 of rounding at this metric size (differences at h
  and 2h disagree by that much; measured 1.9e-6 worst case against the adjoint).
- A full solve of the FP32 adjoint form converges, and its restored objects are within
  1e-2 of the FP64 solve's (the two CG trajectories differ).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def patch():
    from oracle import setup_np
    from redux_amd import synth
    pb = synth.build_problem(synth.CFG5_PATCH, seed=9, setup=setup_np)
    assert (pb["N"], pb["K"], len(pb["imgCh"])) == (256, 150, 200)
    return pb


def _ctx(pb, prec, gradient="reference", batch=32):
    from redux_amd.momfbd import Solver
    sv = Solver(pb, precision=prec, max_batch=batch, gradient=gradient)
    sv.load([pb])
    sv.init()
    return sv


def test_metric_and_vogel_gradient_against_the_oracle(patch, oracle_lib):
    print("cfg5 metric / vogel vs oracle", flush=True)
    st = oracle_lib.State(patch)
    a = 0.5 * patch["truth"]["alpha"]
    sv = _ctx(patch, "fp64")
    L = sv.metric(a[None])[0]
    g = sv.grad(a[None], method="vogel")[0]
    sv.close()
    Lr, gr = st.metric(a), st.grad(a, method="vogel")
    assert abs(L - Lr) <= 1e-12 * abs(Lr), (L, Lr)
    assert np.linalg.norm(g - gr) / np.linalg.norm(gr) <= 1e-9


def test_adjoint_at_100_random_points(patch):
    print("cfg5 adjoint at 100 points", flush=True)
    rng = np.random.default_rng(21)
    at = patch["truth"]["alpha"]
    nImg, K = at.shape
    pts = np.stack([at * rng.uniform(0.0, 1.2) + 0.2 * at.std() * rng.standard_normal(at.shape)
                    for _ in range(100)])
    s32 = _ctx(patch, "fp32", gradient="adjoint")
    g32 = s32.grad(pts, method="adjoint", patches=np.zeros(100, dtype=np.int32))
    s32.close()
    s64 = _ctx(patch, "fp64")
    g64 = s64.grad(pts, method="adjoint", patches=np.zeros(100, dtype=np.int32))
    h = 1e-4 * patch["objects"][0]["wavelength"]
    comps = [(int(rng.integers(nImg)), int(rng.integers(K))) for _ in range(300)]
    probes = []
    for k in range(100):
        for c in range(3):
            i, m = comps[3 * k + c]
            for step in (h, -h, 2 * h, -2 * h):
                p = pts[k].copy()
                p[i, m] += step
                probes.append(p)
    Lp = s64.metric(np.stack(probes), patches=np.zeros(len(probes), dtype=np.int32))
    s64.close()
    fd = (Lp[0::4] - Lp[1::4]) / (2 * h)
    fd2 = (Lp[2::4] - Lp[3::4]) / (4 * h)
    # how far the differences themselves can be trusted: at this metric size (200 frames) the
    # rounding of L over 2h, not the h^2 term, sets their error (h and 2h differ by ~1e-6)
    noise = np.abs(fd - fd2).max() / np.abs(g64).max()
    print(f"central differences at h and 2h differ by {noise:.2e} of max|g|", flush=True)
    assert noise <= 2e-6
    worst32 = worst64 = 0.0
    for k in range(100):
        gmax = np.abs(g64[k]).max()
        for c in range(3):
            i, m = comps[3 * k + c]
            worst64 = max(worst64, abs(g64[k, i, m] - fd[3 * k + c]) / gmax)
            worst32 = max(worst32, abs(g32[k, i, m] - fd[3 * k + c]) / gmax)
        worst32 = max(worst32, np.abs(g32[k] - g64[k]).max() / gmax)
    print(f"adjoint vs central differences: fp64 {worst64:.2e}, fp32 {worst32:.2e} of max|g|", flush=True)
    assert worst64 <= 5e-6
    assert worst32 <= 1e-3


def test_fp32_adjoint_solve_converges_near_the_fp64_solve(patch):
    from redux_amd.momfbd import Solver
    print("cfg5 solves", flush=True)
    out = {}
    for prec, grad in (("fp64", "reference"), ("fp32", "adjoint")):
        sv = Solver(patch, precision=prec, max_batch=1, gradient=grad)
        sv.load([patch])
        out[prec] = sv.run()[0]
        sv.close()
    assert out["fp64"]["status"] == 0 and out["fp32"]["status"] == 0
    ref = out["fp64"]["img"].astype(np.float64)
    err = np.linalg.norm(out["fp32"]["img"] - ref) / np.linalg.norm(ref)
    print(f"fp32 adjoint solve vs fp64: object rel err {err:.2e}, metric {out['fp32']['finalMetric']:.6e} vs "
          f"{out['fp64']['finalMetric']:.6e}", flush=True)
    assert err <= 1e-2
