"""Per-patch exit statuses of the solve (solver.cpp:470-516), forced from the host:
- NONFINITE: frames holding NaN make the metric non-finite; with a single mode stage
  (MODE_START = K) the stage, and so the solve, ends there, and no re-cut follows;
- NO_PROGRESS: objects of weight 0 give a zero metric and an exactly zero gradient, so
  conjugate_rdx reports GSL_ENOPROG (conjugate_gradient.cpp:184-190);
- FAILED: NaN frames over ten mode stages reach a fourth failure before the last stage;
- the other patches of the same batch are unaffected (batched results equal solo ones)."""
import numpy as np
import pytest

from redux_amd import _lib as L
from tests.golden_problem import load

pytestmark = pytest.mark.gpu


def _solve(pbs, batch):
    from redux_amd.momfbd import Solver
    sv = Solver(pbs[0], precision="fp64", max_batch=batch)
    sv.load(pbs)
    out = sv.run()
    sv.close()
    return out


def _single_stage(pb):
    q = dict(pb)
    q["settings"] = dict(pb["settings"], MODE_START=pb["K"])
    return q


def test_nonfinite_and_no_progress_statuses():
    pb, _ = load()
    good = _single_stage(pb)
    nan = dict(good)
    nan["data"] = np.array(pb["data"], copy=True)
    nan["data"][2, 10:20, 10:20] = np.nan
    zero = dict(good)
    zero["objects"] = [dict(o, weight=0.0) for o in pb["objects"]]
    print("status: nan frames / zero weight / good, single stage", flush=True)
    r_nan = _solve([nan], 1)[0]
    assert r_nan["status"] == L.PATCH_NONFINITE and r_nan["failCount"] == 1, r_nan
    r_zero = _solve([zero], 1)[0]
    assert r_zero["status"] == L.PATCH_NO_PROGRESS and r_zero["finalMetric"] == 0.0, r_zero
    r_good = _solve([good], 1)[0]
    assert r_good["status"] == L.PATCH_CONVERGED
    # a batch holding all three: each patch as alone
    batch = _solve([nan, good, good], 3)
    assert batch[0]["status"] == L.PATCH_NONFINITE
    assert batch[1]["status"] == L.PATCH_CONVERGED
    assert np.array_equal(batch[1]["alpha"], r_good["alpha"]) and np.array_equal(batch[1]["img"], r_good["img"])


def test_repeated_nonfinite_stages_fail():
    """NaN frames through the mode stages without re-cuts: shifts are only re-cut when the
    tilt alphas ask for it, and non-finite alphas never do, so every
    stage fails once and the fourth failure ends the solve."""
    pb, _ = load()
    nan = dict(pb)
    nan["data"] = np.array(pb["data"], copy=True)
    nan["data"][:, 5, 5] = np.nan
    nan["settings"] = dict(pb["settings"], MODE_START=2, MODE_STEP=2)    # ten stages
    print("status: nan frames, ten mode stages", flush=True)
    r = _solve([nan], 1)[0]
    assert r["status"] == L.PATCH_FAILED and r["failCount"] == 4, r
    # four stages (the golden schedule): four failures, the last stage's non-finite exit
    nan["settings"] = dict(pb["settings"])
    r = _solve([nan], 1)[0]
    assert r["status"] == L.PATCH_NONFINITE and r["failCount"] == 4, r
