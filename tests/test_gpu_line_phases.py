"""Line-search probes from the line's phases (mfbd_params.linePhases: base + step *
direction, Solver::my_precalc / metricAt, solver.cpp:246-327) instead of the mode GEMM per
probe.  The probe's phases equal the GEMM's in exact arithmetic and differ by rounding, so
a solve follows its own path: checked on converged quantities against the same solve with
the GEMM per probe and against the oracle, and bit for bit where it must hold by
construction (batched == one patch at a time, repeated solves)."""
import numpy as np
import pytest

from tests.golden_problem import load

pytestmark = pytest.mark.gpu


def _run(pbs, prec, batch, line):
    from redux_amd.momfbd import Solver
    sv = Solver(pbs[0], precision=prec, max_batch=batch, line_phases=line)
    sv.load(pbs)
    out = sv.run()
    sv.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_golden_solve_with_line_phases(prec):
    pb, z = load()
    print("line phases, golden", prec, flush=True)
    a = _run([pb], prec, 1, True)[0]
    b = _run([pb], prec, 1, False)[0]
    ref = z["solve_img"]
    ea = np.linalg.norm(a["img"] - ref) / np.linalg.norm(ref)
    eb = np.linalg.norm(b["img"] - ref) / np.linalg.norm(ref)
    fa = (a["finalMetric"] - z["solve_metrics"][-1]) / z["solve_metrics"][-1]
    print(prec, "line:", a["iterations"], a["nF"], f"{fa:+.2e}", f"{ea:.2e}", "gemm:", b["iterations"], b["nF"],
          f"{eb:.2e}", flush=True)
    assert a["status"] == 0 and a["initialMetric"] == b["initialMetric"]
    assert abs(fa) < 1e-3 and ea < 1e-2
    # the probes' phases carry the line's rounding only: the solve ends where the GEMM
    # per probe ends, to the trajectory spread of the form
    assert abs(a["finalMetric"] - b["finalMetric"]) <= 1e-3 * b["finalMetric"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_line_phases_batched_equals_one_at_a_time(prec):
    pb, _ = load()
    rng = np.random.default_rng(15)
    pbs = []
    for k in range(3):
        p = dict(pb)
        p["data"] = rng.poisson(np.clip(pb["data"], 0, None)).astype(np.float32)
        pbs.append(p)
    print("line phases, batch vs solo", prec, flush=True)
    batch = _run(pbs, prec, 3, True)
    again = _run(pbs, prec, 3, True)
    for k, p in enumerate(pbs):
        solo = _run([p], prec, 1, True)[0]
        for x in (batch[k], again[k]):
            assert np.array_equal(x["alpha"], solo["alpha"]) and np.array_equal(x["img"], solo["img"]), k
            assert x["nF"] == solo["nF"] and x["iterations"] == solo["iterations"]


def test_config2_patches_with_line_phases_against_the_oracle(oracle_lib):
    """Four config-2 patches (the benchmark layout), FP64 and FP32 with line phases, against
    the oracle's solves: FP64 as the FP64 form without them (most objects within 1e-4, all
    within 2e-3), FP32 within 5e-3."""
    from oracle import setup_np as S
    from redux_amd import synth
    _, pbs = synth.build_field_batch(synth.CFG2_PATCH, 256, setup=S, picks=[0, 64, 128, 192])
    print("line phases, config 2: oracle", flush=True)
    ores = oracle_lib.solve_many(pbs, threads=len(pbs))
    for prec, bound in (("fp64", 2e-3), ("fp32", 5e-3)):
        res = _run(pbs, prec, len(pbs), True)
        obj = np.array([np.linalg.norm(r["img"].astype(np.float64) - o["img"]) / np.linalg.norm(o["img"])
                        for r, o in zip(res, ores)])
        print(prec, "objects vs oracle", obj, flush=True)
        assert all(r["status"] == 0 for r in res)
        assert obj.max() <= bound
        if prec == "fp64":
            assert (obj <= 1e-4).sum() >= len(obj) / 2
