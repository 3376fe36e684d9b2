"""The conjugate-gradient line search on the device (mfbd_params.lineRounds): each lane's
patches step math::bracket + math::brent (helpers.cpp:58-216) in k_ls_step, one captured
HIP graph per round of probes (step, line phases, transform, P / Q, metric), the host
reading the lines' states every lineRounds rounds.  The state machine takes the host's
probes in the host's order (tests/test_line_search_machine.py pins it to the oracle on the
CPU) and each probe's metric comes from the same kernels on the same line phases, so a
solve with the device line search equals the same solve with the host's line search bit
for bit: alphas, objects, metric and evaluation counts, for every form with a fused
transform, for any number of rounds between read-backs, with the rounds replayed as graphs
or launched directly, and again on a second solve of the same context (graphs captured
anew).  Between graph replays the lanes' own eager
evaluations run on the same streams, which is what a solve does."""
import numpy as np
import pytest

from tests.golden_problem import load

pytestmark = pytest.mark.gpu


def _solve(pbs, prec, rounds, runs=1, graph=True, batch=None):
    from redux_amd.momfbd import Solver
    sv = Solver(pbs[0], precision=prec, max_batch=batch or len(pbs), line_phases=True, line_rounds=rounds,
                line_graph=graph)
    sv.load(pbs)
    out = [sv.run() for _ in range(runs)]
    st = sv.kernel_stats()
    sv.close()
    return out, st


def _same(a, b, tag):
    for k, (x, y) in enumerate(zip(a, b)):
        assert x["status"] == y["status"], (tag, k)
        assert np.array_equal(x["alpha"], y["alpha"]), (tag, k)
        assert np.array_equal(x["img"], y["img"]), (tag, k)
        assert x["finalMetric"] == y["finalMetric"], (tag, k)
        assert x["nF"] == y["nF"] and x["nDF"] == y["nDF"] and x["iterations"] == y["iterations"], (tag, k)


@pytest.mark.parametrize("prec", ["fp32", "mixed", "fp64"])
def test_golden_device_line_search_equals_host(prec):
    pb, _ = load()
    print("device line search, golden", prec, flush=True)
    (host,), st0 = _solve([pb], prec, 0)
    assert "host_line_search" not in st0
    for rounds, graph in ((1, True), (5, True), (3, False)):
        (dev,), st = _solve([pb], prec, rounds, graph=graph)
        assert st["host_line_search"]["launches"] > 0          # the lines ran on the device
        print(prec, "rounds", rounds, "graph", graph, "nF", dev[0]["nF"], "read-backs",
              st["host_line_search"]["launches"], flush=True)
        _same(dev, host, (prec, rounds, graph))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_config2_batch_device_line_search_equals_host(prec):
    """Eight config-2 patches over the evaluation lanes (four FP32, two FP64): every patch as
    with the host's line search, on two solves of one context."""
    from oracle import setup_np as S
    from redux_amd import synth
    _, pbs = synth.build_field_batch(synth.CFG2_PATCH, 256, setup=S, picks=[0, 9, 40, 77, 130, 161, 200, 255])
    print("device line search, config 2 batch", prec, flush=True)
    (host,), _ = _solve(pbs, prec, 0)
    (dev, again), st = _solve(pbs, prec, 6, runs=2)
    assert st["host_line_search"]["launches"] > 0
    assert all(r["status"] == 0 for r in host)
    _same(dev, host, prec)
    _same(again, host, prec)
    (eager,), _ = _solve(pbs, prec, 4, graph=False)
    _same(eager, host, (prec, "eager"))


def test_device_line_search_needs_line_phases():
    """Without line phases the host keeps the line search (the device's probes read the
    line's phases): lineRounds then changes nothing."""
    pb, _ = load()
    from redux_amd.momfbd import Solver
    out = {}
    for rounds in (0, 4):
        sv = Solver(pb, precision="fp32", max_batch=1, line_phases=False, line_rounds=rounds)
        sv.load([pb])
        out[rounds] = sv.run()
        assert "host_line_search" not in sv.kernel_stats()
        sv.close()
    _same(out[4], out[0], "no line phases")


def test_nonfinite_and_zero_weight_patches_as_on_the_host():
    """A patch whose frames hold NaN (its lines run on NaN metrics until brent's iteration
    limit ends them), beside good patches of the same batch, and a context whose objects
    weigh 0 (a zero gradient, so no line starts) end with the statuses and results the
    host's line search gives them, in FP32."""
    from redux_amd import _lib as L
    pb, _ = load()
    good = dict(pb)
    good["settings"] = dict(pb["settings"], MODE_START=pb["K"])
    nan = dict(good)
    nan["data"] = np.array(pb["data"], copy=True)
    nan["data"][2, 10:20, 10:20] = np.nan
    zero = dict(good)
    zero["objects"] = [dict(o, weight=0.0) for o in pb["objects"]]
    print("device line search: nan / good batch, zero weight alone", flush=True)
    for pbs, want in (([nan, good, good], [L.PATCH_NONFINITE, L.PATCH_CONVERGED, L.PATCH_CONVERGED]),
                      ([zero], [L.PATCH_NO_PROGRESS])):
        (host,), _ = _solve(pbs, "fp32", 0)
        (dev,), st = _solve(pbs, "fp32", 4, graph=False)
        assert [r["status"] for r in dev] == want, [r["status"] for r in dev]
        if len(pbs) > 1:
            assert st["host_line_search"]["launches"] > 0          # the good patches' lines
        for k, (x, y) in enumerate(zip(dev, host)):
            assert x["status"] == y["status"] and x["nF"] == y["nF"] and x["iterations"] == y["iterations"], k
            assert np.array_equal(x["alpha"], y["alpha"], equal_nan=True), k
            if x["status"] != L.PATCH_NONFINITE:
                assert np.array_equal(x["img"], y["img"]) and x["finalMetric"] == y["finalMetric"], k


def test_batches_smaller_than_the_load():
    """Five patches solved two at a time (three batches on one context: the lanes' line
    searches set up again for each) as with the host's line search."""
    pb, _ = load()
    rng = np.random.default_rng(23)
    pbs = []
    for k in range(5):
        p = dict(pb)
        p["data"] = rng.poisson(np.clip(pb["data"], 0, None)).astype(np.float32)
        pbs.append(p)
    print("device line search: 5 patches in batches of 2", flush=True)
    (host,), _ = _solve(pbs, "fp32", 0, batch=2)
    (dev,), st = _solve(pbs, "fp32", 4, graph=False, batch=2)
    assert st["host_line_search"]["launches"] > 0
    _same(dev, host, "batches of 2")
