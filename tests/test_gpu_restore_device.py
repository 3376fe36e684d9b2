"""The restoration after the sums on the device (k_rs_build ... k_rs_crop: conj(P0) / Q0, the
Scharmer filter with its box smoothings, weights and connected region, the inverse
transform and the crop; Object::restorePatch, object.cpp:298-331, ScharmerFilter,
utils.cpp:334-392) against the oracle's restoration of the same patches at the same
alphas: four config-2 patches at the alphas of their FP64 solve, with the filter and with
NO_FILTER, in FP64 (relative RMS 1e-7, the objects' float rounding: the sums differ by
the transforms' rounding only, and the filter's thresholds and region follow the
oracle's) and, filtered, FP32 (1e-5; unfiltered, conj(P0) / Q0 divides by the small Q0
at the support's edge and carries the FP32 transforms' rounding up to ~0.3 there, the
reason for the filter); the objects a solve returns (restored while the solve runs) equal
mfbd_restore at the returned alphas."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def patches():
    from oracle import setup_np as S
    from redux_amd import synth
    _, pbs = synth.build_field_batch(synth.CFG2_PATCH, 256, setup=S, picks=[3, 70, 141, 222])
    return pbs


def _with(pbs, no_filter):
    out = []
    for p in pbs:
        q = dict(p)
        q["settings"] = dict(p["settings"], NO_FILTER=int(no_filter))
        out.append(q)
    return out


@pytest.mark.parametrize("no_filter", [False, True])
def test_restoration_against_the_oracle(patches, oracle_lib, no_filter):
    from redux_amd.momfbd import Solver
    pbs = _with(patches, no_filter)
    print("device restoration, no_filter", no_filter, flush=True)
    sv = Solver(pbs[0], precision="fp64", max_batch=len(pbs))
    sv.load(pbs)
    res = sv.run()
    alphas = np.stack([r["alpha"] for r in res])
    again = sv.restore(alphas)
    sv.close()
    for k, r in enumerate(res):
        np.testing.assert_array_equal(again[k], r["img"])         # solve's restoration == mfbd_restore
    for prec, tol in (("fp64", 1e-7), ("fp32", 1e-5))[: 1 if no_filter else 2]:
        sv = Solver(pbs[0], precision=prec, max_batch=len(pbs))
        sv.load(pbs)
        sv.init()
        imgs = sv.restore(alphas)
        sv.close()
        for k, p in enumerate(pbs):
            want = oracle_lib.State(p).restore(alphas[k])
            err = np.linalg.norm(imgs[k].astype(np.float64) - want) / np.linalg.norm(want)
            print(prec, "patch", k, f"{err:.2e}", flush=True)
            assert err <= tol, (prec, k, err)
