"""Configs 3 and 4 at their stated sizes: an uneven batch over the grids 64 / 128 / 256
(patches 64, 88, 128 with 32-, 64- and 128-pixel pupils), 36-100 KL modes and 12-48 frames
per object drawn per patch from the seed (synth.UNEVEN_CLASSES), two patches per grid here
(seed 5: N = 64 with 39 and 78 modes, 65 and 59 images; N = 128 with 69 and 62 modes, 57 and
29 images; N = 256 with 88 and 39 modes, 53 and 48 images).

- FP64 (config 3): the grids' contexts solved at once (mfbd_solve_group) give every patch
  what solving it alone in a context of its own geometry gives, bit for bit; one patch per
  grid evaluated against the oracle at alpha = 0 and at the solve's alphas (metric 5e-12,
  Vogel gradient 1e-9, relative).
- Mixed (config 4: BF16 mode GEMM, FP32 transforms, FP64 reductions and line search) and
  FP32 on the same batch: single evaluations at the FP64 solution within their bounds
  against FP64 (metric 5e-4 / 1e-6, Vogel gradient 8e-2 / 1e-4), every solve converged,
  objects within 5e-2 of the FP64 solve (the paths differ)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def groups():
    from redux_amd import synth
    print("config-3 batch: generating 6 patches", flush=True)
    g = synth.uneven_batch(synth.UNEVEN_CLASSES, 6, seed=5)
    assert [x["setup"]["N"] for x in g] == [64, 128, 256] and all(len(x["patches"]) == 2 for x in g)
    for x in g:
        for sh in x["shapes"]:
            assert 36 <= sh["nModes"] <= 100 and 24 <= int(sh["imgOn"].sum()) <= 96
    return g


@pytest.fixture(scope="module")
def fp64(groups):
    from redux_amd.momfbd import solve_group, uneven_solvers
    print("config-3 batch: fp64 solve_group", flush=True)
    solvers = uneven_solvers([(g["patches"], g["shapes"]) for g in groups], precision="fp64")
    res = solve_group(solvers)
    yield solvers, res
    for sv in solvers:
        sv.close()


def test_fp64_batch_equals_each_patch_alone(groups, fp64):
    from tests.test_gpu_ragged import _alone
    _, res = fp64
    for g, rs in zip(groups, res):
        for k, (sh, own) in enumerate(zip(g["shapes"], g["own"])):
            print("config-3 alone: N =", g["setup"]["N"], "patch", k, flush=True)
            want = _alone(own, "fp64")
            sel = np.flatnonzero(sh["imgOn"])
            assert rs[k]["status"] == 0 and want["status"] == 0
            assert rs[k]["finalMetric"] == want["finalMetric"] and rs[k]["iterations"] == want["iterations"]
            np.testing.assert_array_equal(rs[k]["alpha"][sel][:, :sh["nModes"]], want["alpha"])
            np.testing.assert_array_equal(rs[k]["img"], want["img"])


def test_fp64_evaluations_against_the_oracle(groups, fp64, oracle_lib):
    from redux_amd.momfbd import Solver
    _, res = fp64
    for g, rs in zip(groups, res):
        own, sh = g["own"][0], g["shapes"][0]
        sel = np.flatnonzero(sh["imgOn"])
        print("config-3 oracle: N =", g["setup"]["N"], flush=True)
        st = oracle_lib.State(own)
        sv = Solver(own, precision="fp64", max_batch=1)
        sv.load([own])
        sv.init()
        for a in (np.zeros((len(sel), sh["nModes"])), rs[0]["alpha"][sel][:, :sh["nModes"]]):
            L, Lo = sv.metric(a[None])[0], st.metric(a)
            assert abs(L - Lo) <= 5e-12 * abs(Lo), (g["setup"]["N"], L, Lo)
            gv, go = sv.grad(a[None], method="vogel")[0], st.grad(a, method="vogel")
            assert np.linalg.norm(gv - go) <= 1e-9 * np.linalg.norm(go), g["setup"]["N"]
        sv.close()


@pytest.mark.parametrize("prec,tol_m,tol_g", [("mixed", 5e-4, 8e-2), ("fp32", 1e-6, 1e-4)])
def test_reduced_precision_forms_on_the_batch(groups, fp64, prec, tol_m, tol_g):
    from redux_amd.momfbd import solve_group, uneven_solvers
    s64, r64 = fp64
    print("config-4 batch:", prec, flush=True)
    sets = [(g["patches"], g["shapes"]) for g in groups]
    sl = uneven_solvers(sets, precision=prec)
    rl = solve_group(sl)
    for sa, sb, ra, rb in zip(s64, sl, r64, rl):
        a = np.stack([r["alpha"] for r in ra])
        idx = np.arange(len(ra), dtype=np.int32)
        sa.init()
        sb.init()
        L64, Ll = sa.metric(a, patches=idx), sb.metric(a, patches=idx)
        g64, gl = sa.grad(a, patches=idx), sb.grad(a, patches=idx)
        assert np.all(np.abs(Ll - L64) <= tol_m * np.abs(L64)), (Ll, L64)
        for x, y in zip(gl, g64):
            assert np.linalg.norm(x - y) <= tol_g * np.linalg.norm(y)
        for x, y in zip(rb, ra):
            assert x["status"] == 0
            err = np.linalg.norm(x["img"].astype(np.float64) - y["img"]) / np.linalg.norm(y["img"])
            print(prec, "object vs fp64", f"{err:.2e}", flush=True)
            assert err <= 5e-2, err
    for sv in sl:
        sv.close()
