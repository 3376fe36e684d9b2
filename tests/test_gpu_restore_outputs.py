"""Optional restoration outputs (mfbd_restore_outputs: the PSF, convolved-object and residual
blocks of Object::restorePatch, object.cpp:332-407) through the C-ABI against the oracle.

The outputs are float arrays. The device and the oracle differ in their transforms (device
FFT passes against the oracle's FFT; an OTF held on its support against a full
autocorrelation), by ~1e-15 relative in FP64. Tolerances, relative to the largest value of
the convolved objects (residuals are differences of such values) or of the PSFs: FP64 1e-6
(float rounding; measured: PSF 4e-14, convolved objects and residuals bit-identical); FP32
1e-5 (the FP32 OTF and restoration; measured 3e-7)."""
import numpy as np
import pytest

from tests.golden_problem import load

pytestmark = pytest.mark.gpu

ALL = 2 | 32 | 128


def _check(got, want, tol, tag):
    scale_c = np.abs(want["cobj"]).max()
    for key, scale in (("psf", np.abs(want["psf"]).max()), ("cobj", scale_c), ("res", scale_c)):
        err = np.abs(got[key] - want[key]).max() / scale
        print(tag, key, f"{err:.2e}", flush=True)
        assert err <= tol, (tag, key, err)


def test_golden_problem_against_the_oracle(oracle_lib):
    from redux_amd.momfbd import Solver
    pb, z = load()
    print("restore outputs: golden fp64", flush=True)
    sv = Solver(pb, precision="fp64", max_batch=1)
    sv.load([pb])
    sv.init()
    st = oracle_lib.State(pb)
    for k in (0, 2):
        a = z["alphas"][k]
        _check(sv.restore_outputs(a, ALL), st.restore_outputs(a, ALL), 1e-6, f"golden alpha {k}")
        avg = sv.restore_outputs(a, 64)["psf"]
        want = st.restore_outputs(a, 64)["psf"]
        assert np.abs(avg - want).max() <= 1e-6 * np.abs(want).max()
        both = sv.restore_outputs(a, 32 | 64)["psf"]          # GET_PSF_AVG wins
        assert np.array_equal(both, avg)
    sv.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("fp32", 1e-5)])
def test_padded_two_object_patch_against_the_oracle(oracle_lib, prec, tol):
    """88 x 88 patches on a 128 x 128 grid, two objects with diversity: the crop and the
    per-object averaging."""
    from oracle import setup_np as S
    from redux_amd import synth
    from redux_amd.momfbd import Solver
    pb = synth.build_problem(dict(synth.CFG2_PATCH, n_frames=3), seed=6, setup=S)
    print("restore outputs: cfg2-sized", prec, flush=True)
    rng = np.random.default_rng(3)
    a = rng.normal(0, 2e-7, (len(pb["imgCh"]), pb["K"]))
    sv = Solver(pb, precision=prec, max_batch=2)
    sv.load([pb, pb])
    sv.init()
    st = oracle_lib.State(pb)
    got = sv.restore_outputs(a, ALL, patch=1)
    _check(got, st.restore_outputs(a, ALL), tol, prec)
    avg = sv.restore_outputs(a, 64, patch=0)["psf"]
    want = st.restore_outputs(a, 64)["psf"]
    assert avg.shape == (2, pb["s"], pb["s"])
    assert np.abs(avg - want).max() <= tol * np.abs(want).max()
    # the same patch in either slot gives the same outputs, one at a time or in one call
    again = sv.restore_outputs(a, ALL, patch=0)
    both = sv.restore_outputs(np.stack([a, a]), ALL, patches=[0, 1])
    for key in ("psf", "cobj", "res"):
        assert np.array_equal(again[key], got[key]), key
        assert np.array_equal(both[key][0], got[key]) and np.array_equal(both[key][1], got[key]), key
    sv.close()


def test_after_a_ragged_solve_equal_each_patch_alone():
    """After the solve, the outputs at a patch's solution in a ragged batch equal those of the
    patch solved alone in its own geometry bit for bit; the images it lacks are zero."""
    from redux_amd import synth
    from redux_amd.momfbd import Solver
    from tests.test_gpu_ragged import CFG, SHAPES
    su = synth.build_setup(CFG)
    ragged = [synth.ragged_patch(su, 40 + k, m, f) for k, (m, f) in enumerate(SHAPES[:3])]
    print("restore outputs: ragged", flush=True)
    bigs = [b for b, _, _ in ragged]
    sv = Solver(bigs[0], precision="fp64", max_batch=len(bigs))
    sv.load(bigs)
    sv.set_shapes([sh for _, sh, _ in ragged])
    got = sv.run()
    for k in (1, 2):
        big, sh, small = ragged[k]
        out = sv.restore_outputs(got[k]["alpha"], ALL, patch=k)
        avg = sv.restore_outputs(got[k]["alpha"], 64, patch=k)["psf"]
        al = Solver(small, precision="fp64", max_batch=1)
        al.load([small])
        w = al.run()[0]
        want = al.restore_outputs(w["alpha"], ALL)
        want_avg = al.restore_outputs(w["alpha"], 64)["psf"]
        al.close()
        sel = np.flatnonzero(sh["imgOn"])
        off = np.setdiff1d(np.arange(len(big["imgCh"])), sel)
        for key in ("psf", "cobj", "res"):
            np.testing.assert_array_equal(out[key][sel], want[key], err_msg=f"patch {k} {key}")
            assert np.all(out[key][off] == 0), (k, key)
        np.testing.assert_array_equal(avg, want_avg, err_msg=f"patch {k} psf avg")
    # the three patches in one call, in another order, equal the calls one patch at a time
    order = [2, 0, 1]
    al3 = np.stack([got[k]["alpha"] for k in order])
    many, many_avg = sv.restore_outputs(al3, ALL, patches=order), sv.restore_outputs(al3, 64, patches=order)["psf"]
    for i, k in enumerate(order):
        one = sv.restore_outputs(got[k]["alpha"], ALL | 64, patch=k)["psf"]
        np.testing.assert_array_equal(many_avg[i], one, err_msg=f"patch {k} psf avg in one call")
        one = sv.restore_outputs(got[k]["alpha"], ALL, patch=k)
        for key in ("psf", "cobj", "res"):
            np.testing.assert_array_equal(many[key][i], one[key], err_msg=f"patch {k} {key} in one call")
    sv.close()


def test_argument_errors():
    from redux_amd import _lib as L
    from redux_amd.momfbd import Solver
    pb, z = load()
    sv = Solver(pb, precision="fp64", max_batch=1)
    sv.load([pb])
    a = np.ascontiguousarray(z["alphas"][0])
    out = np.zeros((len(pb["imgCh"]), pb["s"], pb["s"]), dtype=np.float32)
    lib = L.lib()
    one, bad = np.zeros(1, np.int32), np.ones(1, np.int32)
    assert lib.mfbd_restore_outputs(sv.h, 1, L.ptr(bad, L.ip), L.ptr(a), 32, L.ptr(out, L.fp), None, None) == \
        L.MFBD_EINVAL
    assert lib.mfbd_restore_outputs(sv.h, 1, L.ptr(one, L.ip), L.ptr(a), 2, L.ptr(out, L.fp), None, None) == \
        L.MFBD_EINVAL
    assert lib.mfbd_restore_outputs(sv.h, 1, L.ptr(one, L.ip), L.ptr(a), 1 | 4, None, None, None) == 0   # nothing asked
    sv.close()
