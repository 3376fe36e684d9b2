"""Evaluations and restorations larger than one chunk of the tile workspace.

MFBD_MAX_TILES (read when a context is created) caps the images per chunk, so a few small
patches run the multi-chunk paths: each chunk rewrites the lane's pinned tables, which the
previous chunk's queued uploads must have read first.  The results must equal those of a
context without the cap bit for bit, with per-kernel timing (one stream for every lane) and
without it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _patches(n, seed=21):
    from oracle import setup_np as S
    from redux_amd import synth
    su = synth.build_setup(dict(synth.CFG2_PATCH, n_frames=3), S)
    return [synth.build_patch(su, seed + k) for k in range(n)]


def _alphas(pbs, scale=2e-7, seed=5):
    rng = np.random.default_rng(seed)
    return np.stack([rng.normal(0, scale, (len(p["imgCh"]), p["K"])) for p in pbs])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("timing", [False, True])
def test_chunked_restore_and_metric_equal_unchunked(monkeypatch, prec, timing):
    from redux_amd.momfbd import Solver
    pbs = _patches(4)
    a = _alphas(pbs)
    idx = [0, 1, 2, 3]
    ref = Solver(pbs[0], precision=prec, max_batch=4)
    ref.load(pbs)
    ref.init()
    want_img = ref.restore(a, patches=idx)
    want_L = ref.metric(a, patches=idx)
    ref.close()
    nImg = len(pbs[0]["imgCh"])
    monkeypatch.setenv("MFBD_MAX_TILES", str(nImg))            # one patch per chunk
    if timing:
        monkeypatch.setenv("MFBD_KERNEL_TIMING", "1")
    print("chunked restore", prec, "timing" if timing else "", flush=True)
    sv = Solver(pbs[0], precision=prec, max_batch=4)
    sv.load(pbs)
    sv.init()
    got_img = sv.restore(a, patches=idx)
    got_L = sv.metric(a, patches=idx)
    # a second restoration of other patches in other order, after an evaluation
    got2 = sv.restore(a[[3, 1, 2, 0]], patches=[3, 1, 2, 0])
    sv.close()
    np.testing.assert_array_equal(got_img, want_img)
    np.testing.assert_array_equal(got_L, want_L)
    np.testing.assert_array_equal(got2, want_img[[3, 1, 2, 0]])


def test_restore_outputs_one_patch_per_chunk_equal_single_calls():
    """max_batch = 1: mfbd_restore_outputs of patches [0, 1, 0] runs three chunks, with the
    output offsets and per-chunk object rows of every chunk after the first."""
    from redux_amd.momfbd import Solver
    pbs = _patches(2, seed=31)
    a = _alphas(pbs, seed=9)
    sv = Solver(pbs[0], precision="fp64", max_batch=1)
    sv.load(pbs)
    sv.init()
    mask = 2 | 32 | 128
    print("restore outputs, one patch per chunk", flush=True)
    many = sv.restore_outputs(a[[0, 1, 0]], mask, patches=[0, 1, 0])
    one = [sv.restore_outputs(a[k], mask, patch=k) for k in (0, 1)]
    sv.close()
    for key in ("psf", "cobj", "res"):
        np.testing.assert_array_equal(many[key][0], one[0][key], err_msg=key)
        np.testing.assert_array_equal(many[key][1], one[1][key], err_msg=key)
        np.testing.assert_array_equal(many[key][2], one[0][key], err_msg=key)
