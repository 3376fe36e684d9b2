"""Ragged batches: patches of different mode and frame counts solved in one context of the
largest counts (mfbd_set_patch_shape) against each patch solved alone in a context of its
own geometry, which is how the reference runs them (one job per geometry,
object.cpp:789-790).  A missing image has a zero OTF and a missing mode a zero alpha, every
reduction sums a patch's images in the same order, so the results are identical bit for
bit.  One ragged patch is also checked against the FP64 oracle's solve of its own geometry
(converged quantities, as for every full solve)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(s=64, n_frames=6, n_objects=2, modes=list(range(2, 26)), basis="kl", d_r0=10.0, mean_counts=1000.0,
           diversity=[0.0, 1.0], margin=5, settings=dict(FTOL=1e-3))
# (modes, frames per object): the full shape, mode counts off a multiple of 4, one object short
SHAPES = [(24, [6, 6]), (13, [4, 6]), (17, [6, 2]), (20, [3, 5]), (5, [2, 2])]


@pytest.fixture(scope="module")
def ragged():
    from redux_amd import synth
    su = synth.build_setup(CFG)
    return [synth.ragged_patch(su, 40 + k, m, f) for k, (m, f) in enumerate(SHAPES)]


def _alone(small, prec):
    from redux_amd.momfbd import Solver
    sv = Solver(small, precision=prec, max_batch=1)
    sv.load([small])
    out = sv.run()[0]
    sv.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ragged_batch_equals_each_patch_solved_alone(ragged, prec):
    from redux_amd.momfbd import Solver
    print("ragged batch", prec, flush=True)
    bigs = [b for b, _, _ in ragged]
    sv = Solver(bigs[0], precision=prec, max_batch=len(bigs))
    sv.load(bigs)
    sv.set_shapes([sh for _, sh, _ in ragged])
    got = sv.run()
    sv.close()
    K = bigs[0]["K"]
    for k, (big, sh, small) in enumerate(ragged):
        want = _alone(small, prec)
        g = got[k]
        sel = np.flatnonzero(sh["imgOn"])
        m = sh["nModes"]
        assert (g["status"], g["iterations"], g["nF"], g["nDF"]) == \
            (want["status"], want["iterations"], want["nF"], want["nDF"]), (k, prec)
        assert g["finalMetric"] == want["finalMetric"], (k, prec)
        np.testing.assert_array_equal(g["alpha"][sel][:, :m], want["alpha"], err_msg=f"patch {k}")
        rest = np.ones_like(g["alpha"], dtype=bool)
        rest[np.ix_(sel, np.arange(m))] = False
        assert np.all(g["alpha"][rest] == 0), k
        np.testing.assert_array_equal(g["img"], want["img"], err_msg=f"patch {k}")
        assert g["alpha"].shape == (len(big["imgCh"]), K)


def test_ragged_patch_against_the_oracle(ragged, oracle_lib):
    """Patch 1 (13 modes, 4 + 6 frames) of the ragged batch against the oracle's solve of
    that geometry: the same initial metric (1e-12), converged, a final metric no worse than
    the oracle's (1e-5) and within 1e-2, restored objects within 1e-2 (the CG trajectories
    of the two are not pinned)."""
    from redux_amd.momfbd import Solver
    print("ragged vs oracle", flush=True)
    bigs = [b for b, _, _ in ragged]
    sv = Solver(bigs[0], precision="fp64", max_batch=len(bigs))
    sv.load(bigs)
    sv.set_shapes([sh for _, sh, _ in ragged])
    got = sv.run()[1]
    sv.close()
    small = ragged[1][2]
    want = oracle_lib.solve(small)
    assert got["status"] == 0
    assert abs(got["initialMetric"] - want["initialMetric"]) <= 1e-12 * abs(want["initialMetric"])
    rel = (got["finalMetric"] - want["finalMetric"]) / abs(want["finalMetric"])
    assert -1e-2 < rel < 1e-5, rel
    err = np.linalg.norm(got["img"] - want["img"]) / np.linalg.norm(want["img"])
    assert err <= 1e-2, err


def test_patch_shape_validation(ragged):
    """Shapes outside the context are refused with MFBD_EINVAL before any device work."""
    from redux_amd.momfbd import Solver, SolverError
    bigs = [b for b, _, _ in ragged]
    sv = Solver(bigs[0], precision="fp64", max_batch=2)
    sv.load(bigs[:2])
    _, sh, _ = ragged[1]
    bad = [dict(sh, nModes=0), dict(sh, nModes=bigs[0]["K"] + 1),
           dict(sh, imgOn=np.r_[np.zeros(6, np.uint8), np.ones(6, np.uint8)]),   # object 0 without images
           dict(sh, rows=np.r_[sh["rows"][:-1], 13]),                            # a mode the patch lacks
           dict(sh, cols=np.r_[sh["cols"][:-1], sh["nFree"]])]
    for b in bad:
        with pytest.raises(SolverError, match=r"\(1\)"):
            sv.set_shapes([None, b])
    with pytest.raises(SolverError):
        sv.set_shapes([None, None, sh])
    sv.close()


def test_uneven_batch_over_three_grids_equals_each_patch_solved_alone():
    """An uneven batch over the grids 64, 128 and 256 (the last on the N = 256 transform
    path), ragged mode and frame counts within each, solved by mfbd_solve_group with one
    context per grid running at once: every patch as if solved alone in its own geometry."""
    from redux_amd import synth
    from redux_amd.momfbd import solve_group, uneven_solvers
    print("uneven batch, three grids", flush=True)
    small = dict(n_objects=2, basis="kl", d_r0=10.0, mean_counts=1000.0, diversity=[0.0, 1.0], margin=5,
                 settings=dict(FTOL=1e-3))
    classes = [dict(small, s=64, pupil_diameter=32.0, n_frames=4, modes=list(range(2, 14))),
               dict(small, s=88, pupil_diameter=64.0, n_frames=4, modes=list(range(2, 14))),
               dict(small, s=128, pupil_diameter=128.0, n_frames=3, modes=list(range(2, 10)))]
    groups = synth.uneven_batch(classes, 9, seed=5, n_modes=(5, 12), n_frames=(2, 4))
    assert sorted(k for g in groups for k in g["index"]) == list(range(9))
    assert [g["setup"]["N"] for g in groups] == [64, 128, 256]
    solvers = uneven_solvers([(g["patches"], g["shapes"]) for g in groups], precision="fp64")
    got = solve_group(solvers)
    for sv in solvers:
        sv.close()
    for g, res in zip(groups, got):
        for k, (sh, own) in enumerate(zip(g["shapes"], g["own"])):
            want = _alone(own, "fp64")
            sel = np.flatnonzero(sh["imgOn"])
            assert res[k]["finalMetric"] == want["finalMetric"], (g["setup"]["N"], k)
            assert res[k]["iterations"] == want["iterations"]
            np.testing.assert_array_equal(res[k]["alpha"][sel][:, :sh["nModes"]], want["alpha"])
            np.testing.assert_array_equal(res[k]["img"], want["img"])


def test_mixed_form_on_the_uneven_grids():
    """The mixed form (BF16 mode GEMM, FP32 transforms, FP64 reductions and line search) on
    an uneven batch over the grids 64 / 128 / 256, against the FP64 form of the same batch
    (itself pinned to the oracle): single evaluations at the FP64 solution within the
    mixed form's bounds at the other geometries (metric 5e-4, Vogel gradient 8e-2
    relative), every solve converged, restored objects within 5e-2 (the trajectories
    differ; measured on 96 config-3 patches: median 1.9e-3, max 2.8e-2)."""
    from redux_amd import synth
    from redux_amd.momfbd import solve_group, uneven_solvers
    print("uneven batch, mixed form", flush=True)
    small = dict(n_objects=2, basis="kl", d_r0=10.0, mean_counts=1000.0, diversity=[0.0, 1.0], margin=5,
                 settings=dict(FTOL=1e-3))
    classes = [dict(small, s=64, pupil_diameter=32.0, n_frames=4, modes=list(range(2, 14))),
               dict(small, s=88, pupil_diameter=64.0, n_frames=4, modes=list(range(2, 14))),
               dict(small, s=128, pupil_diameter=128.0, n_frames=3, modes=list(range(2, 10)))]
    groups = synth.uneven_batch(classes, 9, seed=6, n_modes=(5, 12), n_frames=(2, 4))
    sets = [(g["patches"], g["shapes"]) for g in groups]
    s64, smx = uneven_solvers(sets, precision="fp64"), uneven_solvers(sets, precision="mixed")
    r64, rmx = solve_group(s64), solve_group(smx)
    for sa, sb, ra, rb in zip(s64, smx, r64, rmx):
        a = np.stack([r["alpha"] for r in ra])
        idx = np.arange(len(ra), dtype=np.int32)
        sa.init()          # after the solves each context holds its own final shifts; back to
        sb.init()          # the first initialisation so both evaluate the same frames
        L64, Lmx = sa.metric(a, patches=idx), sb.metric(a, patches=idx)
        g64, gmx = sa.grad(a, patches=idx), sb.grad(a, patches=idx)
        assert np.all(np.abs(Lmx - L64) <= 5e-4 * np.abs(L64)), (Lmx, L64)
        for x, y in zip(gmx, g64):
            assert np.linalg.norm(x - y) <= 8e-2 * np.linalg.norm(y)
        for x, y in zip(rb, ra):
            assert x["status"] == 0
            err = np.linalg.norm(x["img"].astype(np.float64) - y["img"]) / np.linalg.norm(y["img"])
            assert err <= 5e-2, err
    for sv in s64 + smx:
        sv.close()
