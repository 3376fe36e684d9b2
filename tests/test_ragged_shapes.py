"""Ragged patch shapes (host side, no device): the embedding of a patch's own null space
into the class geometry that mfbd_set_patch_shape receives."""
import numpy as np
import pytest

CFG = dict(s=64, n_frames=6, n_objects=2, modes=list(range(2, 26)), basis="kl", d_r0=10.0, mean_counts=1000.0,
           margin=5, settings=dict(FTOL=1e-3))


@pytest.fixture(scope="module")
def su():
    from redux_amd import synth
    try:
        return synth.build_setup(CFG)
    except RuntimeError as e:
        pytest.skip(str(e))


def test_full_shape_embeds_to_the_class_null_space(su):
    from redux_amd import synth
    big, sh, small = synth.ragged_patch(su, 5, 24, [6, 6])
    ns = su["nullspace"]
    assert sh["nFree"] == ns["nFree"] and np.all(sh["imgOn"] == 1)
    np.testing.assert_array_equal(sh["rows"], ns["rows"])
    np.testing.assert_array_equal(sh["cols"], ns["cols"])
    np.testing.assert_array_equal(sh["vals"], ns["vals"])
    np.testing.assert_array_equal(small["data"], big["data"])


def test_ragged_shape_maps_rows_onto_its_images_and_modes(su):
    from redux_amd import synth
    big, sh, small = synth.ragged_patch(su, 6, 13, [4, 2])
    K, m = big["K"], sh["nModes"]
    img, mode = sh["rows"] // K, sh["rows"] % K
    assert np.all(sh["imgOn"][img] == 1) and np.all(mode < m)
    assert list(np.flatnonzero(sh["imgOn"])) == [0, 1, 2, 3, 6, 7]
    np.testing.assert_array_equal(small["data"], big["data"][[0, 1, 2, 3, 6, 7]])
    np.testing.assert_array_equal(small["imgCh"], [0, 0, 0, 0, 1, 1])
    assert small["objects"][0]["modes"].shape[0] == 13
    # N has orthonormal columns on the patch's own parameters (as the class null space does)
    N = np.zeros((len(small["imgCh"]) * m, sh["nFree"]))
    np.add.at(N, (small["nullspace"]["rows"], small["nullspace"]["cols"]), small["nullspace"]["vals"])
    np.testing.assert_allclose(N.T @ N, np.eye(sh["nFree"]), atol=1e-12)


def test_shapes_outside_the_class_are_refused(su):
    from redux_amd import synth
    for m, f in [(0, [6, 6]), (25, [6, 6]), (10, [0, 6]), (10, [7, 6]), (10, [6])]:
        with pytest.raises(ValueError):
            synth.ragged_patch(su, 1, m, f)
