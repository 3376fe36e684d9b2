"""Mosaic of restored patches on the device (mfbd_mosaic) against the oracle's
Object::doMozaic restatement (src/lib/momfbd/object.cpp:1160-1207): bit-exact, since both
sum the patches in their given order with products and sums rounded separately and divide
in IEEE FP64."""
import numpy as np
import pytest

from oracle import oracle as O
from redux_amd import momfbd

pytestmark = pytest.mark.gpu


def grid_positions(field, s, n):
    p = np.round(np.linspace(0, field - s, n)).astype(np.int32)
    y, x = np.meshgrid(p, p, indexing="ij")
    return y.ravel(), x.ravel()


@pytest.mark.parametrize("clip", [True, False])
def test_mosaic_of_256_patches_of_88_matches_the_oracle(clip):
    """256 overlapping 88x88 patches cut at a 16x16 grid from a 1024x1024 field."""
    print("mosaic 256x88x88 clip", clip, flush=True)
    rng = np.random.default_rng(11)
    py, px = grid_positions(1024, 88, 16)
    p = rng.normal(1000.0, 50.0, (256, 88, 88)).astype(np.float32)
    got, gfirst = momfbd.mosaic(p, py, px, clip=clip)
    want, wfirst = O.mosaic(p, py, px, clip=clip)
    assert gfirst == wfirst and got.shape == want.shape
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("s,n", [(64, 1), (64, 7), (128, 5), (90, 33)])
def test_mosaic_at_scattered_positions_matches_the_oracle(s, n):
    """Ragged layouts: one patch, patches at random positions (overlapping or not), odd sizes."""
    print("mosaic scattered", s, n, flush=True)
    rng = np.random.default_rng(s * 100 + n)
    py = rng.integers(0, 300, n).astype(np.int32)
    px = rng.integers(0, 200, n).astype(np.int32)
    p = rng.uniform(-1.0, 3.0, (n, s, s)).astype(np.float32)
    got, gfirst = momfbd.mosaic(p, py, px)
    want, wfirst = O.mosaic(p, py, px)
    assert gfirst == wfirst
    np.testing.assert_array_equal(got, want)


def test_cut_and_mosaic_reproduces_the_field():
    """Patches cut from one field blend back to that field wherever a window covers it."""
    print("mosaic cut-and-reassemble 1024", flush=True)
    rng = np.random.default_rng(12)
    field = rng.uniform(500.0, 1500.0, (1024, 1024)).astype(np.float32)
    py, px = grid_positions(1024, 88, 16)
    p = np.stack([field[y:y + 88, x:x + 88] for y, x in zip(py, px)])
    img, (r0, c0) = momfbd.mosaic(p, py, px)
    m = 88 // 8
    assert (r0, c0) == (m, m) and img.shape == (1024 - 2 * m, 1024 - 2 * m)
    np.testing.assert_allclose(img, field[m:1024 - m, m:1024 - m], rtol=3e-7)
