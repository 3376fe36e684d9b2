"""The oracle's mosaic (Object::doMozaic, src/lib/momfbd/object.cpp:1160-1207) against what
its definition fixes: the raised-cosine blend window, patches reproduced where one covers a
pixel alone, weighted means where several overlap, the untrimmed extent and the trim of empty
borders (include/redux/image/utils.hpp:389-553).  The reference has no tests of its own for
this path."""
import numpy as np
import pytest

from oracle import oracle as O


def raised_cosine(s):
    """apodizeInPlace(ones, blend, blend, margin, margin) in closed form: zero on the margin,
    0.5 - 0.5 cos(i pi / (blend + 1)) for the i-th blend pixel (i = 1..blend), one inside."""
    margin = s // 8
    blend = (s - 2 * margin) // 3
    f = np.ones(s)
    for r in range(s // 2):
        if r < margin:
            f[r] = 0.0
        elif r < margin + blend:
            f[r] = 0.5 - 0.5 * np.cos((r + 1 - margin) * np.pi / (blend + 1))
        f[s - 1 - r] = f[r]
    return np.outer(f, f), margin


@pytest.mark.parametrize("s", [16, 40, 64, 88])
def test_blend_window_is_the_separable_raised_cosine(s):
    want, margin = raised_cosine(s)
    got = O.blend_window(s, s, (s - 2 * margin) // 3, margin)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-15)  # cos(i*(pi/n)) vs cos(i*pi/n)
    assert np.all(got[:margin] == 0) and np.all(got[:, -margin:] == 0)
    assert got[s // 2, s // 2] == 1.0


def test_single_patch_is_reproduced_inside_its_margin():
    rng = np.random.default_rng(3)
    s, py, px = 32, 7, 11
    p = rng.uniform(0.5, 2.0, (1, s, s)).astype(np.float32)
    img, (r0, c0) = O.mosaic(p, [py], [px])
    m = s // 8
    assert (r0, c0) == (py + m, px + m)
    assert img.shape == (s - 2 * m, s - 2 * m)
    np.testing.assert_array_equal(img, p[0, m:s - m, m:s - m])


def test_untrimmed_extent_and_placement():
    rng = np.random.default_rng(4)
    s = 24
    p = rng.uniform(0.5, 2.0, (2, s, s)).astype(np.float32)
    img, first = O.mosaic(p, [0, 30], [5, 0], clip=False)
    assert img.shape == (30 + s + 1, 5 + s + 1) and first == (0, 0)
    m = s // 8
    np.testing.assert_array_equal(img[m:s - m, 5 + m:5 + s - m], p[0, m:s - m, m:s - m])
    np.testing.assert_array_equal(img[30 + m:30 + s - m, m:s - m], p[1, m:s - m, m:s - m])
    assert np.all(img[:m] == 0) and np.all(img[-1] == 0) and np.all(img[:, -1] == 0)


def test_constant_patches_give_a_constant_mosaic():
    s, step = 64, 40
    pos = [(y, x) for y in range(0, 3 * step, step) for x in range(0, 3 * step, step)]
    p = np.full((len(pos), s, s), 1.75, dtype=np.float32)
    img, _ = O.mosaic(p, [q[0] for q in pos], [q[1] for q in pos])
    np.testing.assert_allclose(img, 1.75, rtol=2e-7)


def test_overlap_is_the_weighted_mean_of_the_covering_patches():
    rng = np.random.default_rng(5)
    s = 48
    p = rng.uniform(0.0, 1.0, (2, s, s)).astype(np.float32)
    img, first = O.mosaic(p, [0, 0], [0, 20], clip=False)
    w, _ = raised_cosine(s)
    y, x = np.meshgrid(np.arange(s), np.arange(20 + s), indexing="ij")
    w0 = np.where(x < s, w[y, np.minimum(x, s - 1)], 0.0)
    w1 = np.where(x >= 20, w[y, np.clip(x - 20, 0, s - 1)], 0.0)
    v0 = np.where(x < s, p[0][y, np.minimum(x, s - 1)], 0.0)
    v1 = np.where(x >= 20, p[1][y, np.clip(x - 20, 0, s - 1)], 0.0)
    den = w0 + w1
    want = np.where(den > 1e-6, (w0 * v0 + w1 * v1) / np.where(den > 0, den, 1), w0 * v0 + w1 * v1)
    np.testing.assert_allclose(img[:s, :20 + s], want, rtol=1e-6, atol=1e-7)


def test_trim_keeps_an_interior_zero_band():
    """img_trim drops only border rows/columns: an all-zero column between two patches stays."""
    s = 16
    p = np.ones((2, s, s), dtype=np.float32)
    img, first = O.mosaic(p, [0, 0], [0, 2 * s])
    m = s // 8
    assert first == (m, m)
    assert img.shape == (s - 2 * m, 2 * s + s - 2 * m)
    assert np.all(img[:, s - 2 * m:2 * s] == 0)


def test_mosaic_size_without_a_device():
    import ctypes as C
    from redux_amd import _lib
    try:
        L = _lib.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    py = (C.c_int * 3)(0, 60, 120)
    px = (C.c_int * 3)(10, 0, 5)
    r, c = C.c_int(), C.c_int()
    assert L.mfbd_mosaic_size(3, 88, py, px, C.byref(r), C.byref(c)) == 0
    assert (r.value, c.value) == (120 + 88 + 1, 10 + 88 + 1)
    neg = (C.c_int * 1)(-1)
    assert L.mfbd_mosaic_size(1, 88, neg, px, C.byref(r), C.byref(c)) == _lib.MFBD_EINVAL
