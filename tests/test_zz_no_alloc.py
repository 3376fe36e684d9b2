"""mfbd_mosaic_noalloc (the allocation-free mosaic: staging in the library's fixed device
scratch array) through the C-ABI against mfbd_mosaic, bit for bit: two calls in a row, the
second smaller than the first, trimmed and untrimmed, and the size limit."""
import numpy as np
import pytest

from redux_amd import _lib as L
from redux_amd import momfbd

pytestmark = pytest.mark.gpu


def _noalloc(patches, py, px, clip):
    pt = np.ascontiguousarray(patches, dtype=np.float32)
    n, s = pt.shape[0], pt.shape[1]
    py = np.ascontiguousarray(py, dtype=np.int32)
    px = np.ascontiguousarray(px, dtype=np.int32)
    rows, cols = np.zeros(1, np.int32), np.zeros(1, np.int32)
    assert L.lib().mfbd_mosaic_size(n, s, L.ptr(py, L.ip), L.ptr(px, L.ip), L.ptr(rows, L.ip),
                                    L.ptr(cols, L.ip)) == 0
    img = np.full(int(rows[0]) * int(cols[0]), np.nan, dtype=np.float32)
    r0, c0 = np.zeros(1, np.int32), np.zeros(1, np.int32)
    rc = L.lib().mfbd_mosaic_noalloc(n, s, L.ptr(pt, L.fp), L.ptr(py, L.ip), L.ptr(px, L.ip), int(clip), 0,
                                     L.ptr(img, L.fp), L.ptr(rows, L.ip), L.ptr(cols, L.ip), L.ptr(r0, L.ip),
                                     L.ptr(c0, L.ip))
    if rc:
        return rc, None, None
    return 0, img[: int(rows[0]) * int(cols[0])].reshape(int(rows[0]), int(cols[0])), (int(r0[0]), int(c0[0]))


def _grid(field, s, n):
    p = np.round(np.linspace(0, field - s, n)).astype(np.int32)
    y, x = np.meshgrid(p, p, indexing="ij")
    return y.ravel(), x.ravel()


@pytest.mark.parametrize("clip", [True, False])
def test_noalloc_mosaic_equals_mosaic_twice(clip):
    rng = np.random.default_rng(77)
    # first call: 256 patches of 88 x 88 on a 1024 x 1024 field
    py, px = _grid(1024, 88, 16)
    p = rng.normal(1000.0, 50.0, (256, 88, 88)).astype(np.float32)
    print("noalloc mosaic 256x88 clip", clip, flush=True)
    rc, got, first = _noalloc(p, py, px, clip)
    assert rc == 0
    want, wfirst = momfbd.mosaic(p, py, px, clip=clip)
    assert first == wfirst
    np.testing.assert_array_equal(got, want)
    # second call, smaller: 9 patches of 64 x 64 at scattered positions; the array keeps the
    # first call's larger contents beyond what this call writes
    py2 = rng.integers(0, 150, 9).astype(np.int32)
    px2 = rng.integers(0, 120, 9).astype(np.int32)
    p2 = rng.uniform(-1.0, 3.0, (9, 64, 64)).astype(np.float32)
    print("noalloc mosaic 9x64 clip", clip, flush=True)
    rc, got2, first2 = _noalloc(p2, py2, px2, clip)
    assert rc == 0
    want2, wfirst2 = momfbd.mosaic(p2, py2, px2, clip=clip)
    assert first2 == wfirst2
    np.testing.assert_array_equal(got2, want2)


def test_noalloc_mosaic_rejects_what_does_not_fit():
    # 1024 patches of 128 x 128 need 64 MiB of staging: refused before any device work
    n, s = 1024, 128
    py, px = _grid(4096, s, 32)
    p = np.zeros((n, s, s), dtype=np.float32)
    rc, _, _ = _noalloc(p, py, px, True)
    assert rc == L.MFBD_ESCRATCH
    rc, _, _ = _noalloc(np.zeros((1, 4096, 4096), np.float32), np.zeros(1, np.int32), np.zeros(1, np.int32), True)
    assert rc == L.MFBD_ESCRATCH
