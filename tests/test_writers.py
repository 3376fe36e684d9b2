"""ANA and FITS writers (Object::writeAna / writeFits, src/lib/momfbd/object.cpp:1212-1334)
read back by independent parsers of the two formats: the ANA first block (magic
0x5555aaaa, nhb, datyp, ndim, dims at byte 192, include/redux/file/fileana.hpp:70-82) and
text blocks (src/lib/file/fileana.cpp:150-200); FITS 80-character cards in 2880-byte
records with big-endian data (the FITS standard, as fits_create_img / fits_write_img
write it).  Host code only: no device."""
import numpy as np
import pytest


def _lib_or_skip():
    from redux_amd import _lib
    try:
        _lib.lib()
    except RuntimeError as e:
        pytest.skip(str(e))


def read_ana(path):
    raw = open(path, "rb").read()
    magic, = np.frombuffer(raw[:4], "<u4")
    assert magic == 0x5555AAAA
    subf, source, nhb, datyp, ndim = raw[4], raw[5], raw[6], raw[7], raw[8]
    dims = np.frombuffer(raw[192:192 + 4 * ndim], "<u4")
    text_len = (nhb - 1) * 512 + 256
    text = raw[256:256 + text_len].split(b"\0")[0].decode()
    dt = ["<i1", "<i2", "<i4", "<f4", "<f8"][datyp]
    data = np.frombuffer(raw[256 + text_len:], dt).reshape(tuple(dims[::-1]))
    return dict(subf=subf, nhb=nhb, text=text, data=data)


def read_fits(path):
    raw = open(path, "rb").read()
    cards, pos = {}, 0
    while True:
        block = raw[pos:pos + 2880]
        pos += 2880
        done = False
        for i in range(36):
            c = block[80 * i:80 * i + 80].decode("ascii")
            key = c[:8].strip()
            if key == "END":
                done = True
                break
            if c[8:10] == "= ":
                cards[key] = c[10:].split("/")[0].strip()
        if done:
            break
    bitpix, naxis = int(cards["BITPIX"]), int(cards["NAXIS"])
    dims = [int(cards[f"NAXIS{i + 1}"]) for i in range(naxis)]
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    n = int(np.prod(dims))
    data = np.frombuffer(raw[pos:pos + n * np.dtype(dt).itemsize], dt).reshape(tuple(dims[::-1]))
    return cards, data, len(raw)


@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.float64, np.int32])
def test_fits_round_trip(tmp_path, dtype):
    _lib_or_skip()
    from redux_amd.momfbd import write_fits
    rng = np.random.default_rng(1)
    a = (rng.standard_normal((37, 53)) * 1000).astype(dtype)
    write_fits(tmp_path / "x.fits", a)
    cards, back, size = read_fits(tmp_path / "x.fits")
    assert cards["SIMPLE"] == "T" and cards["EXTEND"] == "T"
    assert (int(cards["NAXIS1"]), int(cards["NAXIS2"])) == (53, 37)
    assert size % 2880 == 0
    np.testing.assert_array_equal(back, a)


def test_ana_round_trip_and_text_blocks(tmp_path):
    _lib_or_skip()
    from redux_amd.momfbd import write_ana
    rng = np.random.default_rng(2)
    a = rng.standard_normal((20, 31)).astype(np.float32)
    write_ana(tmp_path / "a.f0", a, "TIME_OBS=12:00:00 NX=31 NY=20")
    r = read_ana(tmp_path / "a.f0")
    assert (r["subf"], r["nhb"], r["text"]) == (0, 1, "TIME_OBS=12:00:00 NX=31 NY=20")
    np.testing.assert_array_equal(r["data"], a)
    long = "X" * 700                      # more than 255 characters: 1 + (700 + 256) // 512 = 2 blocks
    b = rng.integers(-3000, 3000, (4, 5, 6)).astype(np.int16)
    write_ana(tmp_path / "b.f0", b, long)
    r = read_ana(tmp_path / "b.f0")
    assert r["nhb"] == 2 and r["text"] == long
    np.testing.assert_array_equal(r["data"], b)


def test_write_object_files(tmp_path):
    """The mosaic and the alphas under the reference's file names."""
    _lib_or_skip()
    from redux_amd.momfbd import write_object
    img = np.linspace(-5, 5, 30 * 40, dtype=np.float32).reshape(30, 40) * 100
    alpha = np.arange(2 * 3 * 4 * 5, dtype=np.float64).reshape(2, 3, 4, 5)
    write_object(tmp_path / "obj", img, fmt="fits", alpha=alpha)
    _, back, _ = read_fits(tmp_path / "obj.fits")
    np.testing.assert_array_equal(back, img)
    _, ab, _ = read_fits(tmp_path / "obj.alpha.fits")
    np.testing.assert_array_equal(ab, alpha.astype(np.float32))
    write_object(tmp_path / "obj", img, fmt="ana", as_int16=True, text="DATE=2026-01-01")
    r = read_ana(tmp_path / "obj.f0")
    np.testing.assert_array_equal(r["data"], img.astype(np.int16))
    assert r["text"] == "DATE=2026-01-01"


def test_bad_arguments(tmp_path):
    _lib_or_skip()
    import ctypes as C
    from redux_amd import _lib as L
    dims = (C.c_int * 2)(3, 0)
    buf = (C.c_float * 1)()
    assert L.lib().mfbd_write_fits(str(tmp_path / "z.fits").encode(), -32, 2, dims, buf) == L.MFBD_EINVAL
    assert L.lib().mfbd_write_ana(str(tmp_path / "z.f0").encode(), 9, 1, dims, buf, None) == L.MFBD_EINVAL
    assert b"datyp" in L.lib().mfbd_file_last_error()
