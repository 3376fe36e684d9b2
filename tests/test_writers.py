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
    raw = open(tmp_path / "obj.f0", "rb").read()
    assert raw[4] & 1 and raw[512 + 12] == 5          # int16 output crunched at slice 5, as the reference
    np.testing.assert_array_equal(decrunch(raw), img.astype(np.int16))
    assert raw[256:512].split(b"\0")[0] == b"DATE=2026-01-01"


def test_bad_arguments(tmp_path):
    _lib_or_skip()
    import ctypes as C
    from redux_amd import _lib as L
    dims = (C.c_int * 2)(3, 0)
    buf = (C.c_float * 1)()
    assert L.lib().mfbd_write_fits(str(tmp_path / "z.fits").encode(), -32, 2, dims, buf) == L.MFBD_EINVAL
    assert L.lib().mfbd_write_ana(str(tmp_path / "z.f0").encode(), 9, 1, dims, buf, None) == L.MFBD_EINVAL
    assert b"datyp" in L.lib().mfbd_file_last_error()


# ---- against the reference's own test files (testsuite/testdata, copied by
# tests/golden/copy_file_fixtures.py): arrays with element (j, k) = j + k
import os

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "files")


def _gradient(rows, cols, dtype):
    j, k = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    return (j + k).astype(dtype)


@pytest.mark.parametrize("name,dtype", [("gradient_8u_4x5.f0", np.int8), ("gradient_16s_4x5_le.f0", np.int16),
                                        ("gradient_32s_4x5_le.f0", np.int32), ("gradient_32f_4x5_le.f0", np.float32),
                                        ("gradient_64f_4x5_le.f0", np.float64),
                                        ("gradient_32s_40x50_le.f0", np.int32)])
def test_ana_files_equal_the_reference_files(tmp_path, name, dtype):
    """Uncompressed ANA files written here are byte for byte the reference's."""
    _lib_or_skip()
    from redux_amd.momfbd import write_ana
    shape = (40, 50) if "40x50" in name else (4, 5)
    write_ana(tmp_path / "g.f0", _gradient(*shape, dtype))
    assert open(tmp_path / "g.f0", "rb").read() == open(os.path.join(FIX, name), "rb").read()


def test_ana_crunch_equals_the_reference_compressed_file(tmp_path):
    """The crunch code of the 40 x 50 int32 gradient at the fixture's slice (5) is the
    reference's compressed file byte for byte, and decodes back (independent decoder)."""
    _lib_or_skip()
    from redux_amd.momfbd import write_ana
    ref = open(os.path.join(FIX, "gradient_32s_40x50_le.fz"), "rb").read()
    assert ref[4] & 1 and ref[524] == 5
    a = _gradient(40, 50, np.int32)
    write_ana(tmp_path / "g.fz", a, slice=5)
    got = open(tmp_path / "g.fz", "rb").read()
    assert got == ref
    np.testing.assert_array_equal(decrunch(got), a)


def decrunch(raw):
    """An independent reader of ANA's crunch code (LSB-first bit stream per row)."""
    datyp, ndim = raw[7], raw[8]
    nx, ny = np.frombuffer(raw[192:200], "<u4")
    nhb = raw[6]
    off = 256 + (nhb - 1) * 512 + 256
    tsize, nblocks, bsize = np.frombuffer(raw[off:off + 12], "<u4")
    slice_, wide = raw[off + 12], raw[off + 13] == 4
    bits = np.unpackbits(np.frombuffer(raw[off + 14:], np.uint8), bitorder="little")
    pos, out = 0, np.zeros((ny, nx), dtype=np.int64)

    def take(n):
        nonlocal pos
        v = int(sum(int(b) << i for i, b in enumerate(bits[pos:pos + n])))
        pos += n
        return v

    first_bits = 32 if wide else 16
    for y in range(ny):
        v = take(first_bits)
        v = v - (1 << first_bits) if v >= 1 << (first_bits - 1) else v
        out[y, 0] = v
        for x in range(1, nx):
            low = take(slice_)
            z = 0
            while bits[pos] == 0:
                z += 1
                pos += 1
            pos += 1
            if z == 31:
                nb = 33 if wide else 17
                d = take(nb)
                d = d - (1 << nb) if d >= 1 << (nb - 1) else d
            else:
                r = z // 2 if z % 2 == 0 else -(z + 1) // 2
                d = (r << slice_) + low
            out[y, x] = out[y, x - 1] + d
        pos = (pos + 7) // 8 * 8
    return out


@pytest.mark.parametrize("dtype,slice_", [(np.int16, 1), (np.int16, 3), (np.int16, 7), (np.int16, 12),
                                          (np.int32, 2), (np.int32, 9), (np.int32, 20)])
def test_ana_crunch_round_trip(tmp_path, dtype, slice_):
    """Large and small differences (the escape of the unary code included) decode back."""
    _lib_or_skip()
    from redux_amd.momfbd import write_ana
    rng = np.random.default_rng(slice_)
    info = np.iinfo(dtype)
    a = np.cumsum(rng.integers(-40, 40, (30, 64)), axis=1)
    a[:, ::7] += rng.integers(-3000, 3000, a[:, ::7].shape)          # occasional big jumps
    a = np.clip(a, info.min // 2, info.max // 2).astype(dtype)
    write_ana(tmp_path / "r.fz", a, slice=slice_)
    raw = open(tmp_path / "r.fz", "rb").read()
    if raw[4] & 1:
        np.testing.assert_array_equal(decrunch(raw), a)
    else:                                      # not smaller compressed: stored raw
        assert raw[256 + 256:] == a.tobytes()


@pytest.mark.parametrize("name,dtype", [("gradient_16s_4x5.fits", np.int16), ("gradient_32s_4x5.fits", np.int32),
                                        ("gradient_32f_4x5.fits", np.float32), ("gradient_64f_4x5.fits", np.float64),
                                        ("gradient_32s_40x50.fits", np.int32)])
def test_fits_files_match_the_reference_files(tmp_path, name, dtype):
    """The same keywords and the same data records as the reference's FITS files."""
    _lib_or_skip()
    from redux_amd.momfbd import write_fits
    shape = (40, 50) if "40x50" in name else (4, 5)
    a = _gradient(*shape, dtype)
    write_fits(tmp_path / "g.fits", a)
    cg, dg, sg = read_fits(tmp_path / "g.fits")
    cr, dr, sr = read_fits(os.path.join(FIX, name))
    for key in ("SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2"):
        assert cg[key] == cr[key], key
    np.testing.assert_array_equal(dg, dr)
    np.testing.assert_array_equal(dg, a)
    assert sg == sr
