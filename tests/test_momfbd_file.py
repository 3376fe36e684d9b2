"""The .momfbd container (mfbd_write_momfbd) against the layout the reference's reader
defines: FileMomfbd::read and PatchInfo::parse / load (src/lib/file/filemomfbd.cpp:21-151,
280-390, 561-717), restated here as a reader that follows their version-dependent
branches and their per-patch size check.  The files are written from known arrays and
read back field by field; one file is also compared byte for byte with a packing of the
same fields in the order PatchInfo::write / FileMomfbd::write emit them (:391-476,
749-842).  The reference holds no .momfbd test files, so nothing else pins the format.
CPU only: the writer is host code in the library."""
import struct

import numpy as np
import pytest

from redux_amd import _lib as L
from redux_amd import momfbd


class Reader:
    """FileMomfbd::read + PatchInfo::parse / load, little- or big-endian by the flag byte."""

    def __init__(self, data):
        self.b, self.o = data, 0
        self.e = ">" if self.take("b") else "<"       # the endian flag: 1 = big-endian writer

    def take(self, fmt, n=1):
        e = getattr(self, "e", "<")
        size = struct.calcsize(e + fmt) * n
        v = struct.unpack_from(e + fmt * n, self.b, self.o)
        self.o += size
        return v[0] if n == 1 else np.array(v)

    def floats(self, n):
        a = np.frombuffer(self.b, dtype=np.dtype(np.float32).newbyteorder(self.e), count=n, offset=self.o)
        self.o += 4 * n
        return a.astype(np.float32)

    def string(self):
        n = self.take("i")
        raw = self.b[self.o:self.o + n]
        self.o += n
        assert raw[-1:] == b"\0"
        return raw[:-1].decode()

    def read(self):
        out = {"version": self.string()}
        v = float(out["version"])
        if v >= 20190408.0:
            out["modified"] = self.string()
        out["time"], out["date"] = self.string(), self.string()
        if v >= 20190401.17:
            out["region"] = self.take("i", 4)
        if self.take("b"):
            if v >= 20110714.0:
                out["pix2cf"], out["cf2pix"] = self.take("f"), self.take("f")
            nPH, nModes = self.take("i"), self.take("i")
            out["pupil"] = self.floats(nPH * nPH).reshape(nPH, nPH)
            modes = []
            for _ in range(nModes):
                if v >= 20120201.0:
                    assert self.take("B") == 1
                modes.append(self.floats(nPH * nPH).reshape(nPH, nPH))
            out["modes"] = np.array(modes)
        nCh = out["nCh"] = self.take("i")
        out["clip"] = np.stack([self.take("h", nCh) for _ in range(4)]).reshape(4, nCh)
        py, px, out["nPoints"] = self.take("i"), self.take("i"), self.take("i")
        out["patches"] = [[self.patch(v) for _ in range(px)] for _ in range(py)]
        if self.o < len(self.b):
            n = self.take("i")
            out["files"] = [self.string() for _ in range(n)]
        assert self.o == len(self.b)
        return out

    def patch(self, v):
        start = self.o
        p = {"region": self.take("i", 4)}
        if v >= 20110714.0:
            p["offx"], p["offy"] = self.take("i"), self.take("i")
        nCh = self.take("i")
        p["nim"], p["dx"], p["dy"] = self.take("i", nCh), self.take("i", nCh), self.take("i", nCh)
        nx = p["region"][1] - p["region"][0] + 1
        ny = p["region"][3] - p["region"][2] + 1
        size = 4 * 4 + (8 if v >= 20110714.0 else 0) + 4 + 12 * nCh + 1
        if self.take("b"):
            p["img"] = self.floats(nx * ny).reshape(ny, nx)
            size += 4 * nx * ny
        for key in ("psf", "obj", "res"):
            n = self.take("i")
            size += 4 + 4 * n * nx * ny
            p[key] = self.floats(n * nx * ny).reshape(n, ny, nx) if n else None
        nalpha = self.take("i")
        size += 4
        p["alpha"] = None
        if nalpha:
            nm = self.take("i")
            p["alpha"] = self.floats(nalpha * nm).reshape(nalpha, nm)
            size += 4 + 4 * nalpha * nm
        p["div"] = None
        if v >= 20100726.0:
            ndiv = self.take("i")
            size += 4
            if ndiv:
                nphx, nphy = self.take("i"), self.take("i")
                size += 8
                divs = []
                for _ in range(ndiv):
                    if v >= 20110916.0:
                        assert self.take("B") == 0
                        size += 1
                    divs.append(self.floats(nphx * nphy).reshape(nphy, nphx))
                    size += 4 * nphx * nphy
                p["div"] = np.array(divs)
        assert self.o - start == size            # PatchInfo::parse's size check
        return p


def _job(seed=0, py=2, px=3, s=12, nCh=2, nimg=(3, 2), nPH=8, nModes=4):
    rng = np.random.default_rng(seed)
    region = np.zeros((py, px, 4), np.int32)
    for y in range(py):
        for x in range(px):
            region[y, x] = [1 + 9 * x, 9 * x + s, 1 + 9 * y, 9 * y + s]
    nI = sum(nimg)
    return dict(img=rng.normal(size=(py, px, s, s)).astype(np.float32), region=region, nim=np.array(nimg),
                clip=np.array([[1, 2], [100, 120], [1, 3], [90, 95]], np.int16),
                dx=rng.integers(-3, 4, (py, px, nCh)), dy=rng.integers(-3, 4, (py, px, nCh)),
                psf=rng.random((py, px, nI, s, s)).astype(np.float32),
                cobj=rng.random((py, px, nI, s, s)).astype(np.float32),
                res=rng.normal(size=(py, px, nI, s, s)).astype(np.float32),
                alpha=rng.normal(size=(py, px, nI, 7)).astype(np.float32),
                pupil=rng.random((nPH, nPH)).astype(np.float32),
                modes=rng.normal(size=(nModes, nPH, nPH)).astype(np.float32),
                pix2cf=0.25, cf2pix=4.0, div=rng.normal(size=(nCh, nPH, nPH)).astype(np.float32),
                file_names=["a_00001.f0", "b_00002.f0", "c.fits"])


def test_full_container_reads_back(tmp_path):
    j = _job()
    path = tmp_path / "obj.momfbd"
    momfbd.write_momfbd(path, version="20240612.1", date="2024-06-12", time="08:30:00",
                        modified="2024-06-12T09:00:00", **j)
    f = Reader(path.read_bytes()).read()
    assert f["version"] == "20240612.1" and f["modified"] == "2024-06-12T09:00:00"
    assert f["time"] == "08:30:00" and f["date"] == "2024-06-12"
    assert list(f["region"]) == [1, 30, 1, 21]         # union of the patches' regions
    assert f["pix2cf"] == np.float32(0.25) and f["cf2pix"] == np.float32(4.0)
    np.testing.assert_array_equal(f["pupil"], j["pupil"])
    np.testing.assert_array_equal(f["modes"], j["modes"])
    assert f["nCh"] == 2
    np.testing.assert_array_equal(f["clip"], j["clip"])
    assert f["nPoints"] == 12 and len(f["patches"]) == 2 and len(f["patches"][0]) == 3
    for y in range(2):
        for x in range(3):
            p = f["patches"][y][x]
            np.testing.assert_array_equal(p["region"], j["region"][y, x])
            assert p["offx"] == 0 and p["offy"] == 0
            np.testing.assert_array_equal(p["nim"], j["nim"])
            np.testing.assert_array_equal(p["dx"], j["dx"][y, x])
            np.testing.assert_array_equal(p["dy"], j["dy"][y, x])
            np.testing.assert_array_equal(p["img"], j["img"][y, x])
            np.testing.assert_array_equal(p["psf"], j["psf"][y, x])
            np.testing.assert_array_equal(p["obj"], j["cobj"][y, x])
            np.testing.assert_array_equal(p["res"], j["res"][y, x])
            np.testing.assert_array_equal(p["alpha"], j["alpha"][y, x])
            np.testing.assert_array_equal(p["div"], j["div"])
    assert f["files"] == j["file_names"]


def test_image_only_container(tmp_path):
    """The reference always writes the image; every other block may be absent (count 0),
    and without names the file ends after the last patch."""
    j = _job(seed=3)
    path = tmp_path / "img.momfbd"
    momfbd.write_momfbd(path, img=j["img"], region=j["region"], nim=j["nim"], clip=j["clip"],
                        date="2024-01-02", modified="2024-01-02T03:04:05")
    f = Reader(path.read_bytes()).read()
    assert "pupil" not in f and "files" not in f
    p = f["patches"][1][2]
    assert p["psf"] is None and p["obj"] is None and p["res"] is None and p["alpha"] is None and p["div"] is None
    np.testing.assert_array_equal(p["img"], j["img"][1, 2])
    np.testing.assert_array_equal(p["dx"], [0, 0])


def test_bytes_match_the_writer_layout(tmp_path):
    """Byte for byte against the fields packed in the order FileMomfbd::write emits them."""
    j = _job(seed=5, py=1, px=2, s=4, nCh=1, nimg=(2,), nPH=2, nModes=1)
    j["clip"] = np.array([[1], [50], [1], [40]], np.int16)
    j["dx"], j["dy"] = j["dx"][..., :1], j["dy"][..., :1]
    j["div"] = j["div"][:1]
    path = tmp_path / "b.momfbd"
    momfbd.write_momfbd(path, version="20240612.1", date="D", time="T", modified="M", **j)

    def s(t):
        return struct.pack("<i", len(t) + 1) + t.encode() + b"\0"

    e = b"\0" + s("20240612.1") + s("M") + s("T") + s("D")
    e += struct.pack("<4i", 1, 13, 1, 4)
    e += b"\1" + struct.pack("<ff", 0.25, 4.0) + struct.pack("<ii", 2, 1) + j["pupil"].tobytes()
    e += b"\1" + j["modes"][0].tobytes()
    e += struct.pack("<i", 1) + j["clip"].astype("<i2").tobytes()
    e += struct.pack("<3i", 1, 2, 4)
    for x in range(2):
        e += j["region"][0, x].astype("<i4").tobytes() + struct.pack("<3i", 0, 0, 1)
        e += struct.pack("<3i", 2, j["dx"][0, x, 0], j["dy"][0, x, 0])
        e += b"\1" + j["img"][0, x].tobytes()
        for key in ("psf", "cobj", "res"):
            e += struct.pack("<i", 2) + j[key][0, x].tobytes()
        e += struct.pack("<ii", 2, 7) + j["alpha"][0, x].tobytes()
        e += struct.pack("<iii", 1, 2, 2) + b"\0" + j["div"][0].tobytes()
    e += struct.pack("<i", 3) + b"".join(s(n) for n in j["file_names"])
    assert path.read_bytes() == e


def test_current_time_and_argument_errors(tmp_path):
    j = _job(seed=7)
    path = tmp_path / "now.momfbd"
    momfbd.write_momfbd(path, img=j["img"], region=j["region"], nim=j["nim"], clip=j["clip"])
    f = Reader(path.read_bytes()).read()
    m = f["modified"]
    assert len(m) == 19 and m[4] == "-" and m[10] == "T" and m[13] == ":"
    with pytest.raises(momfbd.SolverError, match="20190408"):
        momfbd.write_momfbd(path, img=j["img"], region=j["region"], nim=j["nim"], clip=j["clip"],
                            version="20110714.0")
    bad = j["region"].copy()
    bad[0, 0, 1] += 1
    with pytest.raises(momfbd.SolverError, match="nPoints"):
        momfbd.write_momfbd(path, img=j["img"], region=bad, nim=j["nim"], clip=j["clip"])
    assert L.lib().mfbd_write_momfbd(str(path).encode(), None) == L.MFBD_EINVAL
