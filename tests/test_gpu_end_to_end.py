"""The product chain after preprocessing, end to end on one GPU, as a job runs it: patches
cut from one field, solved as one batch (Solver::run per patch), the optional restoration
outputs of a patch (Object::restorePatch), the mosaic of the restored patches
(Object::doMozaic) and the output files (Object::writeFits / writeAna / writeMomfbd).

Checked: every patch converges; each patch's restored object equals the restoration at
its returned alpha (mfbd_restore); the mosaic equals the oracle's mosaic of the same
patches bit for bit; the files read back to the mosaic (FITS float32, ANA int16 crunched), the .momfbd
container to the patches and their alphas;
the PSFs at the solution sum to the OTF's central value over the full grid less what falls
outside the crop."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_field_to_files(tmp_path):
    from oracle import oracle as O
    from redux_amd import momfbd, synth
    from tests.test_writers import decrunch, read_fits
    cfg, n, field = synth.CFG1, 4, 256
    su, pbs = synth.build_field_batch(cfg, n, seed=31, field_size=field)
    print("end to end: solve", n, "patches", flush=True)
    sv = momfbd.Solver(pbs[0], precision="fp64", max_batch=n)
    sv.load(pbs)
    res = sv.run()
    assert all(r["status"] == 0 for r in res), [r["status"] for r in res]
    # the restored objects are the restoration at the returned alphas
    again = sv.restore(np.stack([r["alpha"] for r in res]))
    for k, r in enumerate(res):
        np.testing.assert_array_equal(again[k], r["img"])
    # optional outputs of patch 1 at its solution
    out = sv.restore_outputs(res[1]["alpha"], 2 | 64 | 128, patch=1)
    s = pbs[0]["s"]
    assert out["psf"].shape == (1, s, s) and out["cobj"].shape == (len(pbs[0]["imgCh"]), s, s)
    assert np.all(np.isfinite(out["psf"])) and out["psf"].max() > 0
    assert 0.5 < out["psf"].sum() <= 1.0 + 1e-6           # a PSF normalised to unit volume, cropped
    sv.close()
    # patch origins in the field: the window's offset on the batch's grid, the frames' cut
    # centred in the window, the patch start within the frames (synth.build_patch)
    side = int(math.ceil(math.sqrt(n)))
    N, margin = su["N"], cfg.get("margin", 5)
    M = s + 2 * margin
    G = N if N >= M else int(2 ** math.ceil(math.log2(M)))
    stride = (field - G) // max(side - 1, 1)
    off = (G - M) // 2
    y0, x0 = pbs[0]["start"]
    py = np.array([(k // side) * stride + off + y0 for k in range(n)], dtype=np.int32)
    px = np.array([(k % side) * stride + off + x0 for k in range(n)], dtype=np.int32)
    imgs = np.stack([r["img"][0] for r in res])
    print("end to end: mosaic and files", flush=True)
    mos, first = momfbd.mosaic(imgs, py, px)
    want, wfirst = O.mosaic(imgs, py, px)
    assert first == wfirst
    np.testing.assert_array_equal(mos, want)
    momfbd.write_object(tmp_path / "obj", mos, fmt="fits")
    _, back, _ = read_fits(tmp_path / "obj.fits")
    np.testing.assert_array_equal(back, mos)
    momfbd.write_object(tmp_path / "obj", mos, fmt="ana", as_int16=True)
    np.testing.assert_array_equal(decrunch(open(tmp_path / "obj.f0", "rb").read()), mos.astype(np.int16))
    # the .momfbd container of the same patches (a 2 x 2 grid, y slow): restored objects,
    # alphas, modes and pupil, read back field by field
    from tests.test_momfbd_file import Reader
    ob = pbs[0]["objects"][0]
    region = np.array([[px[k] + 1, px[k] + s, py[k] + 1, py[k] + s] for k in range(n)], np.int32).reshape(2, 2, 4)
    alpha = np.stack([r["alpha"] for r in res]).reshape(2, 2, len(pbs[0]["imgCh"]), -1)
    momfbd.write_momfbd(tmp_path / "obj.momfbd", img=imgs.reshape(2, 2, s, s), region=region,
                        nim=[len(pbs[0]["imgCh"])], clip=np.array([[1], [field], [1], [field]]), alpha=alpha,
                        pupil=ob["pupil"], modes=ob["modes"] / ob["wavelength"], date="2024-06-12",
                        modified="2024-06-12T09:00:00")
    f = Reader(open(tmp_path / "obj.momfbd", "rb").read()).read()
    for k in range(n):
        p = f["patches"][k // 2][k % 2]
        np.testing.assert_array_equal(p["img"], imgs[k])
        np.testing.assert_array_equal(p["alpha"], res[k]["alpha"].astype(np.float32))
    np.testing.assert_array_equal(f["modes"], (ob["modes"] / ob["wavelength"]).astype(np.float32))
