"""Host-side mirror of the reference's per-patch solver interface.

`Solver` plays redux::momfbd::Solver (include/redux/momfbd/solver.hpp:86-179):
construct it once per job geometry, load patches, then `run()` solves them
(Solver::run, solver.cpp:330-580) and returns per-patch alphas, metrics and the
restored objects (Object::restorePatch, object.cpp:278-443).  The work happens in
libmomfbd_hip.so on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import Geometry, Params, Result, dp, fp, ip, ptr


class SolverError(RuntimeError):
    pass


def default_params(**kw):
    """GlobalCfg defaults (config.cpp:629-639), except GETSTEP: conjugate here (the
    reference's default is BFGS_inv, GSL vector_bfgs2; "bfgs_inv" selects it)."""
    p = dict(FTOL=1e-3, EPS=1e-10, GRADDIFF_STEP=1e-2, REG_ALPHA=0.0, MIN_ITER=5, MAX_ITER=500, N_DONE_ITER=3,
             MODE_START=5, MODE_STEP=5, APODIZATION=-1, NO_FILTER=0, GETSTEP="conjugate")
    p.update(kw)
    return p


# getstepMap (config.cpp:102-114), the steppers built here
GETSTEP = {"steepest": L.GETSTEP_STEEPEST, "steepest_descent": L.GETSTEP_STEEPEST,
           "getstep_steepest_descent": L.GETSTEP_STEEPEST,
           "conjugate": L.GETSTEP_CONJUGATE, "conjugate_gradient": L.GETSTEP_CONJUGATE,
           "getstep_conjugate_gradient": L.GETSTEP_CONJUGATE, "bfgs": L.GETSTEP_BFGS,
           "getstep_bfgs": L.GETSTEP_BFGS, "bfgs_inv": L.GETSTEP_BFGS_INV,
           "bfgsinv": L.GETSTEP_BFGS_INV, "getstep_bfgs_inv": L.GETSTEP_BFGS_INV}


class Solver:
    """One context: geometry + structure + basis + constraints; patches are loaded in bulk."""

    def __init__(self, pb: dict, precision: str = "fp64", max_batch: int = 1, device: int = 0, settings=None,
                 gradient: str = "reference", line_phases=None, line_rounds=None, line_graph=True):
        """line_phases: line-search probes from base + step * direction phases instead of the
        mode GEMM per probe (mfbd_params.linePhases); the fused transforms read the line in
        place of the GEMM's output.  By default on for the FP32 and mixed forms and off for
        FP64, whose default is the reference's arithmetic (the plain drop-in); FP64 with
        line phases and the device line search solves 256 config-2 patches at ~268 patches/s
        against 230, within 1e-4 of the oracle on the benchmark sample as the default.

        line_rounds > 0 (with line phases and the conjugate gradient): the line searches run
        on the device, the host reading their states every line_rounds rounds of probes
        (mfbd_params.lineRounds); the same probes and results as the host's line search.
        By default 4 where line phases are on (256 config-2 patches: FP32 ~18 % faster, FP64
        ~268 against 238 patches/s, the host no longer between probes).
        line_graph: each round of probes as one captured HIP graph (mfbd_params.lineGraph, the
        default), else its three kernels launched one by one: the same kernels and results, and
        the two measured the same (FP32 438 / 433-439 patches/s on 256 config-2 patches).  A
        context with kernel events on (MFBD_KERNEL_TIMING) launches them one by one, so that
        each kernel's time is its own; rocprofv3's kernel trace cannot follow graph launches."""
        self.pb = pb
        st = default_params(**pb.get("settings", {}))
        if settings:
            st.update(settings)
        self.settings = st
        objs, chans = pb["objects"], pb["channels"]
        self.nObj, self.nCh, self.nImg = len(objs), len(chans), len(pb["imgCh"])
        self.K, self.s, self.P = pb["K"], pb["s"], pb["P"]
        data = pb["data"]
        g = Geometry(self.s, self.P, self.K, self.nObj, self.nCh, self.nImg, data.shape[-2], data.shape[-1],
                     pb["start"][0], pb["start"][1])
        lp = int(precision != "fp64" if line_phases is None else bool(line_phases))
        prm = Params(st["FTOL"], st["EPS"], st["GRADDIFF_STEP"], st["REG_ALPHA"], st["MIN_ITER"], st["MAX_ITER"],
                     st["N_DONE_ITER"], st["MODE_START"], st["MODE_STEP"], st["APODIZATION"], int(st["NO_FILTER"]),
                     {"fp64": L.FP64, "fp32": L.FP32, "mixed": L.MIXED}[precision], max_batch,
                     L.GRAD_ADJOINT if gradient == "adjoint" else L.GRAD_SCHEDULE_REFERENCE,
                     GETSTEP[str(st["GETSTEP"]).lower()],
                     lp, int((4 if lp else 0) if line_rounds is None else line_rounds),
                     int(bool(line_graph)))
        h = C.c_void_p()
        self._check(L.lib().mfbd_ctx_create(C.byref(g), C.byref(prm), device, C.byref(h)), "ctx_create", None)
        self.h = h
        chObj = np.array([c["obj"] for c in chans], dtype=np.int32)
        imgCh = np.ascontiguousarray(pb["imgCh"], dtype=np.int32)
        w = np.array([o["weight"] for o in objs], dtype=np.float64)
        wl = np.array([o["wavelength"] for o in objs], dtype=np.float64)
        fc = np.array([o["freqCutoff"] for o in objs], dtype=np.float64)
        self._check(L.lib().mfbd_set_structure(h, ptr(chObj, ip), ptr(imgCh, ip), ptr(w), ptr(wl), ptr(fc)),
                    "set_structure")
        pupil = np.ascontiguousarray(objs[0]["pupil"], dtype=np.float64)
        modes = np.ascontiguousarray(np.stack([o["modes"] for o in objs]), dtype=np.float64)
        tilt = np.ascontiguousarray([o["tilt"] for o in objs], dtype=np.int32)
        jac = np.ascontiguousarray([np.concatenate([np.ravel(o["J"]), np.ravel(o["Ji"])]) for o in objs],
                                   dtype=np.float64)
        self._check(L.lib().mfbd_set_basis(h, ptr(pupil), ptr(modes), ptr(tilt, ip), ptr(jac)), "set_basis")
        ns = pb["nullspace"]
        rows = np.ascontiguousarray(ns["rows"], dtype=np.int32)
        cols = np.ascontiguousarray(ns["cols"], dtype=np.int32)
        vals = np.ascontiguousarray(ns["vals"], dtype=np.float64)
        self.nPar, self.nFree = ns["nPar"], ns["nFree"]
        self._check(L.lib().mfbd_set_constraints(h, self.nPar, self.nFree, len(rows), ptr(rows, ip), ptr(cols, ip),
                                                 ptr(vals)), "set_constraints")
        self.nPatch = 0

    def _check(self, rc, what, h="self"):
        if rc != 0:
            msg = ""
            hh = self.h if h == "self" else None
            if hh is not None:
                msg = L.lib().mfbd_last_error(hh).decode()
            raise SolverError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            L.lib().mfbd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def load(self, patches):
        """patches: list of problem dicts sharing this geometry (their data/resOffset differ)."""
        data = np.ascontiguousarray(np.stack([p["data"] for p in patches]), dtype=np.float32)
        chans = self.pb["channels"]
        phiF = np.ascontiguousarray(np.stack([c["phiFixed"] for c in chans]), dtype=np.float64)
        res = np.ascontiguousarray([[c["resOffset"] for c in p["channels"]] for p in patches], dtype=np.float64)
        nf = np.ascontiguousarray([c["noiseFudge"] for c in chans], dtype=np.float64)
        self._check(L.lib().mfbd_load_patches(self.h, len(patches), ptr(data, fp), ptr(phiF), ptr(res), ptr(nf)),
                    "load_patches")
        self.nPatch = len(patches)
        self._data = data

    def set_shapes(self, shapes):
        """Ragged batch: per loaded patch, a dict(nModes, imgOn, nFree, rows, cols, vals) (or
        None to keep the context's) -- the modes and images it has and its null space."""
        for p, sh in enumerate(shapes):
            if sh is None:
                continue
            on = np.ascontiguousarray(sh["imgOn"], dtype=np.uint8)
            rows = np.ascontiguousarray(sh["rows"], dtype=np.int32)
            cols = np.ascontiguousarray(sh["cols"], dtype=np.int32)
            vals = np.ascontiguousarray(sh["vals"], dtype=np.float64)
            self._check(L.lib().mfbd_set_patch_shape(self.h, p, int(sh["nModes"]), on.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                                     int(sh["nFree"]), len(rows), ptr(rows, ip), ptr(cols, ip),
                                                     ptr(vals)), "set_patch_shape")

    def run(self, alpha0=None, restore=True):
        n = self.nPatch
        res = (Result * n)()
        alpha = np.zeros((n, self.nImg, self.K))
        img = np.zeros((n, self.nObj, self.s, self.s), dtype=np.float32) if restore else None
        a0 = None if alpha0 is None else np.ascontiguousarray(alpha0, dtype=np.float64)
        self._check(L.lib().mfbd_solve(self.h, ptr(a0) if a0 is not None else dp(), res, ptr(alpha),
                                       ptr(img, fp) if restore else fp()), "solve")
        out = []
        for k in range(n):
            r = res[k]
            out.append(dict(alpha=alpha[k], img=None if img is None else img[k], status=r.status,
                            iterations=r.iterations, nF=r.nF, nDF=r.nDF, failCount=r.failCount,
                            initialMetric=r.initialMetric, finalMetric=r.finalMetric))
        return out

    # ---- stage-wise ----
    def init(self):
        self._check(L.lib().mfbd_init_patches(self.h), "init_patches")

    def metric(self, alpha, patches=None):
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, self.nImg, self.K)
        idx = np.ascontiguousarray(patches if patches is not None else np.arange(len(a)), dtype=np.int32)
        out = np.zeros(len(idx))
        self._check(L.lib().mfbd_eval_metric(self.h, len(idx), ptr(idx, ip), ptr(a), ptr(out)), "eval_metric")
        return out

    def grad(self, alpha, n_enabled=None, method="vogel", patches=None):
        m = {"fd": L.GRAD_FD, "vogel": L.GRAD_VOGEL, "adjoint": L.GRAD_ADJOINT}[method]
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, self.nImg, self.K)
        idx = np.ascontiguousarray(patches if patches is not None else np.arange(len(a)), dtype=np.int32)
        g = np.zeros((len(idx), self.nImg, self.K))
        ne = self.K if n_enabled is None else n_enabled
        self._check(L.lib().mfbd_eval_grad(self.h, len(idx), ptr(idx, ip), ptr(a), ne, m, ptr(g)), "eval_grad")
        return g

    def restore(self, alpha, patches=None):
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, self.nImg, self.K)
        idx = np.ascontiguousarray(patches if patches is not None else np.arange(len(a)), dtype=np.int32)
        img = np.zeros((len(idx), self.nObj, self.s, self.s), dtype=np.float32)
        self._check(L.lib().mfbd_restore(self.h, len(idx), ptr(idx, ip), ptr(a), ptr(img, fp)), "restore")
        return img

    def restore_outputs(self, alpha, mask, patch=0, patches=None):
        """Optional restoration outputs (Object::restorePatch, object.cpp:332-407): mask bits
        SAVE_COBJ, SAVE_PSF, SAVE_PSF_AVG (wins over SAVE_PSF), SAVE_RESIDUAL.  One patch
        (alpha [nImg][K], index `patch`): {"psf", "cobj", "res"} float32 [nImg or nObj][s][s],
        None where not asked.  With `patches` (alpha [n][nImg][K]) one call for all of them,
        each array with a leading [n]."""
        many = patches is not None
        idx = np.ascontiguousarray(patches if many else [patch], dtype=np.int32)
        n = len(idx)
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(n, self.nImg, self.K)
        if mask & L.SAVE_PSF_AVG:
            mask &= ~L.SAVE_PSF
        s = self.s
        psf = np.zeros((n, self.nObj if mask & L.SAVE_PSF_AVG else self.nImg, s, s), dtype=np.float32)
        cobj = np.zeros((n, self.nImg, s, s), dtype=np.float32)
        res = np.zeros((n, self.nImg, s, s), dtype=np.float32)
        self._check(L.lib().mfbd_restore_outputs(self.h, n, ptr(idx, ip), ptr(a), int(mask), ptr(psf, fp),
                                                 ptr(cobj, fp), ptr(res, fp)), "restore_outputs")
        out = {"psf": psf if mask & (L.SAVE_PSF | L.SAVE_PSF_AVG) else None,
               "cobj": cobj if mask & L.SAVE_COBJ else None, "res": res if mask & L.SAVE_RESIDUAL else None}
        return out if many else {k: (None if v is None else v[0]) for k, v in out.items()}

    def debug(self, what, p=0):
        cnt = np.zeros(1, dtype=np.int32)
        self._check(L.lib().mfbd_debug_get(self.h, what.encode(), p, dp(), ptr(cnt, ip)), "debug_get")
        out = np.zeros(int(cnt[0]))
        self._check(L.lib().mfbd_debug_get(self.h, what.encode(), p, ptr(out), ptr(cnt, ip)), "debug_get")
        return out

    def kernel_stats(self):
        cap = 64
        names = (C.c_char * 48 * cap)()
        ms = np.zeros(cap)
        launches = np.zeros(cap, dtype=np.int32)
        units = np.zeros(cap, dtype=np.int64)
        n = L.lib().mfbd_kernel_stats(self.h, cap, names, ptr(ms), ptr(launches, ip),
                                      units.ctypes.data_as(C.POINTER(C.c_longlong)))
        return {names[i].value.decode(): dict(ms=float(ms[i]), launches=int(launches[i]), units=int(units[i]))
                for i in range(n)}


def solve_group(solvers, restore=True):
    """mfbd_solve_group: the loaded patches of several contexts (the grid classes of an uneven
    batch) solved at once, one host thread per context.  Returns each context's results as
    Solver.run does."""
    n = len(solvers)
    res = [(Result * sv.nPatch)() for sv in solvers]
    alphas = [np.zeros((sv.nPatch, sv.nImg, sv.K)) for sv in solvers]
    imgs = [np.zeros((sv.nPatch, sv.nObj, sv.s, sv.s), dtype=np.float32) for sv in solvers] if restore else None
    H = (C.c_void_p * n)(*[sv.h.value for sv in solvers])
    R = (C.c_void_p * n)(*[C.cast(r, C.c_void_p).value for r in res])
    A = (dp * n)(*[ptr(a) for a in alphas])
    I = (fp * n)(*[ptr(i, fp) for i in imgs]) if restore else C.POINTER(fp)()
    rc = L.lib().mfbd_solve_group(n, H, C.POINTER(dp)(), R, A, I)
    if rc:
        msgs = "; ".join(L.lib().mfbd_last_error(sv.h).decode() for sv in solvers)
        raise SolverError(f"solve_group failed ({rc}): {msgs}")
    out = []
    for k, sv in enumerate(solvers):
        out.append([dict(alpha=alphas[k][i], img=None if imgs is None else imgs[k][i], status=r.status,
                         iterations=r.iterations, nF=r.nF, nDF=r.nDF, failCount=r.failCount,
                         initialMetric=r.initialMetric, finalMetric=r.finalMetric) for i, r in enumerate(res[k])])
    return out


def uneven_solvers(groups, precision="fp64", max_batch=None, gradient="reference"):
    """One loaded context per grid class of an uneven batch: `groups` is a list of
    (patches, shapes), the patches of one class (loaded with the class's largest mode and
    frame counts) and their mfbd_set_patch_shape shapes."""
    out = []
    for pbs, shapes in groups:
        sv = Solver(pbs[0], precision=precision, max_batch=max_batch or len(pbs), gradient=gradient)
        sv.load(pbs)
        sv.set_shapes(shapes)
        out.append(sv)
    return out


def mosaic(patches, pos_y, pos_x, clip=True, device=0):
    """Object::doMozaic of restored patches [n][s][s] placed at (pos_y, pos_x): the blended
    mosaic (trimmed of empty borders unless clip is False) and the (row, column) of its
    first pixel in the untrimmed frame."""
    pt = np.ascontiguousarray(patches, dtype=np.float32)
    n, s = pt.shape[0], pt.shape[1]
    py = np.ascontiguousarray(pos_y, dtype=np.int32)
    px = np.ascontiguousarray(pos_x, dtype=np.int32)
    rows, cols = np.zeros(1, np.int32), np.zeros(1, np.int32)
    rc = L.lib().mfbd_mosaic_size(n, s, ptr(py, ip), ptr(px, ip), ptr(rows, ip), ptr(cols, ip))
    if rc:
        raise SolverError(f"mosaic_size failed ({rc})")
    img = np.zeros(int(rows[0]) * int(cols[0]), dtype=np.float32)
    r0, c0 = np.zeros(1, np.int32), np.zeros(1, np.int32)
    rc = L.lib().mfbd_mosaic(n, s, ptr(pt, fp), ptr(py, ip), ptr(px, ip), int(clip), device, ptr(img, fp),
                             ptr(rows, ip), ptr(cols, ip), ptr(r0, ip), ptr(c0, ip))
    if rc:
        raise SolverError(f"mosaic failed ({rc})")
    return img[: int(rows[0]) * int(cols[0])].reshape(int(rows[0]), int(cols[0])), (int(r0[0]), int(c0[0]))


_ANA_TYPES = {np.dtype(np.int8): 0, np.dtype(np.int16): 1, np.dtype(np.int32): 2, np.dtype(np.float32): 3,
              np.dtype(np.float64): 4}
_FITS_BITPIX = {np.dtype(np.uint8): 8, np.dtype(np.int16): 16, np.dtype(np.int32): 32, np.dtype(np.int64): 64,
                np.dtype(np.float32): -32, np.dtype(np.float64): -64}


def write_ana(path, array, text="", slice=None):
    """Ana::write of an array (row-major, last axis fastest): uncompressed, or with `slice`
    (2-D int16 / int32) the crunch compression at that bit slice."""
    a = np.ascontiguousarray(array)
    if a.dtype not in _ANA_TYPES:
        raise SolverError(f"ANA has no type for {a.dtype}")
    if slice is not None:
        if a.ndim != 2 or a.dtype not in (np.dtype(np.int16), np.dtype(np.int32)):
            raise SolverError("ANA compression takes 2-D int16 or int32 data")
        rc = L.lib().mfbd_write_ana_compressed(str(path).encode(), _ANA_TYPES[a.dtype], a.shape[1], a.shape[0],
                                               a.ctypes.data_as(C.c_void_p), text.encode() if text else None,
                                               int(slice))
        if rc:
            raise SolverError(f"write_ana failed ({rc}): {L.lib().mfbd_file_last_error().decode()}")
        return
    dims = np.ascontiguousarray(a.shape[::-1], dtype=np.int32)
    rc = L.lib().mfbd_write_ana(str(path).encode(), _ANA_TYPES[a.dtype], a.ndim, ptr(dims, ip),
                                a.ctypes.data_as(C.c_void_p), text.encode() if text else None)
    if rc:
        raise SolverError(f"write_ana failed ({rc}): {L.lib().mfbd_file_last_error().decode()}")


def write_fits(path, array):
    """Fits::write of an array as the primary image (NAXIS1 = last axis)."""
    a = np.ascontiguousarray(array)
    if a.dtype not in _FITS_BITPIX:
        raise SolverError(f"FITS has no BITPIX for {a.dtype}")
    dims = np.ascontiguousarray(a.shape[::-1], dtype=np.int32)
    rc = L.lib().mfbd_write_fits(str(path).encode(), _FITS_BITPIX[a.dtype], a.ndim, ptr(dims, ip),
                                 a.ctypes.data_as(C.c_void_p))
    if rc:
        raise SolverError(f"write_fits failed ({rc}): {L.lib().mfbd_file_last_error().decode()}")


def write_object(base, mosaic_img, fmt="fits", as_int16=False, alpha=None, text=""):
    """Object::writeFits / writeAna (object.cpp:1212-1334): the mosaic to `base`.fits or
    `base`.f0 (float32, or int16 as the reference's default output type, converted by
    truncation; int16 ANA output crunch-compressed at slice 5 as the reference writes it),
    and with `alpha` ([patchesY][patchesX][images][modes]) the alphas as float32 to
    `base`.alpha.fits / .alpha.f0."""
    img = np.asarray(mosaic_img, dtype=np.float32)
    if as_int16:
        img = img.astype(np.int16)
    ext = ".fits" if fmt == "fits" else ".f0"
    if fmt == "fits":
        write_fits(str(base) + ext, img)
    else:
        write_ana(str(base) + ext, img, text, slice=5 if as_int16 else None)
    if alpha is not None:
        a = np.asarray(alpha, dtype=np.float32)
        (write_fits if fmt == "fits" else lambda p, x: write_ana(p, x))(str(base) + ".alpha" + ext, a)


def write_momfbd(path, img, region, nim, clip, version="20240612.1", date="", time="N/A", modified=None,
                 dx=None, dy=None, psf=None, cobj=None, res=None, alpha=None, pupil=None, modes=None,
                 pix2cf=0.0, cf2pix=0.0, div=None, file_names=()):
    """Object::writeMomfbd (object.cpp:1336-1565) through mfbd_write_momfbd: one object's
    patches on a [nPatchesY][nPatchesX] grid.
    img [py][px][s][s] restored objects; region [py][px][4] 1-based first x, last x, first y,
    last y of each patch; nim [nCh] images per channel; clip [4][nCh] (clipStartX, clipEndX,
    clipStartY, clipEndY, 1-based); dx, dy [py][px][nCh] channel offsets; psf / cobj / res
    [py][px][n][s][s]; alpha [py][px][images][modes]; pupil [nPH][nPH] and modes
    [nModes][nPH][nPH] (the reference writes the modes divided by the wavelength);
    div [nCh][nPH][nPH]; file_names the frames' file names."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    py, px, s = img.shape[0], img.shape[1], img.shape[2]
    keep = []

    def arr(a, dt, t):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(t))

    m = L.Momfbd()
    m.version, m.time, m.date = version.encode(), time.encode(), date.encode()
    m.modified = modified.encode() if modified else None
    nim = np.asarray(nim, dtype=np.int32)
    m.nCh = len(nim)
    m.clip = arr(np.asarray(clip).reshape(4, m.nCh), np.int16, C.c_int16)
    if pupil is not None:
        pup = np.asarray(pupil, dtype=np.float32)
        m.nPH = pup.shape[0]
        m.pupil = arr(pup, np.float32, C.c_float)
        if modes is not None:
            md = np.asarray(modes, dtype=np.float32)
            m.nModes = md.shape[0]
            m.modes = arr(md, np.float32, C.c_float)
        m.pix2cf, m.cf2pix = float(pix2cf), float(cf2pix)
    m.nPatchesY, m.nPatchesX, m.nPoints = py, px, s
    m.region = arr(region, np.int32, C.c_int)
    m.nim = arr(nim, np.int32, C.c_int)
    m.dx = arr(dx, np.int32, C.c_int)
    m.dy = arr(dy, np.int32, C.c_int)
    m.img = arr(img, np.float32, C.c_float)
    for name, block in (("psf", psf), ("cobj", cobj), ("res", res)):
        if block is not None:
            b = np.asarray(block, dtype=np.float32)
            setattr(m, {"psf": "npsf", "cobj": "nobj", "res": "nres"}[name], b.shape[2])
            setattr(m, name, arr(b, np.float32, C.c_float))
    if alpha is not None:
        a = np.asarray(alpha, dtype=np.float32)
        m.nalpha, m.nm = a.shape[2], a.shape[3]
        m.alpha = arr(a, np.float32, C.c_float)
    if div is not None:
        d = np.asarray(div, dtype=np.float32)
        if pupil is None:
            m.nPH = d.shape[1]
        m.ndiv = d.shape[0]
        m.div = arr(d, np.float32, C.c_float)
    names = (C.c_char_p * max(1, len(file_names)))(*[f.encode() for f in file_names])
    m.nFileNames = len(file_names)
    m.fileNames = names if file_names else None
    rc = L.lib().mfbd_write_momfbd(str(path).encode(), C.byref(m))
    if rc:
        raise SolverError(f"write_momfbd failed ({rc}): {L.lib().mfbd_file_last_error().decode()}")
