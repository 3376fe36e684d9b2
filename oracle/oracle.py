"""Oracle (test infrastructure only): ctypes wrapper around oracle/_build/libmfbd_oracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The problem dictionary format is the one built
by redux_amd.synth (objects / channels / images / nullspace / settings).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmfbd_oracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)


class OProblem(C.Structure):
    _fields_ = [
        ("s", C.c_int), ("P", C.c_int), ("N", C.c_int), ("K", C.c_int),
        ("nObj", C.c_int), ("nCh", C.c_int), ("nImg", C.c_int),
        ("weight", dp), ("wavelength", dp), ("pupil", dp), ("modes", dp), ("tilt", ip),
        ("jac", dp), ("freqCutoff", dp),
        ("chObj", ip), ("phiFixed", dp), ("resOffset", dp), ("noiseFudge", dp),
        ("imgCh", ip), ("data", fp), ("dataY", C.c_int), ("dataX", C.c_int),
        ("startY", C.c_int), ("startX", C.c_int),
        ("nPar", C.c_int), ("nFree", C.c_int), ("nnz", C.c_int),
        ("nsRow", ip), ("nsCol", ip), ("nsVal", dp),
        ("FTOL", C.c_double), ("EPS", C.c_double), ("graddiffStep", C.c_double), ("regAlpha", C.c_double),
        ("minIter", C.c_int), ("maxIter", C.c_int), ("nDoneIter", C.c_int), ("modeStart", C.c_int),
        ("modeStep", C.c_int), ("apodSize", C.c_int), ("noFilter", C.c_int),
        ("alphaInit", dp), ("maxMetrics", C.c_int), ("getstep", C.c_int),
        ("metricNoise", C.c_double), ("noiseSeed", C.c_ulonglong),
    ]


GETSTEP = {"steepest": 1, "steepest_descent": 1, "bfgs": 3, "conjugate": 0, "conjugate_gradient": 0, "bfgs_inv": 4}    # getstepMap (config.cpp:102-114)


class OFdf(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("f", C.c_void_p), ("df", C.c_void_p), ("fdf", C.c_void_p)]


F_T = C.CFUNCTYPE(C.c_double, C.c_void_p, dp)
DF_T = C.CFUNCTYPE(None, C.c_void_p, dp, dp)
FDF_T = C.CFUNCTYPE(None, C.c_void_p, dp, dp, dp)


class OResult(C.Structure):
    _fields_ = [
        ("alpha", dp), ("img", dp), ("metrics", dp), ("nMetrics", C.c_int),
        ("initialMetric", C.c_double), ("finalMetric", C.c_double),
        ("totalIterations", C.c_int), ("nF", C.c_int), ("nDF", C.c_int), ("status", C.c_int),
        ("failCount", C.c_int), ("shifts", ip),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.o_solve.argtypes = [C.POINTER(OProblem), C.POINTER(OResult)]
        L.o_solve_many.argtypes = [C.POINTER(OProblem), C.POINTER(OResult), C.c_int, C.c_int]
        L.o_state_new.argtypes = [C.POINTER(OProblem)]
        L.o_bfgs2_new.argtypes = [C.c_int]
        L.o_bfgs2_new.restype = C.c_void_p
        L.o_bfgs2_free.argtypes = [C.c_void_p]
        L.o_bfgs2_set.argtypes = [C.c_void_p, C.POINTER(OFdf), dp, dp, dp, C.c_double, C.c_double]
        L.o_bfgs2_iterate.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.o_bfgs2_iterate.restype = C.c_int
        L.o_step_new.argtypes = [C.c_int, C.c_int]
        L.o_step_new.restype = C.c_void_p
        L.o_step_free.argtypes = [C.c_void_p]
        L.o_step_set.argtypes = [C.c_void_p, C.POINTER(OFdf), dp, dp, dp, C.c_double, C.c_double]
        L.o_step_iterate.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.o_step_iterate.restype = C.c_int
        L.o_state_new.restype = C.c_void_p
        L.o_state_free.argtypes = [C.c_void_p]
        L.o_state_init.argtypes = [C.c_void_p]
        L.o_state_metric.argtypes = [C.c_void_p, dp]
        L.o_state_metric.restype = C.c_double
        for f in (L.o_state_grad_vogel, L.o_state_grad_fd):
            f.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.o_state_restore.argtypes = [C.c_void_p, dp, dp]
        L.o_state_restore_full.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.o_scharmer_filter.argtypes = [C.c_void_p, dp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.o_cg_new.argtypes = [C.c_int]
        L.o_cg_new.restype = C.c_void_p
        L.o_cg_free.argtypes = [C.c_void_p]
        L.o_cg_set.argtypes = [C.c_void_p, C.POINTER(OFdf), dp, dp, C.c_double, C.c_double]
        L.o_cg_iterate.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.o_cg_iterate.restype = C.c_int
        L.o_cg_direction.argtypes = [C.c_void_p, dp]
        L.o_bracket.argtypes = [C.POINTER(OFdf), dp, dp, dp, dp, dp, dp, C.c_int]
        L.o_brent.argtypes = [C.POINTER(OFdf), C.c_double, C.c_double, dp, dp, C.c_double, C.c_int]
        L.o_state_restore_outputs.argtypes = [C.c_void_p, dp, C.c_int, dp, dp, dp]
        L.o_state_get.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.o_state_scalar.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.o_state_scalar.restype = C.c_double
        L.o_fft2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.o_reorder_into_d.argtypes = [dp, C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_int]
        L.o_autocorrelate_d.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.o_autocorrelate_c.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.o_noise.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.o_noise.restype = C.c_double
        ip = C.POINTER(C.c_int)
        L.o_blend_window.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.o_mosaic.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float), ip, ip, C.c_int, C.POINTER(C.c_float),
                               ip, ip, ip, ip]
        _lib = L
    return _lib


def _p(a, t=dp):
    return a.ctypes.data_as(t)


def fft2(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().o_fft2(a.ctypes.data, a.shape[0], a.shape[1], sign)
    return a


def reorder_into(a, out_shape=None, center=False):
    a = np.ascontiguousarray(a, dtype=np.float64)
    oy, ox = out_shape if out_shape else a.shape
    out = np.zeros((oy, ox))
    lib().o_reorder_into_d(_p(a), a.shape[0], a.shape[1], _p(out), oy, ox, int(center))
    return out


def autocorrelate(a, center=False):
    if np.iscomplexobj(a):
        a = np.ascontiguousarray(a, dtype=np.complex128).copy()
        lib().o_autocorrelate_c(a.ctypes.data, a.shape[0], a.shape[1], int(center))
        return a
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    lib().o_autocorrelate_d(_p(a), a.shape[0], a.shape[1], int(center))
    return a


def blend_window(n_rows, n_cols, blend, margin):
    """apodizeInPlace of an all-ones array (src/lib/image/utils.cpp:584-613)."""
    w = np.zeros((n_rows, n_cols))
    lib().o_blend_window(_p(w), n_rows, n_cols, blend, margin)
    return w


def mosaic(patches, pos_y, pos_x, clip=True):
    """Object::doMozaic (src/lib/momfbd/object.cpp:1160-1207): blended mosaic of patches
    [n][s][s] at (pos_y, pos_x), trimmed unless clip is False, and its first (row, column)."""
    pt = np.ascontiguousarray(patches, dtype=np.float32)
    n, s = pt.shape[0], pt.shape[1]
    py = np.ascontiguousarray(pos_y, dtype=np.int32)
    px = np.ascontiguousarray(pos_x, dtype=np.int32)
    img = np.zeros((int(py.max()) + s + 1) * (int(px.max()) + s + 1), dtype=np.float32)
    o = [np.zeros(1, np.int32) for _ in range(4)]
    ip, fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
    rc = lib().o_mosaic(n, s, _p(pt, fp), _p(py, ip), _p(px, ip), int(clip), _p(img, fp), *[_p(a, ip) for a in o])
    assert rc == 0
    rows, cols = int(o[0][0]), int(o[1][0])
    return img[: rows * cols].reshape(rows, cols), (int(o[2][0]), int(o[3][0]))


def noise(ft, mask=-1, limit=-1.0):
    ft = np.ascontiguousarray(ft, dtype=np.complex128)
    return lib().o_noise(ft.ctypes.data, ft.shape[0], ft.shape[1], mask, limit)


class Problem:
    """Holds numpy arrays alive and exposes them as an OProblem."""

    def __init__(self, pb: dict, alpha_init=None, metric_noise=0.0, noise_seed=0):
        objs, chans = pb["objects"], pb["channels"]
        K, P = pb["K"], pb["P"]
        self.keep = {}

        def arr(name, a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep[name] = a
            return a

        st = pb["settings"]
        ns = pb["nullspace"]
        o = OProblem()
        o.s, o.P, o.N, o.K = pb["s"], P, 2 * P, K
        o.nObj, o.nCh, o.nImg = len(objs), len(chans), len(pb["imgCh"])
        o.weight = _p(arr("w", [ob["weight"] for ob in objs], np.float64))
        o.wavelength = _p(arr("wl", [ob["wavelength"] for ob in objs], np.float64))
        o.pupil = _p(arr("pup", np.stack([ob["pupil"] for ob in objs]), np.float64))
        o.modes = _p(arr("modes", np.stack([ob["modes"] for ob in objs]), np.float64))
        o.tilt = _p(arr("tilt", np.array([ob["tilt"] for ob in objs]), np.int32), ip)
        jac = np.stack([np.concatenate([np.ravel(ob["J"]), np.ravel(ob["Ji"])]) for ob in objs])
        o.jac = _p(arr("jac", jac, np.float64))
        o.freqCutoff = _p(arr("fc", [ob["freqCutoff"] for ob in objs], np.float64))
        o.chObj = _p(arr("chObj", [c["obj"] for c in chans], np.int32), ip)
        o.phiFixed = _p(arr("phiF", np.stack([c["phiFixed"] for c in chans]), np.float64))
        o.resOffset = _p(arr("res", np.array([c["resOffset"] for c in chans]), np.float64))
        o.noiseFudge = _p(arr("nf", [c["noiseFudge"] for c in chans], np.float64))
        o.imgCh = _p(arr("imgCh", pb["imgCh"], np.int32), ip)
        data = arr("data", pb["data"], np.float32)
        o.data = _p(data, fp)
        o.dataY, o.dataX = data.shape[1], data.shape[2]
        o.startY, o.startX = pb["start"]
        o.nPar, o.nFree = ns["nPar"], ns["nFree"]
        o.nnz = len(ns["rows"])
        o.nsRow = _p(arr("nsr", ns["rows"], np.int32), ip)
        o.nsCol = _p(arr("nsc", ns["cols"], np.int32), ip)
        o.nsVal = _p(arr("nsv", ns["vals"], np.float64))
        o.FTOL, o.EPS, o.graddiffStep, o.regAlpha = st["FTOL"], st["EPS"], st["GRADDIFF_STEP"], 0.0
        o.minIter, o.maxIter, o.nDoneIter = st["MIN_ITER"], st["MAX_ITER"], st["N_DONE_ITER"]
        o.modeStart, o.modeStep = st["MODE_START"], st["MODE_STEP"]
        o.apodSize, o.noFilter = st.get("APODIZATION", -1), int(st.get("NO_FILTER", 0))
        if alpha_init is not None:
            o.alphaInit = _p(arr("ainit", alpha_init, np.float64))
        else:
            o.alphaInit = dp()
        o.maxMetrics = 100000
        o.getstep = GETSTEP[str(st.get("GETSTEP", "conjugate")).lower()]
        o.metricNoise, o.noiseSeed = float(metric_noise), int(noise_seed)
        self.o = o
        self.nObj, self.nImg, self.K, self.s = o.nObj, o.nImg, K, o.s
        self.nPar = o.nPar


def solve(pb: dict, alpha_init=None, restore=True):
    prob = Problem(pb, alpha_init)
    return _solve_many([prob], restore)[0]


def solve_many(pbs, restore=True, threads=0, metric_noise=0.0, noise_seeds=None):
    """Oracle solves of independent patches on host threads.  metric_noise > 0 perturbs every
    metric value the minimiser sees (relative, uniform; sensitivity experiments only), with
    noise_seeds[k] seeding patch k's stream."""
    seeds = noise_seeds if noise_seeds is not None else [0] * len(pbs)
    probs = [Problem(pb, None, metric_noise, sd) for pb, sd in zip(pbs, seeds)]
    return _solve_many(probs, restore, threads)


def _solve_many(probs, restore=True, threads=0):
    n = len(probs)
    Parr = (OProblem * n)(*[p.o for p in probs])
    Rarr = (OResult * n)()
    bufs = []
    for k, p in enumerate(probs):
        alpha = np.zeros(p.nPar)
        img = np.zeros((p.nObj, p.s, p.s)) if restore else None
        metrics = np.zeros(100000)
        shifts = np.zeros((p.nImg, 2), dtype=np.int32)
        bufs.append((alpha, img, metrics, shifts))
        Rarr[k].alpha = _p(alpha)
        Rarr[k].img = _p(img) if restore else dp()
        Rarr[k].metrics = _p(metrics)
        Rarr[k].shifts = _p(shifts, ip)
    rc = lib().o_solve_many(Parr, Rarr, n, threads)
    if rc:
        raise RuntimeError("oracle solve failed")
    out = []
    for k, p in enumerate(probs):
        alpha, img, metrics, shifts = bufs[k]
        r = Rarr[k]
        out.append(dict(alpha=alpha.reshape(p.nImg, p.K), img=img, metrics=metrics[: r.nMetrics].copy(),
                        initialMetric=r.initialMetric, finalMetric=r.finalMetric,
                        iterations=r.totalIterations, nF=r.nF, nDF=r.nDF, status=r.status,
                        failCount=r.failCount, shifts=shifts))
    return out


class State:
    """Stage-wise evaluation (metric, gradients, restore) at arbitrary alpha."""

    def __init__(self, pb: dict):
        self.prob = Problem(pb)
        self.h = lib().o_state_new(C.byref(self.prob.o))
        lib().o_state_init(self.h)
        o = self.prob.o
        self.nImg, self.K, self.nObj, self.N, self.P, self.s = o.nImg, o.K, o.nObj, o.N, o.P, o.s

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_state_free(self.h)
            self.h = None

    def metric(self, alpha):
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        return lib().o_state_metric(self.h, _p(a))

    def grad(self, alpha, n_enabled=None, method="vogel"):
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        g = np.zeros(self.nImg * self.K)
        ne = self.K if n_enabled is None else n_enabled
        f = lib().o_state_grad_vogel if method == "vogel" else lib().o_state_grad_fd
        f(self.h, _p(a), ne, _p(g))
        return g.reshape(self.nImg, self.K)

    def restore(self, alpha):
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        img = np.zeros((self.nObj, self.s, self.s))
        lib().o_state_restore(self.h, _p(a), _p(img))
        return img

    def restore_full(self, alpha, obj=0):
        """Object obj's restored object on the full N x N grid before the centred crop
        (object.cpp:298-331: conj, safeDivide, Scharmer filter unless NO_FILTER, normalised
        inverse transform)."""
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        full = np.zeros((self.N, self.N))
        lib().o_state_restore_full(self.h, _p(a), int(obj), _p(full))
        return full

    def restore_outputs(self, alpha, mask):
        """Optional restoration outputs (object.cpp:332-407): mask bits 2 (convolved objects),
        32 (PSF per image), 64 (PSF averaged per object; wins over 32), 128 (residuals).
        Returns {"psf", "cobj", "res"}: [nImg or nObj][s][s] float values, None if not asked."""
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        s = self.s
        psf = np.zeros((self.nObj if mask & 64 else self.nImg, s, s))
        cobj = np.zeros((self.nImg, s, s))
        res = np.zeros((self.nImg, s, s))
        lib().o_state_restore_outputs(self.h, _p(a), int(mask), _p(psf), _p(cobj), _p(res))
        return {"psf": psf if mask & 96 else None, "cobj": cobj if mask & 2 else None,
                "res": res if mask & 128 else None}

    def get(self, what):
        shapes = {"otf": (self.nImg, self.N, self.N, 2), "imgft": (self.nImg, self.N, self.N, 2),
                  "ftsum": (self.nObj, self.N, self.N), "P": (self.nObj, self.N, self.N, 2),
                  "Q": (self.nObj, self.N, self.N), "phi": (self.nImg, self.P, self.P),
                  "vogel": (self.nImg, self.P, self.P), "phich": (len(self.prob.keep["chObj"]), self.P, self.P),
                  "window": (self.s, self.s), "noisewindow": (self.s, self.s)}
        out = np.zeros(shapes[what])
        lib().o_state_get(self.h, what.encode(), _p(out))
        if out.ndim and out.shape[-1] == 2 and what in ("otf", "imgft", "P"):
            out = out[..., 0] + 1j * out[..., 1]
        return out

    def scalar(self, what, idx=0):
        return lib().o_state_scalar(self.h, what.encode(), idx)


def step_minimize(kind, f, grad, x0, step=1.0, tol=1e-6, iters=100):
    """Drive the steepest_descent (kind 1) or vector_bfgs (kind 3) restatement on a Python
    function (tests); returns the iterates as bfgs2_minimize does."""
    L = lib()
    n = len(x0)
    calls = {"f": 0, "df": 0}

    def _f(ctx, x):
        calls["f"] += 1
        return float(f(np.ctypeslib.as_array(x, (n,)).copy()))

    def _df(ctx, x, g):
        calls["df"] += 1
        np.ctypeslib.as_array(g, (n,))[:] = grad(np.ctypeslib.as_array(x, (n,)).copy())

    def _fdf(ctx, x, fo, g):
        calls["f"] += 1
        calls["df"] += 1
        xv = np.ctypeslib.as_array(x, (n,)).copy()
        fo[0] = float(f(xv))
        np.ctypeslib.as_array(g, (n,))[:] = grad(xv)

    cf, cdf, cfdf = F_T(_f), DF_T(_df), FDF_T(_fdf)
    fn = OFdf(None, C.cast(cf, C.c_void_p), C.cast(cdf, C.c_void_p), C.cast(cfdf, C.c_void_p))
    w = L.o_step_new(n, kind)
    x = np.array(x0, dtype=np.float64)
    g = np.zeros(n)
    dx = np.zeros(n)
    fv = np.zeros(1)
    try:
        L.o_step_set(w, C.byref(fn), _p(x), _p(fv), _p(g), step, tol)
        out = [(x.copy(), float(fv[0]), g.copy(), 0, dict(calls))]
        for _ in range(iters):
            st = L.o_step_iterate(w, _p(x), _p(fv), _p(g), _p(dx))
            out.append((x.copy(), float(fv[0]), g.copy(), st, dict(calls)))
            if st != 0:
                break
    finally:
        L.o_step_free(w)
    return out


def bfgs2_minimize(f, grad, x0, step=1.0, tol=1e-6, iters=100):
    """Drive the vector_bfgs2 restatement on a Python function (tests): f(x) -> float,
    grad(x) -> array.  Returns the iterates [(x, f, g, status, evals)] after set and
    after each iteration; evals counts f and gradient calls so far."""
    L = lib()
    n = len(x0)
    calls = {"f": 0, "df": 0}

    def _f(ctx, x):
        calls["f"] += 1
        return float(f(np.ctypeslib.as_array(x, (n,)).copy()))

    def _df(ctx, x, g):
        calls["df"] += 1
        np.ctypeslib.as_array(g, (n,))[:] = grad(np.ctypeslib.as_array(x, (n,)).copy())

    def _fdf(ctx, x, fo, g):
        calls["f"] += 1
        calls["df"] += 1
        xv = np.ctypeslib.as_array(x, (n,)).copy()
        fo[0] = float(f(xv))
        np.ctypeslib.as_array(g, (n,))[:] = grad(xv)

    cf, cdf, cfdf = F_T(_f), DF_T(_df), FDF_T(_fdf)
    fn = OFdf(None, C.cast(cf, C.c_void_p), C.cast(cdf, C.c_void_p), C.cast(cfdf, C.c_void_p))
    w = L.o_bfgs2_new(n)
    x = np.array(x0, dtype=np.float64)
    g = np.zeros(n)
    dx = np.zeros(n)
    fv = np.zeros(1)
    try:
        L.o_bfgs2_set(w, C.byref(fn), _p(x), _p(fv), _p(g), step, tol)
        out = [(x.copy(), float(fv[0]), g.copy(), 0, dict(calls))]
        for _ in range(iters):
            st = L.o_bfgs2_iterate(w, _p(x), _p(fv), _p(g), _p(dx))
            out.append((x.copy(), float(fv[0]), g.copy(), st, dict(calls)))
            if st != 0:
                break
    finally:
        L.o_bfgs2_free(w)
    return out


def scharmer_filter(data, q2, noise_power, cutoff):
    """ScharmerFilter (src/lib/image/utils.cpp:334-392) applied to the centred complex grid
    data with the summed |S|^2 q2; returns the filtered copy."""
    d = np.ascontiguousarray(data, dtype=np.complex128).copy()
    q = np.ascontiguousarray(q2, dtype=np.float64)
    lib().o_scharmer_filter(d.ctypes.data, _p(q), d.shape[0], d.shape[1], float(noise_power), float(cutoff))
    return d


def _fdf1(f):
    """OFdf over a function of one variable (x[0]); keeps the callbacks alive."""
    def _f(ctx, x):
        return float(f(x[0]))
    cf = F_T(_f)
    fn = OFdf(None, C.cast(cf, C.c_void_p), None, None)
    return fn, cf


def bracket(f, a, b, limit=100):
    """math::bracket (src/lib/math/helpers.cpp:58-129) from the points a, b; returns
    (a, mid, b, fa, fmid, fb) with a <= mid <= b as the reference sorts them."""
    fn, keep = _fdf1(f)
    v = [np.array([float(t)]) for t in (a, b, 0.0, f(a), f(b), 0.0)]
    lib().o_bracket(C.byref(fn), *[_p(t) for t in v], int(limit))
    a, b, mid, fa, fb, fmid = [float(t[0]) for t in v]
    return a, mid, b, fa, fmid, fb


def brent(f, ax, bx, x, tol, limit=100):
    """math::brent (src/lib/math/helpers.cpp:133-216) in [ax, bx] from x; returns (x, f(x))."""
    fn, keep = _fdf1(f)
    xv, fv = np.array([float(x)]), np.zeros(1)
    lib().o_brent(C.byref(fn), float(ax), float(bx), _p(xv), _p(fv), float(tol), int(limit))
    return float(xv[0]), float(fv[0])


def cg_minimize(f, grad, x0, step=1.0, tol=1e-6, iters=100):
    """Drive the conjugate_rdx restatement (conjugate_gradient.cpp:72-256, with bracket and
    brent) on a Python function.  Returns the iterates [(x, f, g, status, evals, p)] after
    set and after each iteration; p is the search direction that iteration used."""
    L = lib()
    n = len(x0)
    calls = {"f": 0, "df": 0}

    def _f(ctx, x):
        calls["f"] += 1
        return float(f(np.ctypeslib.as_array(x, (n,)).copy()))

    def _df(ctx, x, g):
        calls["df"] += 1
        np.ctypeslib.as_array(g, (n,))[:] = grad(np.ctypeslib.as_array(x, (n,)).copy())

    cf, cdf = F_T(_f), DF_T(_df)
    fn = OFdf(None, C.cast(cf, C.c_void_p), C.cast(cdf, C.c_void_p), None)
    w = L.o_cg_new(n)
    x = np.array(x0, dtype=np.float64)
    g = np.zeros(n)
    dx = np.zeros(n)
    p = np.zeros(n)
    fv = np.zeros(1)
    try:
        L.o_cg_set(w, C.byref(fn), _p(x), _p(fv), step, tol)
        out = [(x.copy(), float(fv[0]), g.copy(), 0, dict(calls), p.copy())]
        for _ in range(iters):
            st = L.o_cg_iterate(w, _p(x), _p(fv), _p(g), _p(dx))
            L.o_cg_direction(w, _p(p))
            out.append((x.copy(), float(fv[0]), g.copy(), st, dict(calls), p.copy()))
            if st != 0:
                break
    finally:
        L.o_cg_free(w)
    return out
