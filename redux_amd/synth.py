"""Seeded synthetic MOMFBD patches (objects, wavefronts, frames).

Builds problem dictionaries of the shape the solver consumes (the same layout the
reference's PatchData/Object/Channel carry into Solver::run):
  objects : weight, wavelength, pupil[P,P], modes[K,P,P] (scaled to RMS 1/lambda),
            tilt (x,y) mode indices, J/Ji tilt Jacobians, freqCutoff
  channels: obj, phiFixed[P,P] (diversity), resOffset (x,y), noiseFudge
  imgCh   : channel of each image (images ordered object -> channel -> frame)
  data    : float32 [nImg, s+2m, s+2m] data blocks, patch starts at (m, m)
  nullspace: nPar, nFree, rows, cols, vals (Constraints::init layout)
  settings: solver knobs with the reference's names

Truth:
  * object: Gaussian random field with a k^-3 power spectrum, mean `mean_counts`;
  * wavefronts: Kolmogorov mode coefficients drawn with the Noll covariance
    (Zernike) or the KL variances, times (D/r0)^(5/3);
  * frames: object convolved with the model PSF of each wavefront (plus the
    channel's diversity), then Poisson noise.

The setup provider (pupil, modes, Jacobian, nullspace, covariance) is passed in;
redux_amd.setup (the C++ implementation behind the C-ABI) is the default.
"""
from __future__ import annotations

import math

import numpy as np

SST_D = 0.97
WAVELENGTH = 6.302e-7


def _setup_default():
    from . import setup as S
    return S


def grf_field(n, seed, mean=1000.0, contrast=0.1, slope=-3.0):
    """Gaussian random field with power spectrum ~ k^slope on an n x n grid."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.fftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** (slope / 2.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * amp))
    f = (f - f.mean()) / f.std()
    return mean * (1.0 + contrast * f)


def geometry(s, pupil_diameter=None):
    """Optics giving pupil radius = s/4 on P = s/2 (critical sampling, N = s) by default,
    or a chosen pupil diameter; returns dict(arcsec, fc, radius, P)."""
    lam, D = WAVELENGTH, SST_D
    r = (pupil_diameter if pupil_diameter else s / 2.0) / 2.0
    fc = 2.0 * r
    rad_per_pix = fc * lam / (s * D)
    arcsec = rad_per_pix / (math.pi / (180.0 * 3600.0))
    return dict(arcsec=arcsec, lam=lam, D=D)


def build_setup(cfg: dict, setup=None):
    """Geometry, pupil, scaled modes, Jacobian, nullspace and wavefront covariance
    shared by every patch of a configuration."""
    S = setup or _setup_default()
    s = cfg["s"]
    geo = geometry(s, cfg.get("pupil_diameter"))
    fc, radius, P = S.calculate_pupil_size(geo["lam"], s, geo["D"], geo["arcsec"])
    pup = S.generate_pupil(P, radius)
    modes_ids = list(cfg["modes"])
    K = len(modes_ids)
    basis = cfg.get("basis", "zernike")
    kl_max = cfg.get("kl_max", max(modes_ids)) if basis == "kl" else 2000
    B, atm_rms, tilt = S.mode_set(P, radius, modes_ids, basis=basis, kl_min=2, kl_max=kl_max)
    lam = geo["lam"]
    Bs = S.normalize_modes(B, pup["pupil"], pup["area"], scale=1.0 / lam)
    jscale = lam * (0.5 * fc) / S.pix2cf(geo["arcsec"], geo["D"])
    tilt_m, J, Ji = S.measure_jacobian(Bs, pup["pupil"], jscale)
    nO, nF = cfg.get("n_objects", 1), cfg["n_frames"]
    cov = np.array([[S.noll_cov(i, j) if basis == "zernike" else 0.0 for j in modes_ids] for i in modes_ids])
    if basis != "zernike":
        for k, j in enumerate(modes_ids):
            cov[k, k] = atm_rms[k] ** 2
            if j in (2, 3):
                for k2, j2 in enumerate(modes_ids):
                    if j2 in (2, 3):
                        cov[k, k2] = S.noll_cov(j, j2)
    cov *= cfg["d_r0"] ** (5.0 / 3.0)
    div = cfg.get("diversity", [0.0] * nO)
    z4 = S.mode_set(P, radius, [4], basis="zernike")[0][0] if any(div) else None
    wfs = [[list(range(nF))] for _ in range(nO)]
    nPar, nFree, rows, cols, vals = S.constraints([[nF] for _ in range(nO)], wfs, K)
    settings = dict(FTOL=1e-3, EPS=1e-10, GRADDIFF_STEP=1e-2, MIN_ITER=5, MAX_ITER=500, N_DONE_ITER=3,
                    MODE_START=5, MODE_STEP=5, APODIZATION=-1, NO_FILTER=0)
    settings.update(cfg.get("settings", {}))
    return dict(cfg=cfg, s=s, P=P, N=2 * P, K=K, fc=fc, radius=radius, pup=pup, Bs=Bs, tilt=tilt_m, J=J, Ji=Ji,
                lam=lam, geo=geo, chol=np.linalg.cholesky(cov + 1e-12 * np.eye(K)), div=div, z4=z4,
                nullspace=dict(nPar=nPar, nFree=nFree, rows=rows, cols=cols, vals=vals), settings=settings,
                atm_rms=atm_rms, modes=modes_ids, basis=basis)


def build_patch(su: dict, seed: int, field=None, field_offset=(0, 0)):
    """One patch: wavefronts drawn from `seed`, frames from `field` (a list of object
    images, one per object) or from a per-patch random field."""
    cfg = su["cfg"]
    s, P, N, K, lam = su["s"], su["P"], su["N"], su["K"], su["lam"]
    pupil, ps = su["pup"]["pupil"], su["pup"]["ps"]
    Bs = su["Bs"]
    nO, nF = cfg.get("n_objects", 1), cfg["n_frames"]
    margin = cfg.get("margin", 5)
    rng = np.random.default_rng(seed)
    a_rad = (su["chol"] @ rng.standard_normal((K, nF))).T          # [nF, K] radians, unit-RMS modes
    alpha_true = a_rad * lam                                        # reference alpha units
    M = s + 2 * margin
    G = N if N >= M else int(2 ** math.ceil(math.log2(M)))
    objects, channels, imgCh, data = [], [], [], []
    py, px = ps // P, ps % P
    for o in range(nO):
        objects.append(dict(weight=1.0, wavelength=lam, pupil=pupil, modes=Bs, tilt=su["tilt"], J=su["J"],
                            Ji=su["Ji"], freqCutoff=su["fc"]))
        d = su["div"][o]
        phi_fixed = d * su["z4"] if d else np.zeros((P, P))
        channels.append(dict(obj=o, phiFixed=phi_fixed, resOffset=(0.0, 0.0), noiseFudge=1.0))
        if field is None:
            obj_img = grf_field(G, seed * 7919 + 31 * o + 1, mean=cfg.get("mean_counts", 1000.0))
        else:
            fy, fx = field_offset
            obj_img = field[o][fy:fy + G, fx:fx + G]
        F_obj = np.fft.fft2(obj_img)
        for f in range(nF):
            # model PSF on the N grid (same sampling as the solver's OTF), embedded
            # centred in the G x G convolution grid.  FFT(psf) must equal the model
            # OTF S = IFFT(|FFT(a)|^2)/N^2, i.e. psf(n) = N^2 |IFFT(a)(n)|^2.
            phi = phi_fixed + np.tensordot(alpha_true[f], Bs, axes=(0, 0))
            a = np.zeros((N, N), dtype=complex)
            a[(py - P // 2) % N, (px - P // 2) % N] = pupil.ravel()[ps] * np.exp(1j * phi.ravel()[ps])
            psfN = np.fft.fftshift(np.abs(np.fft.ifft2(a)) ** 2)
            psfN /= psfN.sum()
            psf = np.zeros((G, G))
            o0 = (G - N) // 2
            psf[o0:o0 + N, o0:o0 + N] = psfN
            img = np.real(np.fft.ifft2(F_obj * np.fft.fft2(np.fft.ifftshift(psf))))
            off = (G - M) // 2
            blk = img[off:off + M, off:off + M]
            data.append(rng.poisson(np.clip(blk, 0, None)).astype(np.float32))
            imgCh.append(o)
    return dict(s=s, P=P, N=N, K=K, objects=objects, channels=channels, imgCh=np.array(imgCh, dtype=np.int32),
                data=np.stack(data), start=(margin, margin), nullspace=su["nullspace"], settings=su["settings"],
                truth=dict(alpha=np.tile(alpha_true, (nO, 1)), a_rad=a_rad),
                meta=dict(radius=su["radius"], area=su["pup"]["area"], arcsec=su["geo"]["arcsec"], D=su["geo"]["D"],
                          basis=su["basis"], modes=su["modes"], atm_rms=su["atm_rms"]))


def build_problem(cfg: dict, seed: int = 1, setup=None, field=None, field_offset=(0, 0)):
    """cfg keys: s, n_frames (per object), n_objects, modes (list of Noll/KL ids), basis
    ('zernike'|'kl'), d_r0, mean_counts, diversity (list per object, RMS rad of Z4),
    margin, pupil_diameter (optional), kl_max (optional), settings overrides."""
    return build_patch(build_setup(cfg, setup), seed, field, field_offset)


def build_field_batch(cfg: dict, n_patches: int, seed: int = 2024, field_size: int = 1024, setup=None, picks=None):
    """n_patches overlapping patches on a square grid over one seeded k^-3 field per
    object (the benchmark's config-2 layout: 256 patches from a 1024 x 1024 field); with
    `picks`, only those patches of the layout (same seeds and positions)."""
    su = build_setup(cfg, setup)
    nO = cfg.get("n_objects", 1)
    fields = [grf_field(field_size, seed + 17 * o, mean=cfg.get("mean_counts", 1000.0)) for o in range(nO)]
    side = int(math.ceil(math.sqrt(n_patches)))
    N, s, margin = su["N"], su["s"], cfg.get("margin", 5)
    G = N if N >= s + 2 * margin else int(2 ** math.ceil(math.log2(s + 2 * margin)))
    stride = (field_size - G) // max(side - 1, 1)
    out = []
    for k in (range(n_patches) if picks is None else picks):
        iy, ix = divmod(k, side)
        out.append(build_patch(su, seed * 1000 + k, field=fields, field_offset=(iy * stride, ix * stride)))
    return su, out


def ragged_patch(su: dict, seed: int, n_modes: int, frames, field=None, field_offset=(0, 0)):
    """A patch of the class `su` (whose K and n_frames are the class maxima) that solves for
    its first `n_modes` modes with frames[o] frames of object o.  Returns the patch as it is
    loaded into a context of the class geometry, the `shape` mfbd_set_patch_shape takes for
    it, and the same patch as a problem of its own geometry (the reference runs a patch of
    another mode or frame count as a job of its own)."""
    cfg = su["cfg"]
    nO, F, K = cfg.get("n_objects", 1), cfg["n_frames"], su["K"]
    if not 1 <= n_modes <= K or len(frames) != nO or not all(1 <= f <= F for f in frames):
        raise ValueError("ragged patch outside the class geometry")
    big = build_patch(su, seed, field, field_offset)
    img_on = np.zeros(nO * F, dtype=np.uint8)
    sel = []
    for o in range(nO):                       # object o's images are o*F .. o*F+F-1 (one channel each)
        img_on[o * F:o * F + frames[o]] = 1
        sel += list(range(o * F, o * F + frames[o]))
    sel = np.asarray(sel)
    S = _setup_default()
    wfs = [[list(range(frames[o]))] for o in range(nO)]
    nPar, nFree, rows, cols, vals = S.constraints([[frames[o]] for o in range(nO)], wfs, n_modes)
    shape = dict(nModes=n_modes, imgOn=img_on, nFree=nFree,
                 rows=(sel[rows // n_modes] * K + rows % n_modes).astype(np.int32), cols=cols, vals=vals)
    small = dict(big)
    small["K"] = n_modes
    small["objects"] = [dict(o, modes=o["modes"][:n_modes]) for o in big["objects"]]
    small["imgCh"] = big["imgCh"][sel]
    small["data"] = big["data"][sel]
    small["nullspace"] = dict(nPar=nPar, nFree=nFree, rows=rows, cols=cols, vals=vals)
    small["truth"] = dict(alpha=big["truth"]["alpha"][sel][:, :n_modes])
    return big, shape, small


_UNEVEN = {}


def _uneven_item(d):
    c, m, fr, sd = d
    return ragged_patch(_UNEVEN["setups"][c], sd, m, fr)


def uneven_batch(classes, n_patches, seed=2024, n_modes=(36, 100), n_frames=(12, 48), workers=1):
    """An uneven batch: n_patches drawn from `seed`, each of a grid class picked at random
    from `classes` (configurations whose modes and n_frames are the class maxima), with a
    mode count in n_modes and a frame count per object in n_frames (inclusive ranges,
    clipped to the class).  Returns per class (setup, patches, shapes, own-geometry
    problems, indices into the batch)."""
    rng = np.random.default_rng(seed)
    sus = [build_setup(c) for c in classes]
    draws = []
    for k in range(n_patches):
        c = int(rng.integers(len(classes)))
        cfg, K = classes[c], sus[c]["K"]
        m = int(rng.integers(min(n_modes[0], K), min(n_modes[1], K) + 1))
        F = cfg["n_frames"]
        fr = [int(rng.integers(min(n_frames[0], F), min(n_frames[1], F) + 1)) for _ in range(cfg.get("n_objects", 1))]
        draws.append((c, m, fr, seed * 7 + k))
    if workers > 1:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        _UNEVEN["setups"] = sus
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
            made = list(ex.map(_uneven_item, draws, chunksize=4))
        _UNEVEN.clear()
    else:
        made = [ragged_patch(sus[c], sd, m, fr) for c, m, fr, sd in draws]
    out = []
    for c in range(len(classes)):
        idx = [k for k, d in enumerate(draws) if d[0] == c]
        items = [made[k] for k in idx]
        out.append(dict(setup=sus[c], patches=[b for b, _, _ in items], shapes=[sh for _, sh, _ in items],
                        own=[sm for _, _, sm in items], index=idx))
    return out


# Configurations (sizes follow the benchmark definitions; cfg2 keeps N = 128 with a
# 64-pixel pupil because the reference ties N = 2P and a 128 grid cannot hold a 100-px
# pupil without aliasing the OTF).
CFG1 = dict(s=64, n_frames=8, n_objects=1, modes=list(range(2, 22)), basis="zernike", d_r0=10.0,
            mean_counts=1000.0, margin=5, settings=dict(FTOL=1e-6))
CFG2_PATCH = dict(s=88, n_frames=24, n_objects=2, modes=list(range(2, 52)), basis="kl", d_r0=20.0,
                  mean_counts=2000.0, diversity=[0.0, math.pi / math.sqrt(3.0)], margin=5,
                  pupil_diameter=64.0, settings=dict(FTOL=1e-3))
# Config 5 patch: 128 x 128 on a 256 x 256 OTF grid (128-pixel pupil), 150 KL modes,
# 2 objects x 100 frames (D/r0 and the diversity as in config 2).
CFG5_PATCH = dict(s=128, n_frames=100, n_objects=2, modes=list(range(2, 152)), basis="kl", d_r0=20.0,
                  mean_counts=2000.0, diversity=[0.0, math.pi / math.sqrt(3.0)], margin=5,
                  pupil_diameter=128.0, settings=dict(FTOL=1e-3))
# Config 3 (uneven batch): grid classes 64 / 128 / 256 (patches 64, 88, 128 with 32-, 64-
# and 128-pixel pupils), each with the largest counts a patch may draw: 100 KL modes and
# 48 frames per object; patches draw 36-100 modes and 12-48 frames per object.
UNEVEN_CLASSES = [
    dict(s=64, n_frames=48, n_objects=2, modes=list(range(2, 102)), basis="kl", d_r0=20.0, mean_counts=2000.0,
         diversity=[0.0, math.pi / math.sqrt(3.0)], margin=5, pupil_diameter=32.0, settings=dict(FTOL=1e-3)),
    dict(CFG2_PATCH, n_frames=48, modes=list(range(2, 102))),
    dict(CFG5_PATCH, n_frames=48, modes=list(range(2, 102))),
]
