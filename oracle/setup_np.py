"""Oracle (test infrastructure only): numpy restatement of the per-job setup.

Restates, in float64/longdouble numpy:
  * Grid + makePupil + Pupil::normalize/generateSupport
      (src/lib/image/grid.cpp:45-67, src/lib/image/utils.cpp:68-121,
       src/lib/image/pupil.cpp:221-299)
  * Pupil::calculatePupilSize (pupil.cpp:78-100)
  * Zernike radial/angular/normalisation (src/lib/image/zernike.cpp:151-274, 290-364)
  * Noll covariance (zernike.cpp:102-148) and the KL expansion
      (zernike.cpp:367-479, src/lib/momfbd/modes.cpp:158-205)
  * ModeSet getNorms/normalize/measureJacobian (modes.cpp:443-535)
  * Constraints + nullspace (src/lib/momfbd/constraints.cpp:85-198, 302-356,
      429-469, 536-622)
Used by tests/ as the checker and by bench.py's cpu_baseline leg only.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import gammasgn

LD = np.longdouble


# ----------------------------------------------------------------- pupil -----
def grid(n: int, origin: float):
    """Grid::init: distance and angle of pixel centres from origin-0.5 (grid.cpp:45-67)."""
    o = LD(np.float32(origin)) - LD(0.5)
    y = np.arange(n, dtype=LD)[:, None] - o
    x = np.arange(n, dtype=LD)[None, :] - o
    d = np.sqrt(y * y + x * x)
    a = np.arctan2(y, x)
    zero = (y == 0) & (x == 0)
    d = np.where(zero, LD(0), d)
    a = np.where(zero, LD(0), a)
    return d, a


def make_pupil(n: int, radius: float, inner: float = 0.0):
    """makePupil (utils.cpp:68-121): anti-aliased disk centred on pixel (n/2, n/2)."""
    D, ang = grid(n, n / 2.0 + 0.5)
    absang = np.abs(ang)
    pi = LD(math.pi)
    with np.errstate(divide="ignore"):
        d = np.where((absang > pi / 4) & (absang < 3 * pi / 4), np.abs(1 / np.sin(absang)),
                     np.abs(1 / np.cos(absang)))
    out = np.zeros((n, n), dtype=LD)
    area = LD(0)
    has_inner = inner > 0
    for y in range(n):
        for x in range(n):
            dd = d[y, x]
            dir_ = D[y, x] - LD(inner)
            if has_inner and dir_ < -dd:
                continue
            dor = D[y, x] - LD(radius)
            if dor > dd:
                continue
            if dor < -dd and (not has_inner or dir_ > dd):
                out[y, x] = 1
                area += 1
                continue
            R = dor
            d2 = dd * 2
            if has_inner and dor < -d2:
                R = dir_
                d2 = -d2
            val = LD(0.5) - R / d2
            out[y, x] = val
            area += val
    return out.astype(np.float64), float(area)


def pupil_normalize(p):
    mn, mx = p.min(), p.max()
    if mn == mx:
        return p
    return (p - mn) * (1.0 / (mx - mn))


def autocorr_real_centered(a):
    """static FourierTransform::autocorrelate(Array<double>&, center) (fouriertransform.hpp:291-297)."""
    n = a.shape[0]
    F = np.fft.fft2(a)
    r = np.real(np.fft.ifft2(np.abs(F) ** 2))
    return np.fft.fftshift(r) if n % 2 == 0 else r


def pupil_support(pupil, threshold=1e-9):
    """Pupil::generateSupport (pupil.cpp:245-277)."""
    P = pupil.shape[0]
    N = 2 * P
    flat = pupil.ravel()
    ps = np.nonzero(flat > threshold)[0]
    area = float(np.sum(flat[ps]))
    py, px = ps // P, ps % P
    pio = py * N + px + (P // 2) + (P // 2) * N
    otf = np.zeros((N, N))
    otf[:P, :P] = pupil
    ac = autocorr_real_centered(otf)
    os_ = np.nonzero(np.abs(ac.ravel()) > threshold)[0]
    return ps.astype(np.int32), pio.astype(np.int32), os_.astype(np.int32), area


def generate_pupil(P, radius, co_radius=0.0):
    """Pupil::generate (pupil.cpp:221-241)."""
    p, _ = make_pupil(P, radius, co_radius)
    p = pupil_normalize(p)
    ps, pio, os_, area = pupil_support(p)
    return dict(pupil=p, ps=ps, pio=pio, os=os_, area=area)


def calculate_pupil_size(wavelength, nPixels, telescopeD, arcSecsPerPixel):
    """Pupil::calculatePupilSize (pupil.cpp:78-100): (frequencyCutoff, radius, pupilPixels)."""
    goodsizes = [16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 81, 90,
                 96, 100, 108, 120, 125, 128, 135, 144]
    rad_per_pix = arcSecsPerPixel * (math.pi / (180.0 * 3600.0))
    q = wavelength / (rad_per_pix * telescopeD)
    fc = nPixels / q
    npp = nPixels >> 2
    r = fc / 2.0
    if npp < r:
        for gs in goodsizes:
            npp = max(gs, npp)
            if npp >= r:
                break
    return fc, r, npp << 1


# --------------------------------------------------------------- zernike -----
def noll_to_nm(j):
    """Zernike::NollToNM (zernike.cpp:69-80)."""
    n = 0
    ln = 1
    i = 1
    while ln < j:
        n = i
        ln += n + 1
        i += 1
    dl = n + 1 - ln + j
    m = 2 * ((dl + (n % 2)) // 2) + (0 if n % 2 else 1) - 1
    if j % 2:
        m = -m
    return n, m


def radial_poly(n, m, r):
    """Zernike radial polynomial R_n^m(r) (Noll 1976), closed form, longdouble."""
    out = np.zeros_like(r, dtype=LD)
    for s in range((n - m) // 2 + 1):
        c = (-1) ** s * math.factorial(n - s) // (
            math.factorial(s) * math.factorial((n + m) // 2 - s) * math.factorial((n - m) // 2 - s))
        out += LD(c) * r ** (n - 2 * s)
    return out


_PUPIL_CACHE = {}


def fetch_pupil(P, radius):
    key = (P, float(np.float32(radius)))
    if key not in _PUPIL_CACHE:
        _PUPIL_CACHE[key] = generate_pupil(P, radius)
    return _PUPIL_CACHE[key]


def zernike(j, P, radius, angle=0.0, normalize=True):
    """Zernike::getZernike (zernike.cpp:234-274), Noll index j >= 2."""
    n, m = noll_to_nm(j)
    rf = LD(np.float32(radius))
    d, a = grid(P, P / 2.0 + 0.5)
    rr = np.minimum(LD(1.0), (LD(1.0) / rf) * d).astype(np.float64).astype(LD)
    rad = radial_poly(n, abs(m), rr).astype(np.float64).astype(LD)
    angr = LD(np.float32(angle)) * LD(math.pi) / LD(180.0)
    if m < 0:
        ang = np.sin(abs(m) * (a - angr))
    else:
        ang = np.cos(m * (a - angr))
    z = rad * ang.astype(np.float64).astype(LD)
    if normalize:
        pup = fetch_pupil(P, radius)
        nrm = np.sum(z * z * pup["pupil"].astype(LD))
        nrm = np.sqrt(LD(pup["area"]) / nrm)
    else:
        nrm = np.sqrt(LD(2 * (n + 1))) if m else np.sqrt(LD(n + 1))
    return (z * nrm).astype(np.float64)


def noll_cov_nm(n1, m1, n2, m2):
    """Zernike::calcCovariance(n1,m1,n2,m2) (zernike.cpp:102-118), Noll (1976)."""
    if m1 != m2:
        return 0.0
    k = (4.8 * math.exp(math.lgamma(6.0 / 5.0))) ** (5.0 / 6.0) * math.exp(
        math.lgamma(14.0 / 3.0) + 2.0 * math.lgamma(11.0 / 6.0)) / (2.0 ** (8.0 / 3.0) * math.pi)
    k *= (-1.0) ** ((n1 + n2 - 2 * m1) / 2.0) * math.sqrt((n1 + 1) * (n2 + 1))
    args = [((n1 + n2) - 5.0 / 3.0) / 2.0, ((n1 - n2) + 17.0 / 3.0) / 2.0,
            ((n2 - n1) + 17.0 / 3.0) / 2.0, ((n1 + n2) + 23.0 / 3.0) / 2.0]
    lg = [math.lgamma(v) for v in args]
    sg = [float(gammasgn(v)) for v in args]
    return k * math.exp(lg[0] - lg[1] - lg[2] - lg[3]) * sg[0] * sg[1] * sg[2] * sg[3]


def noll_cov(i, j):
    """Zernike::calcCovariance(i,j) on Noll indices (zernike.cpp:121-134)."""
    if i < 2 or j < 2:
        return 0.0
    if i > j:
        i, j = j, i
    ni, mi = noll_to_nm(i)
    if mi and (i + j) % 2:
        return 0.0
    nj, mj = noll_to_nm(j)
    return noll_cov_nm(ni, abs(mi), nj, abs(mj))


def kl_expansion(first, last):
    """Zernike::karhunenLoeveExpansion (zernike.cpp:367-479).

    Returns {kl_index: (covariance, [(zernike_index, weight), ...])}.  The SVD of each
    symmetric positive block is taken as an eigen-decomposition sorted by decreasing
    eigenvalue (GSL's SVD returns the same up to column signs, which are fixed by the
    reference's largest-coefficient rule; order inside degenerate blocks is unpinned)."""
    mapping = {i: i for i in range(first, last + 1)}
    for i in range(first, last):
        cov1 = noll_cov(mapping[i], mapping[last])
        for j in range(last - 1, i, -1):
            cov2 = noll_cov(mapping[i], mapping[j])
            if (not cov2) and cov1:
                mapping[j], mapping[j + 1] = mapping[j + 1], mapping[j]
            else:
                cov1 = cov2
    reverse = {v: k for k, v in mapping.items()}
    values = {}
    n_blocks = 1
    keys = sorted(mapping)
    for prev, cur in zip(keys[:-1], keys[1:]):
        values[cur] = noll_cov(mapping[prev], mapping[cur])
        if values[cur] == 0:
            n_blocks += 1
    blocks = []
    n = first
    for _ in range(n_blocks):
        fb = n
        n += 1
        while n <= last and values[n]:
            n += 1
        lb = n - 1
        sz = lb - fb + 1
        M = np.zeros((sz, sz))
        for a in range(sz):
            for b in range(sz):
                M[a, b] = noll_cov(mapping[fb + a], mapping[fb + b])
        if sz > 1:
            w, U = np.linalg.eigh(M)
            order = np.argsort(-w, kind="stable")
            w, U = w[order], U[:, order]
        else:
            w, U = np.array([M[0, 0]]), np.ones((1, 1))
        blocks.append((fb, lb, w, U))
    out = {}
    for i in range(first, last + 1):
        im = reverse[i]
        for fb, lb, w, U in blocks:
            if fb <= im <= lb:
                break
        col = im - fb
        cov = w[col]
        c = U[:, col]
        sign = 1
        largest = 0.0
        for v in c:
            if abs(v) > largest:
                largest = abs(v)
                sign = -1 if v < 0 else 1
        weights = [(mapping[m + fb], sign * c[m]) for m in range(lb - fb + 1)]
        out[i] = (cov, weights)
    return out


def kl_mode(kl, j, P, radius, angle=0.0, cutoff=1e-3):
    """PupilMode KL constructor (modes.cpp:158-205)."""
    cov, weights = kl[j]
    mode = np.zeros((P, P))
    for zj, c in weights:
        if abs(c) >= cutoff:
            mode = mode + zernike(zj, P, radius, angle) * c
    pup = fetch_pupil(P, radius)
    p = pup["pupil"]
    msk = p > 0
    nrm = np.sum((mode[msk].astype(LD) ** 2) * p[msk].astype(LD))
    nrm = np.sqrt(LD(pup["area"]) / nrm)
    return (mode.astype(LD) * nrm).astype(np.float64), math.sqrt(cov)


def mode_set(P, radius, modes, basis="zernike", kl_min=2, kl_max=2000, kl_cutoff=1e-3, angle=0.0):
    """ModeSet::generate (modes.cpp:348-439) for a list of Noll/KL indices.
    Returns (B[K,P,P] unit-RMS, atm_rms[K], tilt (x,y))."""
    B = np.zeros((len(modes), P, P))
    rms = np.zeros(len(modes))
    tilt = [-1, -1]
    kl = kl_expansion(kl_min, kl_max) if basis == "kl" else None
    for k, j in enumerate(modes):
        if j == 2 or j == 3 or basis == "zernike":
            if j == 2:
                tilt[0] = k
            elif j == 3:
                tilt[1] = k
            B[k] = zernike(j, P, radius, angle) if j >= 2 else float(j)
            rms[k] = math.sqrt(noll_cov(j, j))
        else:
            B[k], rms[k] = kl_mode(kl, j, P, radius, angle, kl_cutoff)
    return B, rms, tuple(tilt)


def get_norms(B, pupil, area):
    """ModeSet::getNorms (modes.cpp:443-462)."""
    m = pupil > 0
    return np.array([math.sqrt(float(np.sum(b[m] * pupil[m] * b[m])) / area) for b in B])


def normalize_modes(B, pupil, area, scale=1.0):
    """ModeSet::getNorms + normalize(scale) (modes.cpp:524-535)."""
    norms = get_norms(B, pupil, area)
    return B * (scale / norms)[:, None, None]


def fit_plane(data, mask):
    """fitPlane (utils.cpp:169-234): z = a*y + b*x + c over mask>0; returns (a,b,c,chisq/n)."""
    ys, xs = data.shape
    idx = np.nonzero(mask.ravel() > 0)[0]
    zy = (idx // xs).astype(float) - ys // 2
    zx = (idx % xs).astype(float) - xs // 2
    X = np.stack([zy, zx, np.ones_like(zy)], axis=1)
    d = data.ravel()[idx]
    coef, *_ = np.linalg.lstsq(X, d, rcond=None)
    res = d - X @ coef
    return coef[0], coef[1], coef[2], float(res @ res) / len(idx)


def measure_jacobian(B, pupil, scale):
    """ModeSet::measureJacobian (modes.cpp:464-512) -> (tilt (x,y), J_T_xy 2x2, J_xy_T 2x2)."""
    fits = []
    for i, b in enumerate(B):
        a, bx, c, chi = fit_plane(b, pupil)
        fits.append(((chi, a, bx, c), i))
    fits.sort(key=lambda t: t[0])
    J = np.zeros((2, 2))
    chisq = [np.finfo(float).max] * 3
    tilt = [-1, -1]
    cnt = 0
    for coeffs, i in fits:
        chisq[cnt] = coeffs[0]
        if cnt == 2:
            break
        J[0, cnt] = scale * coeffs[2]
        J[1, cnt] = scale * coeffs[1]
        tilt[cnt] = i
        cnt += 1
    if 100 * chisq[0] < chisq[2] and 100 * chisq[1] < chisq[2]:
        if abs(J[0, 0]) / abs(J[1, 0]) < 1:
            tilt = [tilt[1], tilt[0]]
            J[0, 0], J[0, 1] = J[0, 1], J[0, 0]
            J[1, 0], J[1, 1] = J[1, 1], J[1, 0]
    det = J[0, 0] * J[1, 1] - J[0, 1] * J[1, 0]
    inv = 1.0 / det
    Ji = np.array([[J[1, 1] * inv, -J[1, 0] * inv], [-J[0, 1] * inv, J[0, 0] * inv]])
    return tuple(tilt), J, Ji


def pix2cf(arcsecperpix, telescope_d):
    """util::pix2cf (src/lib/momfbd/util.cpp:126-129)."""
    return (math.pi * math.pi) / 648000.0 * (arcsecperpix * telescope_d)


# ------------------------------------------------------------ nullspace -----
def _haar_nullspace(C):
    """NullSpace::calculateNullspace, new-style branch (constraints.cpp:92-176)."""
    nC, nP = C.shape
    nF = nP - nC
    if nF < 1:
        return np.zeros((nP, 0))
    if nF == 1:
        return np.full((nP, 1), 1.0 / math.sqrt(nP))
    ns = np.zeros((nF, nP))
    counts = [(int(np.count_nonzero(C[i])), i) for i in range(nC)]
    row_index, maxc = 0, 0
    for c, i in counts:
        if c > maxc:
            maxc, row_index = c, i
    bl = 1
    basis = 0
    while bl < maxc:
        off = 0
        while off + bl < maxc:
            rem = max(min(maxc - off - bl, bl), 0)
            for i in range(bl):
                ns[basis, off + i] = 1.0
                if off + bl + i < maxc:
                    ns[basis, off + bl + i] = -bl * 1.0 / rem
            basis += 1
            off += 2 * bl
        bl *= 2
    for i in range(nF):
        for j in range(maxc):
            v = ns[i, j]
            if abs(v) > 0:
                for k in range(nC):
                    v1 = C[k, j]
                    if k != row_index and v1:
                        for p in range(nP):
                            if p == j:
                                continue
                            v2 = C[k, p]
                            if v2:
                                ns[i, p] = -v1 * v / v2
    for i in range(nF):
        s = float(np.sum(ns[i] * ns[i]))
        for j in range(i + 1, nF):
            sp = float(np.sum(ns[i] * ns[j]))
            if abs(sp) > 0:
                sp /= s
                ns[j] = ns[j] - sp * ns[i]
        if s > 0 and s != 1.0:
            ns[i] = ns[i] * (1.0 / math.sqrt(s))
    return ns.T


def constraints(n_images_per_obj_ch, wavefronts, K):
    """Constraints::init (constraints.cpp:536-622) + grouping/blockify/nullspace.

    n_images_per_obj_ch: list over objects of lists over channels of frame counts.
    wavefronts: same nesting, list of wavefront ids per channel.
    Returns (nPar, nFree, rows, cols, vals) with entries sorted by (row, col)."""
    nImg = sum(sum(c) for c in n_images_per_obj_ch)
    nPar = nImg * K
    cons = []
    for mi in range(K):
        if mi in (0, 1):
            n0 = n_images_per_obj_ch[0][0]
            cons.append({ic * K + mi: 1 for ic in range(n0)})
            wf_seen = {}
            off = 0
            for obj in wavefronts:
                for ch in obj:
                    for wf in ch:
                        if wf in wf_seen:
                            c = dict(wf_seen[wf])
                            c[off + mi] = -1
                            cons.append(c)
                        else:
                            wf_seen[wf] = {off + mi: 1}
                        off += K
        else:
            wf_seen = {}
            count = 0
            for obj in wavefronts:
                for ch in obj:
                    for wf in ch:
                        off = count * K
                        if wf in wf_seen:
                            c = dict(wf_seen[wf])
                            c[off + mi] = -1
                            cons.append(c)
                        else:
                            wf_seen[wf] = {off + mi: 1}
                        count += 1
    nFree = nPar - len(cons)
    # groupConnectedVariables (constraints.cpp:472-482)
    tmp = list(cons)
    groups = []
    while tmp:
        g = [tmp[0]]
        idx = set(tmp[0])
        tmp = tmp[1:]
        changed = True
        while changed:
            changed = False
            for k, c in enumerate(tmp):
                if idx & set(c):
                    g.append(c)
                    idx |= set(c)
                    del tmp[k]
                    changed = True
                    break
        groups.append((g, idx))
    entries = {}
    used_rows, used_cols = set(), set()
    rOff = cOff = 0
    for g, idx in groups:
        # sortRows: descending by (first index ascending, then more entries, then larger value)
        def key(c):
            items = sorted(c.items())
            return (items[0][0], -len(items), [-v for _, v in items])
        g = sorted(g, key=key)
        counts = sorted(((sum(1 for c in g if i in c), i) for i in idx), key=lambda t: (-t[0], t[1]))
        order = [i for _, i in counts]
        local = {orig: k for k, orig in enumerate(order)}
        nP = len(order)
        C = np.zeros((len(g), nP))
        for r, c in enumerate(g):
            for i, v in c.items():
                C[r, local[i]] = v
        ns = _haar_nullspace(C)
        for r in range(ns.shape[0]):
            for c in range(ns.shape[1]):
                v = ns[r, c]
                if abs(v) > 1e-10:
                    entries[(order[r], cOff + c)] = v
                    used_rows.add(order[r])
                    used_cols.add(cOff + c)
        rOff += nP
        cOff += nP - len(g)
    unused1 = [i for i in range(nPar) if i not in used_rows]
    unused2 = [i for i in range(nFree) if i not in used_cols]
    assert len(unused1) == len(unused2)
    for r, c in zip(unused1, unused2):
        entries[(r, c)] = 1.0
    keys = sorted(entries)
    rows = np.array([k[0] for k in keys], dtype=np.int32)
    cols = np.array([k[1] for k in keys], dtype=np.int32)
    vals = np.array([entries[k] for k in keys], dtype=np.float64)
    return nPar, nFree, rows, cols, vals
