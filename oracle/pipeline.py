"""Oracle: the per-pattern indexing path in float64 (shaders/sphere.wgsl, so3.wgsl, detector.wgsl,
prep.wgsl, mastersph.wgsl; src/spherical.zig, src/detector.zig, src/engine.zig).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
import numpy as np

from . import sh as osh

TAU = 2 * np.pi


# ── geometry and the master on the sphere ──

def frame(geo_R, h, w, mask_on=True, mask_radius=1.0):
    """Columns of R and the circular mask (src/engine.zig:1444-1448, 1567-1579)."""
    R = np.asarray(geo_R, np.float64)
    if mask_on:
        r = mask_radius * 0.5 * min(h, w)
        mask = np.array([0.5 * (h - 1), 0.5 * (w - 1), r * r], np.float32).astype(np.float64)
    else:
        mask = np.zeros(3)
    return {"c0": R[:, 0], "c1": R[:, 1], "c2": R[:, 2], "h": h, "w": w, "mask": mask}


def bruker_R(sample_tilt_deg=70.0, detector_tilt_deg=0.0):
    """R0 of the Bruker construction (src/geometry.zig:56-62), without detector rotation."""
    a = np.radians(detector_tilt_deg - sample_tilt_deg)
    s, c = np.sin(a), np.cos(a)
    return np.array([[-s, 0, c], [0, 1, 0], [-c, 0, -s]])


def translation(R, pc, rows, cols):
    """t = R (0.5 - pcy rows, 0.5 - pcx cols, pcz rows) (src/geometry.zig:71-75)."""
    return R @ np.array([0.5 - pc[1] * rows, 0.5 - pc[0] * cols, pc[2] * rows])


def sphere_to_square(v):
    """(common.wgsl:44-57) in float64."""
    v = np.asarray(v, np.float64)
    ax, ay = np.abs(v[..., 0]), np.abs(v[..., 1])
    xn = np.sqrt(np.maximum(np.pi * 0.5 * (ax * ax + ay * ay) / (1 + np.abs(v[..., 2])), 0))
    big = ax >= ay
    u = np.where(big, xn, (4 * xn / np.pi) * np.arctan2(ax, np.maximum(ay, 1e-30)))
    w = np.where(big, (4 * xn / np.pi) * np.arctan2(ay, np.maximum(ax, 1e-30)), xn)
    sc = 1 / np.sqrt(np.pi / 2)
    return u * sc * np.where(v[..., 0] < 0, -1, 1), w * sc * np.where(v[..., 1] < 0, -1, 1)


def square_to_sphere(x, y, south):
    """(mastersph.wgsl:37-56) in float64."""
    sc = np.sqrt(np.pi / 2)
    xa, ya = np.abs(x) * sc, np.abs(y) * sc
    swap = xa >= ya
    xn, yn = np.maximum(xa, ya), np.minimum(xa, ya)
    safe = np.where(xn > 0, xn, 1)
    r = (2 * safe / np.pi) * np.sqrt(np.maximum(np.pi - safe * safe, 0))
    ang = np.pi * yn / (4 * safe)
    hx, hy = r * np.cos(ang), r * np.sin(ang)
    hx, hy = np.where(swap, hx, hy), np.where(swap, hy, hx)
    v = np.stack([hx * np.where(x < 0, -1, 1), hy * np.where(y < 0, -1, 1), 1 - 2 * safe * safe / np.pi], -1)
    v = v / np.linalg.norm(v, axis=-1, keepdims=True)
    v = np.where((xn > 0)[..., None], v, np.array([0.0, 0.0, 1.0]))
    v[..., 2] = np.where(south, -v[..., 2], v[..., 2])
    return v


def sheet_at(sheets, side, v):
    """Bilinear lookup of [north | south] sheets (any trailing channels) in direction v, texel
    indices clamped as detector.wgsl master_at (102-115) does."""
    v = np.asarray(v, np.float64)
    S = np.asarray(sheets, np.float64).reshape(2 * side * side, -1)
    u, w = sphere_to_square(v)
    last = side - 1.0
    fj = np.clip((u + 1) * 0.5 * last, 0, last)
    fi = np.clip((w + 1) * 0.5 * last, 0, last)
    c0 = np.minimum(fj.astype(np.int64), side - 2)
    r0 = np.minimum(fi.astype(np.int64), side - 2)
    o = r0 * side + c0 + np.where(v[..., 2] < 0, side * side, 0)
    tj = (fj - c0)[..., None]
    ti = (fi - r0)[..., None]
    top = S[o] * (1 - tj) + S[o + 1] * tj
    bot = S[o + side] * (1 - tj) + S[o + side + 1] * tj
    out = top * (1 - ti) + bot * ti
    return out[..., 0] if S.shape[1] == 1 else out


def texel_dirs(side):
    tid = np.arange(2 * side * side)
    pix = tid % (side * side)
    last = side - 1.0
    return square_to_sphere(2 * (pix % side) / last - 1, 2 * (pix // side) / last - 1, tid >= side * side)


def tangents(d):
    ref = np.where((np.abs(d[..., 0]) > 0.8)[..., None], np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0]))
    e1 = np.cross(d, ref)
    e1 /= np.linalg.norm(e1, axis=-1, keepdims=True)
    return e1, np.cross(d, e1)


def blur(sheets, side, sigma_rad, texels=None):
    """One sphere_blur pass (mastersph.wgsl:87-109): the weighted mean over a 9 x 9 lattice of
    step 0.6 sigma in each texel's tangent plane. `texels`: only these (all by default)."""
    d = texel_dirs(side)
    if texels is not None:
        d = d[texels]
    e1, e2 = tangents(d)
    step = 0.6 * sigma_rad
    k = 1 / (2 * sigma_rad ** 2)
    acc = np.zeros(len(d))
    ws = 0.0
    for i in range(-4, 5):
        for j in range(-4, 5):
            a, b = i * step, j * step
            wgt = np.exp(-(a * a + b * b) * k)
            v = d + a * e1 + b * e2
            acc += wgt * sheet_at(sheets, side, v / np.linalg.norm(v, axis=-1, keepdims=True))
            ws += wgt
    return acc / ws


def gradient(sheets, side, texels=None):
    """(grad M, M) per texel by central differences one texel apart (mastersph.wgsl:124-138)."""
    c = texel_dirs(side)
    idx = np.arange(len(c)) if texels is None else np.asarray(texels)
    c = c[idx]
    e1, e2 = tangents(c)
    h = np.sqrt(2 * np.pi) / side

    def at(v):
        return sheet_at(sheets, side, v / np.linalg.norm(v, axis=-1, keepdims=True))

    g1 = (at(c + h * e1) - at(c - h * e1)) / (2 * h)
    g2 = (at(c + h * e2) - at(c - h * e2)) / (2 * h)
    g = g1[:, None] * e1 + g2[:, None] * e2
    return np.concatenate([g, np.asarray(sheets, np.float64).reshape(-1)[idx][:, None]], 1)


# ── patterns ──

def in_mask(rows, cols, mask):
    dr, dc = rows - mask[0], cols - mask[1]
    return (mask[2] <= 0) | (dr * dr + dc * dc <= mask[2])


def normalize(pat, mask):
    """Zero mean and unit norm under the mask, zero outside (prep.wgsl:405-443)."""
    h, w = pat.shape
    r, c = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    m = in_mask(r, c, mask)
    v = np.where(m, pat - pat[m].mean(), 0.0).astype(np.float64)
    return v / np.sqrt(max((v * v).sum(), 1e-30))


# ── pattern filters before the path (prep.wgsl:72-287; src/prep.zig:105-204) ──

def lattice(rows, cols, want):
    """Cells spread over the map, row-major (src/engine.zig:1325-1341)."""
    gy, gx = rows, cols
    if rows * cols > want:
        ratio = rows / cols
        gy = int(min(max(int(np.floor(np.sqrt(want * ratio) + 0.5)), 1), rows))
        gx = int(min(max(want // gy, 1), cols))
    return np.array([((2 * r + 1) * rows // (2 * gy)) * cols + (2 * c + 1) * cols // (2 * gx) for r in range(gy) for c in range(gx)],
                    np.int64)


def static_average(patterns, rows, cols, want):
    """The mean pattern over the lattice cells (src/engine.zig:1344-1385), float64 summed."""
    cells = lattice(rows, cols, want)
    flat = np.asarray(patterns).reshape(rows * cols, -1).astype(np.float64)
    return flat[cells].mean(0).reshape(np.asarray(patterns).shape[-2:])


def _gauss_rows_direct(img, sigma):
    """blur_rows: taps exp(-d^2 / 2 sigma^2) out to max(1, ceil(3 sigma)), renormalized to the
    taps inside the image (prep.wgsl:92-135)."""
    h, w = img.shape
    r = max(1, int(np.ceil(3 * sigma)))
    k = 1.0 / (2 * sigma * sigma)
    out = np.empty_like(img)
    for c in range(w):
        x = np.arange(max(c - r, 0), min(c + r, w - 1) + 1)
        g = np.exp(-((x - c) ** 2) * k)
        out[:, c] = (img[:, x] * g).sum(1) / g.sum()
    return out


def blur_direct(img, sigma):
    """The separable Gaussian of the narrow path: rows, then columns."""
    return _gauss_rows_direct(_gauss_rows_direct(np.asarray(img, np.float64), sigma).T, sigma).T


def _cell(j, n, s):
    """Centre (pixels) and width of coarse cell j along an axis of n pixels (prep.wgsl:179-185)."""
    lo = j * s
    hi = lo + s
    if j >= 0 and lo < n:
        hi = min(hi, n)
    return 0.5 * (lo + hi - 1), float(hi - lo)


def blur_box(img, sigma):
    """The wide path (sigma >= 12 pixels, prep.wgsl:166-287; src/prep.zig:127-146): pixel sums over
    s x s boxes (s = floor(sigma / 6)) with their counts, a Gaussian over the coarse cells (with a
    margin of 3 whole cells outside the image), spread back over the pixels by a Gaussian of
    deviation s weighted by the cells' widths; sums and counts divided at the end."""
    img = np.asarray(img, np.float64)
    h, w = img.shape
    s = int(np.floor(sigma / 6))
    cs = np.sqrt(sigma * sigma - s * s - (s * s - 1) / 12.0)
    pad = 3
    R = float(np.ceil(4 * cs))
    k0 = 1.0 / (2 * cs * cs)
    k2 = 1.0 / (2 * s * s)
    hs, ws = -(-h // s), -(-w // s)
    box = np.zeros((hs, ws, 2))
    for i in range(hs):
        for j in range(ws):
            blk = img[i * s:min((i + 1) * s, h), j * s:min((j + 1) * s, w)]
            box[i, j] = (blk.sum(), blk.size)
    wp, hp = ws + 2 * pad, hs + 2 * pad

    def coarse(cin, n, m_out):
        # cin [lines][m_in] along an axis of n pixels -> [lines][m_out] over cells -pad .. m_out - pad - 1
        m_in = cin.shape[1]
        out = np.zeros((cin.shape[0], m_out, 2))
        for t in range(m_out):
            c, _ = _cell(t - pad, n, s)
            j0 = max(int(np.floor((c - R) / s)), 0)
            j1 = min(int(np.floor((c + R) / s)), m_in - 1)
            for j in range(j0, j1 + 1):
                d = _cell(j, n, s)[0] - c
                out[:, t] += np.exp(-d * d * k0) * cin[:, j]
        return out

    rows_ = coarse(box, w, wp)                                        # [hs][wp]
    cols_ = coarse(rows_.transpose(1, 0, 2), h, hp).transpose(1, 0, 2)  # [hp][wp]

    def spread(cin, n):
        # cin [lines][m] over cells -pad ..; -> [lines][n] pixels
        out = np.zeros((cin.shape[0], n, 2))
        for x in range(n):
            own = x // s
            for t in range(2 * pad + 1):
                cx, cwid = _cell(own + t - pad, n, s)
                d = x - cx
                out[:, x] += cwid * np.exp(-d * d * k2) * cin[:, own + t]
        return out

    sr = spread(cols_, w)                                             # [hp][w]
    sc = spread(sr.transpose(1, 0, 2), h).transpose(1, 0, 2)          # [h][w]
    return sc[..., 0] / sc[..., 1]


def gauss(img, sigma):
    """The reference's Gaussian of deviation sigma pixels: the box path from 12 pixels
    (src/prep.zig:105-140)."""
    return blur_box(img, sigma) if int(np.floor(sigma / 6)) >= 2 else blur_direct(img, sigma)


def stretch(img):
    """To [0, 1] by the image's extremes (prep.wgsl:289-315), in float32 as there."""
    x = np.asarray(img, np.float32)
    lo = x.min()
    k = np.float32(1.0) / np.maximum(x.max() - lo, np.float32(1e-20))
    return (x - lo) * k


def _reflect(i, n, pad):
    if n <= 1:
        return np.zeros_like(i)
    x = np.abs(i - pad)
    x = np.where(x >= n, 2 * n - 2 - x, x)
    return np.clip(x, 0, n - 1)


def clahe(img, tiles, bins, clip):
    """Contrast-limited adaptive histogram equalization of a [0, 1] image (prep.wgsl:317-397),
    in float32 in the same order of operations: the image padded by reflection at its top and
    left to whole tiles; each tile's histogram clipped at clip x the mean count, what was cut
    spread evenly (the remainder one count each to the first levels); every pixel read from the
    four nearest tiles' tables, each between its two nearest levels."""
    x = np.asarray(img, np.float32)
    h, w = x.shape
    f32 = np.float32
    pad_h, pad_w = (tiles - h % tiles) % tiles, (tiles - w % tiles) % tiles
    hp, wp = h + pad_h, w + pad_w
    th, tw = max(1, hp // tiles), max(1, wp // tiles)
    py, px = np.meshgrid(np.arange(hp), np.arange(wp), indexing="ij")
    v = np.clip(x[_reflect(py, h, pad_h), _reflect(px, w, pad_w)], f32(0), f32(1))
    b = np.clip(v * f32(bins - 1), f32(0), f32(bins - 1)).astype(np.int64)
    tile = (py // th) * tiles + px // tw
    hist = np.zeros((tiles * tiles, bins), np.int64)
    np.add.at(hist, (tile.ravel(), b.ravel()), 1)
    voxels = f32(max(th * tw, 1))
    nb = f32(bins)
    limit = np.maximum(f32(clip) * voxels / nb, f32(1))
    cdf = np.zeros((tiles * tiles, bins), np.float32)
    for t in range(tiles * tiles):
        cut = np.minimum(hist[t].astype(np.float32), limit)
        sumc = f32(0)
        for c in cut:
            sumc = f32(sumc + c)
        clipped = f32(voxels - sumc)
        residual = int(f32(clipped - nb * np.floor(clipped / nb)))
        redist = f32(f32(clipped - f32(residual)) / nb)
        acc = f32(0)
        for i in range(bins):
            c = f32(cut[i] + redist)
            if i < residual:
                c = f32(c + f32(1))
            acc = f32(acc + c)
            cdf[t, i] = np.clip(acc / voxels, f32(0), f32(1))
    r, c = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    v = np.clip(x, f32(0), f32(1))
    gy = (r.astype(np.float32) + f32(0.5)) / f32(h) * f32(tiles - 1)
    gx = (c.astype(np.float32) + f32(0.5)) / f32(w) * f32(tiles - 1)
    y0, x0 = np.floor(gy).astype(np.int64), np.floor(gx).astype(np.int64)
    y1, x1 = np.minimum(y0 + 1, tiles - 1), np.minimum(x0 + 1, tiles - 1)
    fy, fx = gy - y0.astype(np.float32), gx - x0.astype(np.float32)
    bf = v * f32(bins - 1)
    b0 = np.clip(np.floor(bf), f32(0), f32(bins - 2)).astype(np.int64)
    fb = bf - b0.astype(np.float32)
    mix = lambda a, b_, t: a * (f32(1) - t) + b_ * t  # noqa: E731

    def at(ty, tx):
        t = ty * tiles + tx
        return mix(cdf[t, b0], cdf[t, b0 + 1], fb)
    return mix(mix(at(y0, x0), at(y0, x1), fx), mix(at(y1, x0), at(y1, x1), fx), fy)


def prep(raw, mask, background=None, dynamic="none", dynamic_sigma=0.15, clahe_opts=None, bandpass=None, norm=True):
    """The filter chain then normalization, in the reference's order (src/prep.zig:148-204):
    the static background subtracted; the dynamic one (a Gaussian of dynamic_sigma x the width)
    subtracted or divided by; CLAHE (clahe_opts = (grid, bins, clip)) after stretching to
    [0, 1]; the band pass (bandpass = (fine, coarse) sigmas in pixels: the fine Gaussian, or the
    image itself when fine <= 0.05, minus the coarse one); then zero mean and unit norm under
    the mask."""
    img = np.asarray(raw, np.float64).copy()
    if background is not None:
        img = img - np.asarray(background, np.float64)
    if dynamic != "none":
        bl = gauss(img, dynamic_sigma * img.shape[1])
        img = img - bl if dynamic == "subtract" else img / np.maximum(bl, 1e-6)
    if clahe_opts is not None:
        img = clahe(stretch(img), *clahe_opts).astype(np.float64)
    if bandpass is not None and bandpass[1] > 0:
        fine = gauss(img, bandpass[0]) if bandpass[0] > 0.05 else img
        img = fine - gauss(img, bandpass[1])
    return normalize(img, mask) if norm else img


def project(sheets, side, fr, t, q):
    """Simulated pattern: the master at q . normalize(ray) (detector.wgsl:254-263)."""
    h, w = fr["h"], fr["w"]
    r, c = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64), indexing="ij")
    u = fr["c0"] * r[..., None] + fr["c1"] * c[..., None] + np.asarray(t)
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    return sheet_at(sheets, side, qrot(q, u))


# ── the sphere stage ──

def grid(L):
    n = osh.fft_size(L)
    x, w = osh.gl(n)
    return n, x, w


def sample_grid(pat, fr, t, L, window=False):
    """Value and window at every grid point (sphere.wgsl:67-101): [n_theta, n_phi] each."""
    n, x, _ = grid(L)
    n_phi = 2 * n
    st = np.sqrt(np.maximum(1 - x * x, 0))
    phi = TAU * np.arange(n_phi) / n_phi
    d = np.stack(np.broadcast_arrays(st[:, None] * np.cos(phi), st[:, None] * np.sin(phi), x[:, None]), -1)
    t = np.asarray(t, np.float64)
    nd = d @ fr["c2"]
    ok = nd > 1e-4
    s = np.where(ok, (fr["c2"] @ t) / np.where(ok, nd, 1), 0)
    hv = s[..., None] * d - t
    row = hv @ fr["c0"]
    col = hv @ fr["c1"]
    h, w = fr["h"], fr["w"]
    inb = ok & (row >= 0) & (row <= h - 1) & (col >= 0) & (col <= w - 1) & in_mask(row, col, fr["mask"])
    win = inb.astype(np.float64)
    if window:
        return win.copy(), win
    r0 = np.minimum(np.where(inb, row, 0).astype(np.int64), h - 2)
    q0 = np.minimum(np.where(inb, col, 0).astype(np.int64), w - 2)
    fr_ = np.where(inb, row, 0) - r0
    fc = np.where(inb, col, 0) - q0
    P = np.asarray(pat, np.float64)
    top = P[r0, q0] * (1 - fc) + P[r0, q0 + 1] * fc
    bot = P[r0 + 1, q0] * (1 - fc) + P[r0 + 1, q0 + 1] * fc
    val = np.where(inb, top * (1 - fr_) + bot * fr_, 0.0)
    return val, win


def coefficients(pat, fr, t, L, center=True, norm=True, window=False):
    """Harmonic coefficients of one item: window mean, centred windowed samples, analysis, unit
    norm (sphere.wgsl sphere_sums, rings, legendre, sh_norm). Returns (alm, stats)."""
    n, x, w = grid(L)
    n_phi = 2 * n
    val, win = sample_grid(pat, fr, t, L, window)
    wq = w[:, None]
    sv = (val * win * wq).sum()
    sw = (win * wq).sum()
    mean = sv / max(sw, 1e-20) if center else 0.0
    alm = osh.analyze((val - mean) * win, L, x, w)
    ms, _ = osh.pairs(L)
    nrm = np.sqrt(max(np.sum(np.where(ms == 0, 1, 2) * np.abs(alm) ** 2), 1e-30))
    if norm:
        alm = alm / nrm
    return alm, np.array([sw * TAU / n_phi, mean, nrm, alm[0].real])


# ── the master's harmonics ──

def harmonics(sheets, side, L):
    """f_lm of the master on a max(4L, 512)-ring grid, f_00 dropped, symmetry-forbidden terms
    dropped, and f2_lm of the band-limited master's square (src/master.zig:665-711).
    Returns (f, f2, fold, inversion)."""
    nt = max(4 * L, 512)
    x, w = osh.gl(nt)
    st = np.sqrt(np.maximum(1 - x * x, 0))
    phi = TAU * np.arange(2 * nt) / (2 * nt)
    d = np.stack(np.broadcast_arrays(st[:, None] * np.cos(phi), st[:, None] * np.sin(phi), x[:, None]), -1)
    samples = sheet_at(sheets, side, d.reshape(-1, 3)).reshape(nt, 2 * nt)
    f = osh.analyze(samples, L, x, w)
    f[0] = 0
    ms, ls = osh.pairs(L)
    e = np.where(ls == 0, 0, np.where(ms == 0, 1, 2) * np.abs(f) ** 2)
    tot = e.sum()
    inversion = e[ls % 2 == 1].sum() < 1e-5 * tot
    mirror = e[(ls + ms) % 2 == 1].sum() < 1e-5 * tot
    n_fold = 1
    for k in (6, 4, 3, 2):
        if e[ms % k != 0].sum() < 1e-5 * tot:
            n_fold = k
            break
    allow = (ms % n_fold == 0) & (~mirror | ((ls + ms) % 2 == 0)) & (~inversion | (ls % 2 == 0))
    f = np.where(allow, f, 0)
    x2, w2 = osh.gl(2 * L)
    s2 = osh.synthesize(f, L, x2, 4 * L)
    f2 = osh.analyze(s2 * s2, L, x2, w2)
    fold = {4: 4, 2: 2, 6: 2}.get(n_fold, 1)
    return f, f2, fold, inversion


# ── the SO(3) correlation ──

class So3:
    """The correlation's transform along b for one bandwidth and fold, as d^l_{m',m}(beta_j)
    (the reference transforms a contraction at pi/2 along k instead; the two are equal by
    d(beta) = i^{m'-m} sum_k d_{km'}(pi/2) d_{km}(pi/2) e^{-ik beta}, src/sh.zig:484-519)."""

    def __init__(self, L, fold, rows=None):
        """rows: the rows j of b to tabulate (all n/2 + 1 by default; volume needs all)."""
        self.L, self.fold = L, fold
        self.n = osh.fft_size(L)
        self.nc = self.n // fold
        self.nb = self.n // 2 + 1
        self.mu_max = (L - 1) // fold
        mus = np.arange(-self.mu_max, self.mu_max + 1)
        self.mus = mus
        # D[j, l, m', mu] = d^l_{m', fold mu}(beta_j), m' >= 0
        self.D = np.zeros((self.nb, L, L, len(mus)))
        for j in (range(self.nb) if rows is None else sorted(set(int(r) for r in rows))):
            beta = TAU * j / self.n
            for l in range(L):
                d = osh.wigner_d_matrix(l, beta)
                mm = fold * mus
                ok = np.abs(mm) <= l
                tmp = np.zeros((l + 1, len(mus)))
                tmp[:, ok] = d[l:, mm[ok] + l]
                self.D[j, l, : l + 1, :] = tmp

    def volume(self, f, g):
        """C[j, c, a] over the half b range, c along the master's orders, a along m'."""
        L, n, nc, nb = self.L, self.n, self.nc, self.nb
        F = osh.signed(f, L)[:, L - 1 + self.fold * self.mus]          # [l, mu]
        G = np.conj(osh.signed(g, L)[:, L - 1:])                         # [l, m' >= 0]
        U = np.einsum("lu,lp,jlpu->jpu", F, G, self.D)                   # [j, m', mu]
        # along c: V[j, m', c] = sum_mu U e^{-2 pi i mu c / nc}
        cidx = np.arange(nc)
        Ec = np.exp(-2j * np.pi * np.outer(self.mus, cidx) / nc)          # [mu, c]
        V = U @ Ec                                                        # [j, m', c]
        # along a: C = Re V[0] + 2 Re sum_{m' > 0} V[m'] e^{-2 pi i m' a / n}
        aidx = np.arange(n)
        Ea = np.exp(-2j * np.pi * np.outer(np.arange(L), aidx) / n)     # [m', a]
        k = np.where(np.arange(L) == 0, 1.0, 2.0)
        V0 = V.copy()
        V0[:, 0, :] = V0[:, 0, :].real
        C = np.real(np.einsum("jpc,pa->jca", V0 * k[None, :, None], Ea))
        return C

    def value_at(self, f, g, j, c, a):
        """C[j, c, a] alone (the same sums as volume, for one row j of b)."""
        L, n, nc = self.L, self.n, self.nc
        F = osh.signed(f, L)[:, L - 1 + self.fold * self.mus]
        G = np.conj(osh.signed(g, L)[:, L - 1:])
        U = np.einsum("lu,lp,lpu->pu", F, G, self.D[j])                  # [m', mu]
        V = U @ np.exp(-2j * np.pi * self.mus * c / nc)                   # [m']
        k = np.where(np.arange(L) == 0, 1.0, 2.0)
        V[0] = V[0].real
        return float(np.real(np.sum(k * V * np.exp(-2j * np.pi * np.arange(L) * a / n))))

    def rden(self, wf, wf2, area):
        """1 / sqrt(<w,f^2> - <w,f>^2 / area), 0 where that is <= 1e-6 area
        (so3.wgsl:614-626): wf, wf2 the window's correlation volumes."""
        a = max(area, 1e-12)
        v = wf2 - wf * wf / a
        return np.where(v > 1e-6 * a, 1 / np.sqrt(np.where(v > 1e-6 * a, v, 1)), 0.0)

    def peak(self, C, rden):
        """Argmax of C x rden (ties: the lowest index (j nc + c) n + a), parabolic offsets from the
        six neighbours (each clamped to half a step; b only inside 0 < j < n/2) and the
        orientation conj(Rz(a) Ry(b) Rz(c)) (so3.wgsl:631-796)."""
        n, nc = self.n, self.nc
        S = C * rden
        flat = S.reshape(-1)
        idx = int(np.argmax(flat))  # first of equal maxima
        v0 = flat[idx]
        j, c, a = idx // (nc * n), (idx // n) % nc, idx % n

        def vertex(lo, mid, hi):
            den = lo - 2 * mid + hi
            if den >= -1e-12:
                return 0.0
            return float(np.clip(0.5 * (lo - hi) / den, -0.5, 0.5))

        da = vertex(S[j, c, (a - 1) % n], v0, S[j, c, (a + 1) % n])
        dc = vertex(S[j, (c - 1) % nc, a], v0, S[j, (c + 1) % nc, a])
        dj = vertex(S[j - 1, c, a], v0, S[j + 1, c, a]) if 0 < j < n // 2 else 0.0
        step = TAU / n
        A, B, Cc = (a + da) * step, (j + dj) * step, (c + dc) * step
        q = qconj(zyz(A, B, Cc))
        if q[0] < 0:
            q = -q
        return q, v0, (A, B, Cc), idx


# ── rotations ──

def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def qconj(q):
    return np.array([q[0], -q[1], -q[2], -q[3]])


def qrot(q, v):
    u = np.asarray(q[1:], np.float64)
    t = 2 * np.cross(u, v)
    return v + q[0] * t + np.cross(u, t)


def zyz(a, b, c):
    return qmul([np.cos(a / 2), 0, 0, np.sin(a / 2)], qmul([np.cos(b / 2), 0, np.sin(b / 2), 0], [np.cos(c / 2), 0, 0, np.sin(c / 2)]))


def qexp(p):
    th = np.linalg.norm(p)
    k = 0.5 - th * th / 48 if th <= 1e-3 else np.sin(th / 2) / th
    return np.concatenate([[np.cos(th / 2)], k * np.asarray(p)])


# ── refinement and mutual information on the detector ──

def masked_pixels(fr):
    h, w = fr["h"], fr["w"]
    r, c = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    m = in_mask(r, c, fr["mask"])
    return r[m].astype(np.float64), c[m].astype(np.float64), np.nonzero(m.reshape(-1))[0]


def rot_sums(grad_field, side, fr, pat, t, q):
    """The 19 sums of orientation refinement (detector.wgsl:318-363)."""
    r, c, idx = masked_pixels(fr)
    u = fr["c0"] * r[:, None] + fr["c1"] * c[:, None] + np.asarray(t)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    cr = qrot(q, u)
    m = sheet_at(grad_field, side, cr)
    d = np.cross(cr, m[:, :3])
    e = np.asarray(pat, np.float64).reshape(-1)[idx]
    M = m[:, 3]
    return {"n": len(e), "m": M.sum(), "mm": (M * M).sum(), "me": (M * e).sum(), "sd": d.sum(0), "sv": (M[:, None] * d).sum(0),
            "se": (e[:, None] * d).sum(0), "dd": d.T @ d}


def ncc_grad_hess(s):
    """NCC, its gradient and Gauss-Newton matrix over the rotation from the sums
    (detector.wgsl:561-583)."""
    n = max(s["n"], 1)
    mean = s["m"] / n
    var_n = s["mm"] - n * mean * mean
    if not var_n > 1e-12:
        return -2.0, np.zeros(3), np.zeros((3, 3))
    inv = 1 / np.sqrt(var_n)
    ncc = s["me"] * inv
    au = (s["sv"] - mean * s["sd"]) * inv
    grad = (s["se"] - ncc * au) * inv
    hess = (s["dd"] - np.outer(s["sd"], s["sd"]) / n - np.outer(au, au)) * inv * inv
    return ncc, grad, hess


def sums6(grad_field, side, fr, pat, t, q):
    """The 43 sums of the six-direction fit (detector.wgsl:433-494): count, sum M, sum M^2,
    sum M e, then per direction i (a shift of the rays' origin along the sample's x, y, z, in
    pixels; a crystal rotation about x, y, z) sum d_i, sum M d_i, sum e d_i, and the upper
    triangle of sum d_i d_j row by row. d = (q^-1 grad M / |u|, c x grad M) for the pixel ray u,
    c = q u / |u|."""
    r, c, idx = masked_pixels(fr)
    u = fr["c0"] * r[:, None] + fr["c1"] * c[:, None] + np.asarray(t)
    dist = np.linalg.norm(u, axis=1, keepdims=True)
    cr = qrot(q, u / dist)
    m = sheet_at(grad_field, side, cr)
    a = np.cross(cr, m[:, :3])
    w = qrot(qconj(q), m[:, :3]) / dist
    d = np.concatenate([w, a], 1)
    e = np.asarray(pat, np.float64).reshape(-1)[idx]
    M = m[:, 3]
    out = [len(e), M.sum(), (M * M).sum(), (M * e).sum()]
    out += list(d.sum(0)) + list((M[:, None] * d).sum(0)) + list((e[:, None] * d).sum(0))
    dd = d.T @ d
    out += [dd[i, j] for i in range(6) for j in range(i, 6)]
    return np.asarray(out)


def local6(s):
    """NCC, its gradient and Gauss-Newton matrix over the six directions from the 43 sums
    (planefit.zig:177-201 localOf). Returns (ok, ncc, g[6], H[6, 6])."""
    s = np.asarray(s, np.float64)
    n = max(s[0], 1.0)
    mean = s[1] / n
    var_n = s[2] - n * mean * mean
    if not var_n > 1e-12:
        return False, 0.0, np.zeros(6), np.zeros((6, 6))
    inv = 1 / np.sqrt(var_n)
    ncc = s[3] * inv
    sd = s[4:10]
    au = (s[10:16] - mean * sd) * inv
    g = (s[16:22] - ncc * au) * inv
    H = np.zeros((6, 6))
    o = 22
    for i in range(6):
        for j in range(i, 6):
            H[i, j] = H[j, i] = (s[o] - sd[i] * sd[j] / n - au[i] * au[j]) * inv * inv
            o += 1
    return True, ncc, g, H


def fit6(grad_field, side, fr, pat, t0, q0, iters=20, unknowns=6, lambda0=1e-3, max_rot=np.radians(2.0), max_shift=2.0):
    """The six-direction fit by Levenberg-Marquardt steps (planefit.zig:177-247 localOf, damped,
    solve): the Newton step of the NCC model over the last `unknowns` directions with the matrix
    damped by lambda on its diagonal plus 1e-9 of its trace (numpy's solver), limited to
    max_shift pixels and max_rot radians, kept when the NCC rises (lambda / 10) and otherwise
    refused (lambda x 10). Returns (t, q, NCC)."""
    t = np.asarray(t0, np.float64).copy()
    q = np.asarray(q0, np.float64).copy()
    lam = lambda0
    ok, ncc, g, H = local6(sums6(grad_field, side, fr, pat, t, q))
    if not ok:
        return t, q, 0.0
    o = 6 - unknowns
    for _ in range(iters):
        A = H[o:, o:].copy()
        tr = np.trace(A)
        A[np.diag_indices(unknowns)] = A[np.diag_indices(unknowns)] * (1 + lam) + 1e-9 * tr + 1e-30
        x = np.zeros(6)
        x[o:] = np.linalg.solve(A, g[o:])
        sh, th = np.linalg.norm(x[:3]), np.linalg.norm(x[3:])
        if sh > max_shift:
            x[:3] *= max_shift / sh
        if th > max_rot:
            x[3:] *= max_rot / th
        t2 = t + x[:3]
        q2 = qmul(qexp(x[3:]), q)
        q2 = q2 / np.linalg.norm(q2) * (1 if q2[0] >= 0 else -1)
        # (the rotation block of g and H does not depend on the shift directions' sums)
        ok2, n2, g2, H2 = local6(sums6(grad_field, side, fr, pat, t2, q2))
        if ok2 and n2 > ncc:
            t, q, ncc, g, H = t2, q2, n2, g2, H2
            lam = max(lam * 0.1, 1e-9)
        else:
            lam = min(lam * 10, 1e9)
    return t, q, ncc


def refine(grad_field, side, fr, pat, t, q0, iters=8, max_step=np.radians(2.0), scale0=1.0, scale_min=0.02, scale_max=1e4):
    """Gauss-Newton on the orientation with the curvature rescaled from each gain against its
    promise, accept/reject against the best point (detector.wgsl:529-659). Returns
    (best q, best NCC, last improving step)."""
    cur = np.asarray(q0, np.float64)
    qbest, best, mu = cur.copy(), -2.0, scale0
    gbest, hbest = np.zeros(3), np.zeros((3, 3))
    last_step = promised = step_size = accepted = 0.0
    for it in range(iters + 1):
        ncc, grad, hess = ncc_grad_hess(rot_sums(grad_field, side, fr, pat, t, cur))
        if ncc > best:
            if accepted > 0.5 and promised > 1e-6:
                rho = (ncc - best) / promised
                mu = min(max(mu * (2 - rho), mu / 3), mu * 3)
            qbest, best, gbest, hbest = cur.copy(), ncc, grad, hess
            last_step = step_size
            accepted += 1
        else:
            mu *= 3
        mu = min(max(mu, scale_min), scale_max)
        if it == iters:
            break
        H = hbest * mu
        tr = max(np.trace(H), 1e-20)
        H = H + 1e-6 * tr * np.eye(3)
        det = np.linalg.det(H)
        step = np.zeros(3) if abs(det) < 1e-30 else np.linalg.solve(H, gbest)
        ln = np.linalg.norm(step)
        if ln > max_step:
            step *= max_step / ln
        promised = gbest @ step - 0.5 * step @ H @ step
        step_size = np.linalg.norm(step)
        q = qmul(qexp(step), qbest)
        cur = q / np.linalg.norm(q)
    return qbest, best, last_step


def nmi(sheets, side, fr, pat, t, q, bins=16):
    """2 I(A;B) / (H(A) + H(B)) of the pattern and the master at q, each binned between its own
    extremes under the mask (detector.wgsl:739-832)."""
    r, c, idx = masked_pixels(fr)
    u = fr["c0"] * r[:, None] + fr["c1"] * c[:, None] + np.asarray(t)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    s = sheet_at(sheets, side, qrot(q, u))
    e = np.asarray(pat, np.float64).reshape(-1)[idx]
    be = np.minimum(((e - e.min()) * (bins / max(e.max() - e.min(), 1e-20))).astype(np.int64), bins - 1)
    bs = np.minimum(((s - s.min()) * (bins / max(s.max() - s.min(), 1e-20))).astype(np.int64), bins - 1)
    H = np.zeros((bins, bins))
    np.add.at(H, (be, bs), 1)
    N = H.sum()

    def ent(x):
        x = x[x > 0] / N
        return -(x * np.log(x)).sum()

    hj = ent(H.reshape(-1))
    hm = ent(H.sum(1)) + ent(H.sum(0))
    return 2 * (hm - hj) / hm if hm > 1e-12 else 0.0
