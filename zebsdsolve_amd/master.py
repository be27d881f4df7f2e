"""A candidate phase: its master pattern on the sphere and what is derived from it once per phase
and bandwidth (src/master.zig).

Files are ebsdsim `.npz` masters: the intensities on the point group's fundamental sector, the
Lambert pixel of each, the point group's operators and the sector's bounding normals. They are
expanded here to two side x side Lambert sheets [north | south] and normalised to zero mean and
unit variance over the sphere (src/master.zig:157-166, 384-485). Loading never executes anything
from the file (numpy.load with allow_pickle off).
"""
import json
from dataclasses import dataclass, field

import numpy as np

from . import lambert, sh, sym


@dataclass
class Master:
    name: str
    side: int
    raw: np.ndarray            # float32 [2, side, side], normalised
    ops: np.ndarray            # point-group operators the master obeys, [n, 3, 3]
    point_group: int = 0
    lattice: tuple = None
    voltage_kv: float = 0.0
    rotations: np.ndarray = field(default=None)  # proper rotations (quaternions), identity first

    @property
    def laue(self):
        return sym.laue_of_rotations(self.rotations)

    def at(self, sheets, v):
        return lambert.at(sheets, self.side, v)


def normalize_sheets(sheets):
    """Zero mean, unit variance over both sheets, accumulated in float64 (src/master.zig:157-166)."""
    s = np.asarray(sheets, np.float64)
    mean = s.mean()
    sd = np.sqrt(((s - mean) ** 2).mean())
    k = 1.0 / sd if sd > 0 else 1.0
    return ((s - mean) * k).astype(np.float32)


def _sample_stored(sheet, has, side, x, y):
    """Bilinear over the stored pixels only (weights renormalised); the nearest stored pixel
    searched outward when none of the four is stored (src/master.zig:489-526)."""
    last = side - 1.0
    fj = np.clip((x + 1) * 0.5 * last, 0.0, last)
    fi = np.clip((y + 1) * 0.5 * last, 0.0, last)
    c0 = np.floor(fj).astype(np.int64)
    r0 = np.floor(fi).astype(np.int64)
    c1 = np.minimum(c0 + 1, side - 1)
    r1 = np.minimum(r0 + 1, side - 1)
    tj = fj - c0
    ti = fi - r0
    idx = [r0 * side + c0, r0 * side + c1, r1 * side + c0, r1 * side + c1]
    wts = [(1 - ti) * (1 - tj), (1 - ti) * tj, ti * (1 - tj), ti * tj]
    acc = np.zeros_like(fj)
    wsum = np.zeros_like(fj)
    for i, w in zip(idx, wts):
        h = has[i]
        acc += np.where(h, w * sheet[i], 0.0)
        wsum += np.where(h, w, 0.0)
    out = np.where(wsum > 1e-9, acc / np.where(wsum > 1e-9, wsum, 1.0), 0.0)
    for k in np.nonzero(~(wsum > 1e-9))[0]:
        # (round half away from zero; the coordinates are non-negative)
        ci = int(np.floor(fi[k] + 0.5))
        cj = int(np.floor(fj[k] + 0.5))
        found = 0.0
        for r in range(1, 8):
            hit = False
            for di in range(-r, r + 1):
                for dj in range(-r, r + 1):
                    ii, jj = ci + di, cj + dj
                    if ii < 0 or jj < 0 or ii >= side or jj >= side:
                        continue
                    if has[ii * side + jj]:
                        found = sheet[ii * side + jj]
                        hit = True
                        break
                if hit:
                    break
            if hit:
                break
        out[k] = found
    return out


def _grid_ops(ops):
    """The Lambert grid's own symmetries (swap, -x, -y, -z: codes 1..15) that are also the
    crystal's (src/master.zig:345-377, 407-422)."""
    found = []
    for code in range(1, 16):
        swap, nx, ny, nz = code & 1, code & 2, code & 4, code & 8
        sx, sy, sz = (-1.0 if nx else 1.0), (-1.0 if ny else 1.0), (-1.0 if nz else 1.0)
        M = np.array([[0, sx, 0], [sy, 0, 0], [0, 0, sz]]) if swap else np.diag([sx, sy, sz])
        if any(np.abs(g - M).sum() < 1e-6 for g in ops):
            found.append((bool(swap), bool(nx), bool(ny), bool(nz)))
    return found


def _grid_image(g, hemi, row, col, last):
    swap, nx, ny, nz = g
    r, c = (col, row) if swap else (row, col)
    if nx:
        c = last - c
    if ny:
        r = last - r
    h = 1 - hemi if nz else hemi
    return (h * (last + 1) + r) * (last + 1) + c


def from_sector(vals, kij, hw, ops, normals, by_grid=True):
    """Expand fundamental-sector intensities to both hemispheres: each pixel takes the value of
    its symmetry-equivalent direction deepest inside the stored sector, sampled over the stored
    pixels only; pixels the grid's own crystal symmetries map to an earlier pixel copy it
    (src/master.zig:384-485). Returns float32 [2, side, side], not yet normalised."""
    side = 2 * hw + 1
    plane = side * side
    sec = np.zeros(2 * plane, np.float64)
    has = np.zeros(2 * plane, bool)
    ix = kij[:, 0].astype(np.int64) + hw
    iy = kij[:, 1].astype(np.int64) + hw
    ok = (ix >= 0) & (iy >= 0) & (ix < side) & (iy < side)
    south = kij[:, 2] <= 0
    needs_south = bool(np.any(south[ok]))
    at = np.where(south, plane, 0) + iy * side + ix
    # (later entries of a pixel overwrite earlier ones, as the reference's loop does)
    sec[at[ok]] = np.asarray(vals, np.float64)[ok].astype(np.float32)
    has[at[ok]] = True

    eps = 1.0 / max(hw, 1)
    last = side - 1
    grid = _grid_ops(ops) if by_grid else []
    hemi, row, col = np.meshgrid(np.arange(2), np.arange(side), np.arange(side), indexing="ij")
    hemi, row, col = hemi.ravel(), row.ravel(), col.ravel()
    here = np.arange(2 * plane)
    first = np.ones(2 * plane, bool)
    rep_of = here.copy()
    for g in grid:
        img = _grid_image(g, hemi, row, col, last)
        first &= ~(img < here)
        rep_of = np.minimum(rep_of, img)
    todo = np.nonzero(first)[0]
    sheets = np.zeros(2 * plane, np.float32)
    opsm = np.asarray(ops, np.float64)
    nrm = np.asarray(normals, np.float64).reshape(-1, 3)
    chunk = 65536
    for s0 in range(0, len(todo), chunk):
        sel = todo[s0:s0 + chunk]
        x = 2 * col[sel] / last - 1
        y = 2 * row[sel] / last - 1
        d = lambert.square_to_sphere(x, y, hemi[sel] == 1)
        if len(nrm) == 0:
            rep = d @ opsm[0].T
        else:
            t = np.einsum("gij,pj->pgi", opsm, d)              # [p, g, 3]
            margin = np.min(np.einsum("pgi,ni->pgn", t, nrm), axis=2)
            inside = margin >= -eps
            any_in = inside.any(1)
            idx_in = np.argmax(np.where(inside, margin, -np.inf), axis=1)
            idx_all = np.argmax(margin, axis=1)
            pick = np.where(any_in, idx_in, idx_all)
            rep = t[np.arange(len(sel)), pick]
        use_south = (rep[:, 2] < 0) & needs_south
        u, w = lambert.sphere_to_square(rep)
        vn = _sample_stored(sec[:plane], has[:plane], side, u, w)
        if needs_south and use_south.any():
            vs = _sample_stored(sec[plane:], has[plane:], side, u[use_south], w[use_south])
            vn[use_south] = vs
        sheets[sel] = vn.astype(np.float32)
    if grid:
        sheets = sheets[rep_of]
    return sheets.reshape(2, side, side)


def measure_symmetry(raw, side, candidates, seed=12345):
    """The operators the raw master obeys: those whose self-correlation over 1500 random
    directions exceeds 0.97 (src/master.zig:77-138). Directions come from numpy's generator
    (the reference draws them from its own; the decision does not hinge on the draw)."""
    r = np.random.default_rng(seed)
    n = 1500
    z = 2 * r.random(n) - 1
    phi = 2 * np.pi * r.random(n)
    st = np.sqrt(np.maximum(0.0, 1 - z * z))
    dirs = np.stack([st * np.cos(phi), st * np.sin(phi), z], -1)
    base = lambert.at(raw, side, dirs)
    b0 = base - base.mean()
    var0 = (b0 * b0).sum()
    found = []
    for g in candidates:
        if any(np.abs(f - g).sum() < 1e-6 for f in found):
            continue
        v = lambert.at(raw, side, dirs @ np.asarray(g).T)
        v = v - v.mean()
        cov = (v * b0).sum()
        var1 = (v * v).sum()
        if cov / np.sqrt(max(var0 * var1, 1e-300)) > 0.97:
            found.append(np.asarray(g, np.float64))
    return np.array(found) if found else np.eye(3)[None]


def load(path):
    """Load an ebsdsim master (.npz) and measure the symmetry it obeys (src/master.zig:183-195,
    276-335)."""
    z = np.load(path, allow_pickle=False)
    files = z.files
    key = "fundamental_sector" if "fundamental_sector" in files else "integrated_fundamental_sector"
    if key not in files or "fundamental_kij" not in files or "pg_operators" not in files or "fs_normals" not in files:
        raise ValueError(f"{path}: not an ebsdsim master pattern")
    fs = z[key]
    kij = z["fundamental_kij"].astype(np.int64)
    n_k = kij.shape[0]
    if fs.ndim == 3:
        if fs.shape[2] != n_k:
            raise ValueError(f"{path}: sector and directions disagree")
        vals = fs.reshape(-1)[:n_k].astype(np.float64)
    else:
        n_sites = fs.shape[1] if fs.ndim == 2 else 1
        vals = fs.reshape(n_k, n_sites).astype(np.float64).mean(1)
    meta = {}
    if "meta_json" in files:
        try:
            txt = bytes(z["meta_json"].astype(np.uint8)).decode("utf-8", "replace")
            for tok in ("-Infinity", "Infinity", "NaN"):
                txt = txt.replace(tok, "null")
            meta = json.loads(txt)
        except ValueError:
            meta = {}
    hw = int(max(np.abs(kij[:, 0]).max(), np.abs(kij[:, 1]).max()))
    if isinstance(meta.get("halfw"), (int, float)):
        hw = int(meta["halfw"])
    ops = z["pg_operators"].astype(np.float64).reshape(-1, 3, 3)
    normals = z["fs_normals"].astype(np.float64)
    normals = normals.reshape(-1, 3) if normals.ndim == 2 else np.zeros((0, 3))
    sheets = normalize_sheets(from_sector(vals, kij, hw, ops, normals))
    side = 2 * hw + 1
    name = str(path).replace("\\", "/").split("/")[-1].rsplit(".", 1)[0]
    cell = meta.get("cell") if isinstance(meta.get("cell"), dict) else None
    lattice = None
    if cell:
        try:
            lattice = tuple(float(cell[k]) for k in ("a_angstrom", "b_angstrom", "c_angstrom", "alpha_deg", "beta_deg", "gamma_deg"))
        except (KeyError, TypeError, ValueError):
            lattice = None
    # candidates: the file's operators, each also with an inversion (src/master.zig:100-110)
    cands = list(ops) + [s * g for g in ops for s in (1.0, -1.0)]
    obeyed = measure_symmetry(sheets, side, cands)
    m = Master(name=name, side=side, raw=sheets, ops=obeyed,
               point_group=int(meta.get("pg_num", 0) or 0), lattice=lattice,
               voltage_kv=float(meta.get("voltage_kv", 0) or 0))
    m.rotations = sym.rotations_of(obeyed)
    return m


# ── the master as filtered for a pattern size ──

def bake_px(side, pixel_angle):
    """A detector pixel `pixel_angle` radians wide in sheet pixels, quantised to 5 % steps as the
    reference keys its baked masters (src/engine.zig:1679-1683). 0 without a pixel angle."""
    if not (pixel_angle > 0):
        return 0.0
    px = pixel_angle * side / np.sqrt(2 * np.pi)
    return float(1.05 ** int(np.round(np.log(px) / np.log(1.05))))


def bake_sigma(side, pixel_angle, master_smooth):
    """The Gaussian's sigma in sheet pixels for a detector pixel `pixel_angle` radians wide,
    quantised to 5 % steps of the pixel's width as the reference keys its baked masters
    (src/engine.zig:1676-1686, 1733-1736). 0 when no smoothing applies."""
    if not (pixel_angle > 0):
        return 0.0
    px = pixel_angle * side / np.sqrt(2 * np.pi)
    px_q = int(np.round(np.log(px) / np.log(1.05)))
    px = 1.05 ** px_q
    smooth = master_smooth * px
    return smooth if smooth > 0.05 else 0.0


def blur_passes(sigma):
    """The passes of doubling width the sphere blur takes to reach `sigma` (sheet pixels), each
    as (sigma_pass,) (src/prep.zig:215-237)."""
    out = []
    if not (sigma > 0.05):
        return out
    left = sigma * sigma
    s = min(sigma, 2.0)
    while left > 1e-6:
        if s * s > left or left - s * s < 0.02 * left:
            s = np.sqrt(left)
        left -= s * s
        out.append(s)
        s *= 2
    return out


# ── harmonics ──

@dataclass
class Harmonics:
    L: int
    rules: sh.Rules
    f: np.ndarray    # complex128 [pairs]
    f2: np.ndarray   # complex128 [pairs]


def harmonics(side, sheets, L):
    """Analyse the (filtered) master on a Gauss-Legendre grid of max(4L, 512) rings, drop f_00
    and what the detected symmetry forbids, then square the band-limited master on a 2L x 4L
    grid and analyse that (src/master.zig:665-711)."""
    nt = max(4 * L, 512)
    grid = sh.Grid.make(L, nt, 2 * nt)
    dirs = grid.directions().reshape(-1, 3)
    samples = lambert.at(sheets, side, dirs).reshape(nt, 2 * nt)
    f = grid.analyze(samples)
    f[0] = 0
    rules = sh.Rules.detect(L, f, 1e-5)
    ms, ls = sh.pairs_of(L)
    f = np.where(rules.allows(ls, ms), f, 0)
    g2 = sh.Grid.make(L, 2 * L, 4 * L)
    s2 = g2.synthesize(f)
    f2 = g2.analyze(s2 * s2)
    return Harmonics(L=L, rules=rules, f=f, f2=f2)


def weights(L, coef, tab, fold):
    """W[mu, l, k] = i^{-m} f_{l,m} d^l_{k,m}(pi/2), m = fold mu, for mu in [-mu_max, mu_max],
    complex, zero where l < max(k, |m|) (src/master.zig:723-753). Returns [mc, L, L] complex128."""
    mu_max = (L - 1) // fold
    mc = 2 * mu_max + 1
    out = np.zeros((mc, L, L), np.complex128)
    for mi in range(mc):
        m = (mi - mu_max) * fold
        am = abs(m)
        ph = 1j ** ((-m) % 4)
        for l in range(am, L):
            c = sh.signed_coef(coef, L, l, m)
            if c == 0:
                continue
            k = np.arange(l + 1)
            # d^l_{k,m} for k >= 0 from d^l_{k,|m|}: (-1)^{l+k} when m < 0 (src/sh.zig:286-302)
            sgn = np.where((l + k) % 2 == 1, -1.0, 1.0) if m < 0 else 1.0
            out[mi, l, :l + 1] = c * ph * sgn * tab[k, am, l]
    return out
