"""The oracle pinned by what the mathematics fixes (the reference's own host tests, sh.zig:441-606,
lambert.zig:66-85, restated): closed forms, exact rational arithmetic, brute-force quadrature,
library routines, and identities a plausible mistake breaks (a dropped term, a wrong sign or
index, a transposed operand). No GPU."""
import math

import numpy as np
import pytest
import scipy.special

from oracle import pipeline as op
from oracle import sh as osh
from tests.conftest import NI


def rand_coef(rng, L, scale=1.0):
    ms, _ = osh.pairs(L)
    re = rng.random(len(ms)) - 0.5
    im = np.where(ms == 0, 0.0, rng.random(len(ms)) - 0.5)
    return scale * (re + 1j * im)


def test_gauss_legendre_integrates_polynomials_exactly():
    x, w = osh.gl(12)
    assert x[0] > x[-1]
    assert abs(w.sum() - 2) < 1e-13
    # exact through degree 2n - 1 = 23, not beyond
    assert abs((w * x ** 22).sum() - 2 / 23) < 1e-13
    assert abs((w * x ** 24).sum() - 2 / 25) > 1e-8


def test_legendre_orthonormal_with_condon_shortley_phase():
    L = 10
    x, w = osh.gl(24)
    P = osh.plm(L, x)
    ms, ls = osh.pairs(L)
    for m in range(L):
        sel = np.nonzero(ms == m)[0]
        G = (P[sel] * w) @ P[sel].T * 2 * np.pi
        assert np.abs(G - np.eye(len(sel))).max() < 1e-12
    # Condon-Shortley: P_1^1(x) = -sqrt(3 / (8 pi)) sqrt(1 - x^2)
    xx = 0.3
    assert abs(osh.plm(2, [xx])[osh.pair(2, 1, 1), 0] + np.sqrt(3 / (8 * np.pi)) * np.sqrt(1 - xx * xx)) < 1e-14


def test_wigner_eigen_route_matches_exact_rational_sum():
    for l in range(0, 9):
        d = osh.wigner_d_matrix(l, np.pi / 2)
        for mp in range(-l, l + 1):
            for m in range(-l, l + 1):
                assert abs(d[mp + l, m + l] - osh.wigner_exact(l, mp, m)) < 1e-13


def test_wigner_factorization_and_unitarity():
    # d^l_{m'm}(beta) = i^{m'-m} sum_k d^l_{km'}(pi/2) d^l_{km}(pi/2) e^{-ik beta} (sh.zig:508-518)
    beta = 0.83
    for l in (3, 5, 8):
        d = osh.wigner_d_matrix(l, beta)
        h = osh.wigner_d_matrix(l, np.pi / 2)
        k = np.arange(-l, l + 1)
        for mp in range(-l, l + 1):
            for m in range(-l, l + 1):
                acc = np.sum(h[:, mp + l] * h[:, m + l] * np.exp(-1j * k * beta)) * 1j ** ((mp - m) % 4)
                assert abs(acc.real - d[mp + l, m + l]) < 1e-12 and abs(acc.imag) < 1e-12
        assert np.abs(d @ d.T - np.eye(2 * l + 1)).max() < 1e-12
    # d^1_{00}(beta) = cos(beta), d^1_{10}(beta) = -sin(beta)/sqrt(2): Wigner's sum at l = 1,
    # m' = 1, m = 0 has the single term -sqrt(2) cos(beta/2) sin(beta/2)
    d1 = osh.wigner_d_matrix(1, beta)
    assert abs(d1[1, 1] - np.cos(beta)) < 1e-14 and abs(d1[2, 1] + np.sin(beta) / np.sqrt(2)) < 1e-14


def test_analysis_inverts_synthesis_and_parseval():
    rng = np.random.default_rng(7)
    L = 12
    c = rand_coef(rng, L)
    x, w = osh.gl(16)
    s = osh.synthesize(c, L, x, 32)
    back = osh.analyze(s, L, x, w)
    assert np.abs(back - c).max() < 1e-12
    ms, _ = osh.pairs(L)
    power = np.sum(np.where(ms == 0, 1, 2) * np.abs(c) ** 2)
    assert abs((w[:, None] * s * s).sum() * 2 * np.pi / 32 - power) < 1e-11


def _f_at(f, L, v):
    """f(v) from its coefficients, via scipy's spherical harmonics."""
    theta = np.arccos(np.clip(v[..., 2], -1, 1))
    phi = np.arctan2(v[..., 1], v[..., 0])
    ms, ls = osh.pairs(L)
    out = np.zeros(v.shape[:-1])
    for p, (m, l) in enumerate(zip(ms, ls)):
        y = scipy.special.sph_harm_y(l, m, theta, phi)
        out += (1 if m == 0 else 2) * np.real(f[p] * y)
    return out


def test_correlation_formula_equals_the_integral():
    # C(a, b, c) = int g(x) f(S^-1 x) dOmega by quadrature of the synthesized functions
    rng = np.random.default_rng(3)
    L = 6
    f, g = rand_coef(rng, L), rand_coef(rng, L)
    x, w = osh.gl(24)
    gs = osh.synthesize(g, L, x, 48)
    a, b, c = 0.9, 0.6, 2.2
    S = op.zyz(a, b, c)
    st = np.sqrt(1 - x * x)
    phi = 2 * np.pi * np.arange(48) / 48
    d = np.stack(np.broadcast_arrays(st[:, None] * np.cos(phi), st[:, None] * np.sin(phi), x[:, None]), -1)
    v = op.qrot(op.qconj(S), d.reshape(-1, 3)).reshape(d.shape)
    integral = (w[:, None] * gs * _f_at(f, L, v)).sum() * 2 * np.pi / 48
    assert abs(integral - osh.correlation_at(f, g, L, a, b, c)) < 1e-10


@pytest.mark.parametrize("fold", [1, 4])
def test_volume_equals_direct_sum_at_every_cell(fold):
    rng = np.random.default_rng(5)
    L = 8
    f = rand_coef(rng, L)
    ms, ls = osh.pairs(L)
    f = np.where(ms % fold == 0, f, 0)  # a master with a fold-fold axis along z
    g = rand_coef(rng, L)
    so = op.So3(L, fold)
    C = so.volume(f, g)
    n, nc = so.n, so.nc
    rng2 = np.random.default_rng(1)
    for _ in range(12):
        j, c, a = rng2.integers(0, so.nb), rng2.integers(0, nc), rng2.integers(0, n)
        ref = osh.correlation_at(f, g, L, 2 * np.pi * a / n, 2 * np.pi * j / n, 2 * np.pi * c / n)
        assert abs(C[j, c, a] - ref) < 1e-12


def test_peak_of_a_rotated_function_is_its_rotation():
    # g = Lambda_S f has coefficients sum_m D^l_{m'm}(S) f_lm; its correlation with f peaks at S,
    # and the orientation reported is conj(S)
    rng = np.random.default_rng(11)
    L = 10
    f = rand_coef(rng, L)
    so = op.So3(L, 1)
    n = so.n
    a, b, c = 2 * np.pi * 5 / n, 2 * np.pi * 3 / n, 2 * np.pi * 9 / n  # on the grid
    F = osh.signed(f, L)
    ms, ls = osh.pairs(L)
    g = np.zeros(len(ms), np.complex128)
    for p, (mp, l) in enumerate(zip(ms, ls)):
        d = osh.wigner_d_matrix(l, b)
        m = np.arange(-l, l + 1)
        g[p] = np.sum(np.exp(-1j * mp * a) * d[mp + l, :] * np.exp(-1j * m * c) * F[l, L - 1 - l:L + l])
    C = so.volume(f, g)
    q, v0, ang, idx = so.peak(C, np.ones_like(C))
    S = op.zyz(a, b, c)
    assert abs(abs(np.dot(q, op.qconj(S))) - 1) < 1e-9
    assert np.allclose(ang, (a, b, c), atol=1e-9)


def test_lambert_round_trip_and_equal_area():
    pts = np.array([[0.3, -0.7], [-0.9, 0.2], [0.0, 0.5], [1.0, 1.0], [-0.01, 0.01]])
    for south in (False, True):
        v = op.square_to_sphere(pts[:, 0], pts[:, 1], south)
        assert np.allclose(np.linalg.norm(v, axis=1), 1)
        u, w = op.sphere_to_square(v)
        assert np.abs(u - pts[:, 0]).max() < 1e-10 and np.abs(w - pts[:, 1]).max() < 1e-10
    h = 1e-6
    for p in pts[:3]:
        dx = (op.square_to_sphere(p[0] + h, p[1], False) - op.square_to_sphere(p[0] - h, p[1], False)) / (2 * h)
        dy = (op.square_to_sphere(p[0], p[1] + h, False) - op.square_to_sphere(p[0], p[1] - h, False)) / (2 * h)
        assert abs(np.linalg.norm(np.cross(dx, dy)) - np.pi / 2) < 1e-6


def _smooth_master(side=101):
    """A smooth function on the sphere as [north | south] sheets: f = x y + 0.3 z^2 + 0.2 x."""
    v = op.texel_dirs(side)
    return (v[:, 0] * v[:, 1] + 0.3 * v[:, 2] ** 2 + 0.2 * v[:, 0]).reshape(2, side, side), side


def test_gradient_field_is_the_tangential_gradient():
    sheets, side = _smooth_master(201)
    rng = np.random.default_rng(2)
    tex = rng.choice(2 * side * side, 200, replace=False)
    g = op.gradient(sheets, side, tex)
    v = op.texel_dirs(side)[tex]
    grad = np.stack([v[:, 1] + 0.2, v[:, 0], 0.6 * v[:, 2]], 1)
    tang = grad - (grad * v).sum(1, keepdims=True) * v
    assert np.abs(g[:, :3] - tang).max() < 2e-2
    assert np.abs(g[:, 3] - sheets.reshape(-1)[tex]).max() == 0


def test_refinement_gradient_and_matrix_from_sums_are_the_ncc_derivatives():
    # the NCC's gradient over a small crystal rotation exp(phi) q (from the 19 sums) against
    # central differences of the NCC itself; J'J from the sums against the outer product of the
    # directional derivatives (exact for a pattern that is its simulation)
    sheets, side = _smooth_master(301)
    grad_field = op.gradient(sheets, side)
    fr = op.frame(op.bruker_R(), 30, 30)
    t = op.translation(op.bruker_R(), (0.5, 0.5, 0.55), 30.0, 30.0)
    q0 = op.qexp(np.array([0.3, -0.2, 0.5]))
    pat = op.normalize(op.project(grad_field[:, 3].reshape(2, side, side), side, fr, t, op.qexp(np.array([0.31, -0.19, 0.52]))), fr["mask"])
    r, c, idx = op.masked_pixels(fr)

    def ncc_at(q):
        s = op.rot_sums(grad_field, side, fr, pat, t, q)
        return op.ncc_grad_hess(s)[0]

    ncc, g, H = op.ncc_grad_hess(op.rot_sums(grad_field, side, fr, pat, t, q0))
    h = 1e-4
    fd = np.array([(ncc_at(op.qmul(op.qexp(h * e), q0)) - ncc_at(op.qmul(op.qexp(-h * e), q0))) / (2 * h) for e in np.eye(3)])
    assert np.abs(fd - g).max() < 0.02 * np.abs(g).max()
    assert np.all(np.linalg.eigvalsh(H) > -1e-9)


def test_six_direction_gradient_from_sums_is_the_ncc_derivative():
    # the fit's gradient over a shift of the rays' origin (h = 1e-5 of the pattern height, in
    # pixels) and a crystal rotation (h = 1e-4 rad), from the 43 sums, against central
    # differences of the NCC; its rotation block equals the orientation refinement's
    sheets, side = _smooth_master(301)
    grad_field = op.gradient(sheets, side)
    h_px = 30
    fr = op.frame(op.bruker_R(), h_px, h_px)
    t0 = op.translation(op.bruker_R(), (0.5, 0.5, 0.55), float(h_px), float(h_px))
    q0 = op.qexp(np.array([0.3, -0.2, 0.5]))
    pat = op.normalize(op.project(grad_field[:, 3].reshape(2, side, side), side, fr, t0 + np.array([0.3, -0.2, 0.1]),
                                  op.qexp(np.array([0.31, -0.19, 0.52]))), fr["mask"])

    def ncc_at(t, q):
        return op.local6(op.sums6(grad_field, side, fr, pat, t, q))[1]

    ok, ncc, g, H = op.local6(op.sums6(grad_field, side, fr, pat, t0, q0))
    assert ok and abs(ncc - ncc_at(t0, q0)) == 0
    hs = 1e-5 * h_px
    fd_shift = np.array([(ncc_at(t0 + hs * e, q0) - ncc_at(t0 - hs * e, q0)) / (2 * hs) for e in np.eye(3)])
    hr = 1e-4
    fd_rot = np.array([(ncc_at(t0, op.qmul(op.qexp(hr * e), q0)) - ncc_at(t0, op.qmul(op.qexp(-hr * e), q0))) / (2 * hr) for e in np.eye(3)])
    assert np.abs(fd_shift - g[:3]).max() < 0.02 * np.abs(g[:3]).max()
    assert np.abs(fd_rot - g[3:]).max() < 0.02 * np.abs(g[3:]).max()
    assert np.all(np.linalg.eigvalsh(H) > -1e-9)
    # the rotation block is the orientation refinement's
    n3, g3, H3 = op.ncc_grad_hess(op.rot_sums(grad_field, side, fr, pat, t0, q0))
    assert abs(n3 - ncc) < 1e-12 and np.abs(g3 - g[3:]).max() < 1e-12 and np.abs(H3 - H[3:, 3:]).max() < 1e-12


def test_nmi_is_the_histogram_definition():
    rng = np.random.default_rng(4)
    sheets, side = _smooth_master(101)
    fr = op.frame(op.bruker_R(), 20, 20)
    t = op.translation(op.bruker_R(), (0.5, 0.5, 0.55), 20.0, 20.0)
    q = op.qexp(np.array([0.1, 0.2, 0.3]))
    pat = op.normalize(rng.normal(size=(20, 20)), fr["mask"])
    v = op.nmi(sheets, side, fr, pat, t, q, bins=8)
    r, c, idx = op.masked_pixels(fr)
    u = fr["c0"] * r[:, None] + fr["c1"] * c[:, None] + t
    s = op.sheet_at(sheets, side, op.qrot(q, u / np.linalg.norm(u, axis=1, keepdims=True)))
    e = pat.reshape(-1)[idx]
    # bins between each one's own extremes; the maximum lands in the last bin
    eb = np.minimum(np.floor((e - e.min()) / (e.max() - e.min()) * 8), 7)
    sb = np.minimum(np.floor((s - s.min()) / (s.max() - s.min()) * 8), 7)
    joint, _, _ = np.histogram2d(eb, sb, bins=8, range=[[0, 8], [0, 8]])
    p = joint / joint.sum()
    pa, pb = p.sum(1), p.sum(0)
    ent = lambda x: -np.sum(x[x > 0] * np.log(x[x > 0]))
    mi = ent(pa) + ent(pb) - ent(p.reshape(-1))
    assert abs(v - 2 * mi / (ent(pa) + ent(pb))) < 1e-12


def test_normalize_is_zero_mean_unit_norm_under_the_mask():
    rng = np.random.default_rng(0)
    fr = op.frame(op.bruker_R(), 16, 20)
    p = op.normalize(rng.random((16, 20)), fr["mask"])
    r, c = np.meshgrid(np.arange(16), np.arange(20), indexing="ij")
    m = op.in_mask(r, c, fr["mask"])
    assert abs(p[m].mean()) < 1e-15 and abs((p * p).sum() - 1) < 1e-12 and np.all(p[~m] == 0)


def test_oracle_recovers_known_orientations_end_to_end():
    # the nickel master, two patterns at known orientations: the correlation peak lands within the
    # grid's reach (L = 32: 5.6 deg steps) and the refinement on it converges to the truth
    from tests import cases
    from tests.conftest import NI
    from zebsdsolve_amd import master, sym
    m = master.load(NI)
    L = 32
    pats, truth = cases.simulate(m.raw, m.side, 2, seed=3)
    fr, t = cases.oracle_frame(60, 60)
    f, f2, fold, _ = op.harmonics(m.raw, m.side, L)
    so = op.So3(L, fold)
    wa, wst = op.coefficients(None, fr, t, L, center=False, norm=False, window=True)
    rden = so.rden(so.volume(f, wa), so.volume(f2, wa), wst[0])
    grad = op.gradient(m.raw, m.side)
    for b in range(2):
        pn = op.normalize(pats[b].astype(np.float64), fr["mask"])
        alm, _ = op.coefficients(pn, fr, t, L)
        q, v0, _, _ = so.peak(so.volume(f, alm), rden)
        assert np.degrees(sym.disorientation(q, truth[b], m.rotations)) < 3.0
        qr, ncc, _ = op.refine(grad, m.side, fr, pn, t, q, iters=8)
        assert np.degrees(sym.disorientation(qr, truth[b], m.rotations)) < 0.02 and ncc > 0.99


# ── the pattern filters (prep.wgsl:92-287) ──

def _smooth_image(h, w, seed=0):
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    img = ndimage.gaussian_filter(rng.normal(size=(h, w)), 3) * 50 + 400
    return img + np.linspace(0, 30, w)[None, :] + np.linspace(0, 10, h)[:, None]


def test_direct_blur_is_the_truncated_gaussian():
    """Inside the image, the separable taps out to ceil(3 sigma) are scipy's Gaussian truncated at
    the same radius; a constant image stays constant up to the edges (the renormalization)."""
    from scipy import ndimage
    img = _smooth_image(60, 70)
    for sigma in (1.7, 4.3, 9.0):
        r = max(1, int(np.ceil(3 * sigma)))
        ref = ndimage.gaussian_filter(img, sigma, mode="constant", truncate=(r - 0.5 + 1e-9) / sigma)
        got = op.blur_direct(img, sigma)
        assert np.abs(got[r:-r, r:-r] - ref[r:-r, r:-r]).max() < 1e-9
        assert np.allclose(op.blur_direct(np.full((30, 40), 7.0), sigma), 7.0, rtol=0, atol=1e-12)


def test_box_blur_stays_within_a_fraction_of_a_percent_of_the_gaussian():
    """The wide path (boxes of floor(sigma / 6) pixels, a coarse Gaussian, a spread) against the
    Gaussian carried to 6 sigma, renormalized at the edges: the reference's stated 0.2 % rms
    (prep.wgsl:166-176) on smooth images; a constant stays constant."""
    from scipy import ndimage
    for h, w, sigma in ((80, 80, 12.0), (120, 120, 18.0), (90, 70, 13.5)):
        img = _smooth_image(h, w, seed=h)
        ref = (ndimage.gaussian_filter(img, sigma, mode="constant", truncate=6) /
               ndimage.gaussian_filter(np.ones_like(img), sigma, mode="constant", truncate=6))
        got = op.blur_box(img, sigma)
        rms = np.sqrt(np.mean((got - ref) ** 2)) / img.std()
        assert rms < 5e-3, (h, w, sigma, rms)
        assert np.allclose(op.blur_box(np.full((h, w), 3.0), sigma), 3.0, atol=1e-9)


def test_lattice_spreads_cells_over_the_map():
    """engine.zig:1325-1341: every cell of a small map; about `want` distinct cells of a large one;
    3 x 4 with 6 wanted: rows 0 and 2, columns 0, 2, 3 (by hand)."""
    assert np.array_equal(op.lattice(10, 10, 1024), np.arange(100))
    assert np.array_equal(op.lattice(3, 4, 6), [0, 2, 3, 8, 10, 11])
    for rows, cols, want in ((48, 64, 1024), (100, 300, 1024), (1, 5000, 256)):
        c = op.lattice(rows, cols, want)
        assert len(np.unique(c)) == len(c) <= want and c.min() >= 0 and c.max() < rows * cols
        assert len(c) >= 0.9 * want
    # a map far taller than wide: round(sqrt(50 x 700 / 3)) = 108 rows of one column each
    c = op.lattice(700, 3, 50)
    assert len(c) == 108 and np.all(c % 3 == 1)


def test_prep_chain_order_and_identities():
    """No filter: prep is the normalization; a static background equal to the pattern leaves
    nothing (all zeros); a linear ramp is its own blur away from the edges (symmetric taps)."""
    fr = op.frame(np.eye(3), 40, 50)
    img = _smooth_image(40, 50)
    assert np.allclose(op.prep(img, fr["mask"]), op.normalize(img, fr["mask"]))
    assert np.allclose(op.prep(img, fr["mask"], background=img), 0.0)
    ramp = np.add.outer(np.arange(40) * 0.7, np.arange(50) * 1.3) + 100
    r = int(np.ceil(3 * 4.0))
    assert np.abs(op.blur_direct(ramp, 4.0)[r:-r, r:-r] - ramp[r:-r, r:-r]).max() < 1e-9


def test_clahe_with_one_tile_and_no_clip_is_histogram_equalization():
    """One tile and a clip no bin reaches: every count kept, so a pixel's table value is the
    fraction of the image at or below its level (the empirical CDF), read between levels; on a
    ramp with equal counts per level that CDF is (b + 1) / bins."""
    from oracle import pipeline as op
    bins = 16
    # each level b holds the same 8 pixels, values at the level's centre
    vals = np.minimum(np.repeat((np.arange(bins) + 0.25) / (bins - 1), 8), 1.0)
    img = vals.reshape(16, 8).astype(np.float32)
    out = op.clahe(img, 1, bins, 1e6)
    b0 = np.clip(np.floor(img * (bins - 1)), 0, bins - 2).astype(int)
    fb = img * (bins - 1) - b0
    want = ((b0 + 1) * (1 - fb) + (b0 + 2) * fb) / bins
    assert np.abs(out - want).max() < 1e-6
    # distinct random values: the table is the rank fraction, and the output is monotone
    rng = np.random.default_rng(3)
    x = rng.random((20, 20)).astype(np.float32)
    y = op.clahe(x, 1, 64, 1e6)
    order = np.argsort(x.ravel())
    assert np.all(np.diff(y.ravel()[order]) >= -1e-7)
    bins_of = np.clip(x.ravel() * 63, 0, 63).astype(int)
    cdf = np.cumsum(np.bincount(bins_of, minlength=64)) / 400.0
    b0 = np.clip(np.floor(x.ravel() * 63), 0, 62).astype(int)
    f = x.ravel() * 63 - b0
    assert np.abs(y.ravel() - (cdf[b0] * (1 - f) + cdf[b0 + 1] * f)).max() < 1e-6


def test_clahe_clip_spreads_what_it_cuts():
    """A constant 8 x 8 image at 0.5 (level 31 of 64), one tile: with clip 40 the full level
    (64 counts) is cut to 40 and the 24 cut counts go one each to levels 0..23 (fewer than the
    levels: no even share), so the table is (24 + 40) / 64 = 1 from level 31 on and the output
    is 1. With clip 8, 56 are cut and go one each to levels 0..55: level 31 holds 8 + 1, the
    table reads (31 + 9) / 64 at level 31 and (32 + 9) / 64 at 32, and 0.5 (31.5 levels) takes
    their mean."""
    from oracle import pipeline as op
    img = np.full((8, 8), 0.5, np.float32)
    assert np.allclose(op.clahe(img, 1, 64, 40.0), 1.0)
    out = op.clahe(img, 1, 64, 8.0)
    assert np.allclose(out, 0.5 * (31 + 9) / 64 + 0.5 * (32 + 9) / 64)


def test_clahe_pads_by_reflection_to_whole_tiles():
    """A 6 x 6 image in a 4 x 4 grid is padded by 2 rows and columns, reflected at the top and
    left (padded index i reads |i - 2|): every tile's histogram holds 2 x 2 pixels. A left-right
    ramp of 6 levels: the first tile column holds image columns 2, 1 (the reflection), the next
    0, 1, then 2, 3 and 4, 5; with 6 levels and no clip each tile's table is 1/2 at its lower
    level and 1 from its upper one, so the first pixel (level 0, between tiles 0 and 1 at
    gx = 0.25) reads 0.75 x 0 + 0.25 x 1/2 from the two tables."""
    from oracle import pipeline as op
    assert list(op._reflect(np.arange(8), 6, 2)) == [2, 1, 0, 1, 2, 3, 4, 5]
    img = np.tile(np.linspace(0, 1, 6, dtype=np.float32), (6, 1))
    out = op.clahe(img, 4, 6, 1e6)
    assert out.shape == (6, 6) and np.allclose(out, out[:1])  # every row the same ramp
    # pixel (0, 0): level 0; tile 0 holds levels {1, 2} (table 0 at level 0), tile 1 holds
    # levels {0, 1} (table 1/2 at level 0); gx = 0.5 / 6 * 3 = 0.25 between tile columns 0 and 1
    assert abs(out[0, 0] - (0.75 * 0.0 + 0.25 * 0.5)) < 1e-6


def test_bandpass_of_a_constant_is_zero_and_a_difference_of_blurs():
    """Both Gaussians renormalize their taps inside the image, so a constant passes both
    unchanged and the band pass removes it; otherwise it is the fine blur minus the coarse one
    (the image itself below 0.05 pixels)."""
    from oracle import pipeline as op
    c = np.full((40, 50), 3.0)
    assert np.abs(op.prep(c, None, bandpass=(1.0, 8.0), norm=False)).max() < 1e-12
    rng = np.random.default_rng(5)
    x = rng.random((40, 50))
    bp = op.prep(x, None, bandpass=(1.0, 14.0), norm=False)
    assert np.allclose(bp, op.blur_direct(x, 1.0) - op.blur_box(x, 14.0))
    bp0 = op.prep(x, None, bandpass=(0.0, 8.0), norm=False)
    assert np.allclose(bp0, x - op.blur_direct(x, 8.0))


def test_six_direction_fit_climbs_to_the_truth():
    """The oracle's Levenberg-Marquardt fit on a noise-free nickel pattern, started half a pixel
    and half a degree off: the NCC maximum is the truth (the pattern is the master there, NCC 1),
    so the fit ends on it; with lambda 0 its first step is the Newton step H^-1 g."""
    from oracle import pipeline as op
    from tests import cases
    from zebsdsolve_amd import master as mm
    from tests.conftest import NI
    m = mm.load(NI)
    side = m.side
    sheets = np.asarray(m.raw, np.float64).reshape(2, side, side)
    grad = op.gradient(sheets, side)
    pats, truth = cases.simulate(sheets, side, 1, seed=4)
    fr, t = cases.oracle_frame(60, 60)
    pat = op.normalize(pats[0].astype(np.float64), fr["mask"])
    rng = np.random.default_rng(8)
    t0 = t + rng.uniform(-0.5, 0.5, 3)
    q0 = op.qmul(op.qexp(rng.normal(size=3) * np.radians(0.3)), truth[0])
    tf, qf, ncc = op.fit6(grad.reshape(-1, 4), side, fr, pat, t0, q0, iters=12)
    d = np.degrees(2 * np.arccos(min(1.0, abs(float(np.dot(qf, truth[0]))))))
    assert ncc > 0.999 and d < 0.02 and np.abs(tf - t).max() < 0.1, (ncc, d, tf - t)
    ok, n0, g, H = op.local6(op.sums6(grad.reshape(-1, 4), side, fr, pat, t0, q0))
    assert ok and ncc > n0
    # lambda 0, one step: the Newton step (the 1e-9 trace term aside) moves the shift by H^-1 g
    x = np.linalg.solve(H + 1e-9 * np.trace(H) * np.eye(6), g)
    assert np.linalg.norm(x[:3]) < 2.0 and np.linalg.norm(x[3:]) < np.radians(2.0)
    t1, q1, n1 = op.fit6(grad.reshape(-1, 4), side, fr, pat, t0, q0, iters=1, lambda0=0.0)
    assert n1 > n0 and np.allclose(t1, t0 + x[:3], atol=1e-9)


def _real_field(coef, L, v):
    """The real field sum over signed m of a_lm Y_lm at directions v, by scipy (no oracle code)."""
    ms, ls = osh.pairs(L)
    th = np.arccos(np.clip(v[..., 2], -1, 1))
    ph = np.arctan2(v[..., 1], v[..., 0])
    out = np.zeros(v.shape[:-1])
    for c, m, l in zip(coef, ms, ls):
        if c != 0:
            out += (1.0 if m == 0 else 2.0) * np.real(c * scipy.special.sph_harm_y(l, m, th, ph))
    return out


def _gl_dirs(nt, nphi):
    x, w = np.polynomial.legendre.leggauss(nt)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    st = np.sqrt(1 - x * x)
    d = np.stack(np.broadcast_arrays(st[:, None] * np.cos(phi), st[:, None] * np.sin(phi), x[:, None]), -1)
    return d, w[:, None] * (2 * np.pi / nphi) * np.ones((1, nphi))


def test_sphere_blur_scales_a_degree_l_field_by_the_heat_kernel_of_its_lattice():
    """sphere_blur (mastersph.wgsl:87-109; src/prep.zig:215-237: a 9 x 9 lattice of step
    0.6 sigma in the tangent plane, weights exp(-(a^2 + b^2) / 2 sigma^2)) is, to first order, a
    rotation-invariant smoothing, so a degree-l harmonic comes back scaled by one number:
    (1) that number, the lattice's zonal mean of P_l(cos theta) (tan theta = the lattice radius),
    is the heat kernel exp(-l(l+1) s^2 / 2) of the lattice's own per-axis variance s^2, taken from
    the 1-D lattice alone (0.947 sigma^2: a step of 0.5 sigma, a weight exp(-x^2 / sigma^2) or a
    wider lattice would each move it); (2) the oracle's blur of Re Y_lm (l = 2, 6, 12) sampled on
    the sheets is that number times the field at every texel tested, within the sheets' bilinear
    error. A misread tangent basis, step, weight or normalization fails (2) by 2e-2 or more."""
    side, sigma = 201, 0.04
    d = op.texel_dirs(side)
    rng = np.random.default_rng(0)
    sel = rng.choice(len(d), 3000, replace=False)
    x1 = 0.6 * np.arange(-4, 5)
    w1 = np.exp(-x1 * x1 / 2)
    s2 = np.sum(w1 * x1 * x1) / w1.sum() * sigma ** 2
    assert abs(s2 / sigma ** 2 - 0.947) < 1e-3
    i, j = np.meshgrid(np.arange(-4, 5), np.arange(-4, 5), indexing="ij")
    r = 0.6 * sigma * np.hypot(i, j)
    wl = np.exp(-(r * r) / (2 * sigma ** 2))
    th, ph = np.arccos(np.clip(d[:, 2], -1, 1)), np.arctan2(d[:, 1], d[:, 0])
    for l, m in ((2, 1), (6, 3), (12, 5)):
        lam = np.sum(wl * scipy.special.eval_legendre(l, 1 / np.sqrt(1 + r * r))) / wl.sum()
        heat = np.exp(-l * (l + 1) * s2 / 2)
        assert abs(lam - heat) < 2e-4, (l, lam, heat)
        Y = np.real(scipy.special.sph_harm_y(l, m, th, ph))
        got = op.blur(Y, side, sigma, texels=sel)
        err = np.abs(got - lam * Y[sel]).max() / np.abs(Y).max()
        assert err < 4e-3, (l, err, 1 - lam)


def test_square_of_the_master_obeys_parseval_and_brute_force_quadrature():
    """harmonics' f2 (src/master.zig:680-710: the band-limited master synthesized on 2L x 4L,
    squared, analysed again): its f2_00 is Parseval's sum k |f_lm|^2 / sqrt(4 pi) (k = 1 for
    m = 0, 2 above), and every f2_lm equals the projection of the squared field onto Y_lm by a
    different quadrature (48 x 96 points, scipy's harmonics), on the nickel master at L = 8."""
    from zebsdsolve_amd import master
    L = 8
    m = master.load(NI)
    f, f2, fold, inversion = op.harmonics(m.raw, m.side, L)
    ms, ls = osh.pairs(L)
    k = np.where(ms == 0, 1.0, 2.0)
    assert abs(f2[0].real - np.sum(k * np.abs(f) ** 2) / np.sqrt(4 * np.pi)) < 1e-12 * np.abs(f2[0])
    d, wt = _gl_dirs(48, 96)
    F = _real_field(f, L, d)
    th, ph = np.arccos(np.clip(d[..., 2], -1, 1)), np.arctan2(d[..., 1], d[..., 0])
    want = np.array([np.sum(wt * F * F * np.conj(scipy.special.sph_harm_y(l, mm, th, ph))) for mm, l in zip(ms, ls)])
    assert np.abs(f2 - want).max() < 1e-10 * np.abs(want).max()
    # and the square is not the f of a master squared texel by texel (what a misread would take)
    assert np.abs(f2[1:]).max() > 1e-3 * np.abs(f2[0])


def test_normalized_correlation_of_a_window_limited_match_is_one():
    """So3.rden (so3.wgsl:614-626: 1 / sqrt(<w,f^2> - <w,f>^2 / area)) with So3.volume: for a
    pattern that is the master rotated by S, seen through a window w (a spherical cap) with its
    window mean removed, volume x rden is its NCC with the master at every cell. At S it is 1,
    the maximum; at other cells it equals the NCC by brute-force quadrature of the rotated master
    (scipy's harmonics at the rotated directions). The master is band-limited to l < 4, so its
    square lies below L = 8 and every inner product is exact on the quadrature grid: a dropped
    area, a missing square root or a transposed volume fails by far more than the 1e-10 here."""
    L, lf = 8, 4
    rng = np.random.default_rng(3)
    ms, ls = osh.pairs(L)
    f = (rng.normal(size=len(ms)) + 1j * rng.normal(size=len(ms))) * ((ls < lf) & (ls > 0))
    f = np.where(ms == 0, f.real + 0j, f)
    so = op.So3(L, 1)
    n = so.n
    nt, nphi = 24, 48
    x, wq = osh.gl(nt)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    st = np.sqrt(1 - x * x)
    D = np.stack(np.broadcast_arrays(st[:, None] * np.cos(phi), st[:, None] * np.sin(phi), x[:, None]), -1)
    wt = wq[:, None] * (2 * np.pi / nphi) * np.ones((1, nphi))
    cap = np.array([0.3, -0.2, 0.93])
    win = (D @ (cap / np.linalg.norm(cap)) > np.cos(0.7)).astype(float)
    area = np.sum(win * wt)

    def rotated_master(q):
        # f(S^-1 x) at the grid directions
        qi = op.qconj(q)
        u = np.asarray(qi[1:])
        t = 2 * np.cross(u, D)
        return _real_field(f, L, D + qi[0] * t + np.cross(u, t))

    ia, jb, ic = 3, 5, 11
    fS = rotated_master(op.zyz(2 * np.pi * ia / n, 2 * np.pi * jb / n, 2 * np.pi * ic / so.nc))
    g = win * (fS - np.sum(win * fS * wt) / area)
    gnorm = np.sqrt(np.sum(g * g * wt))
    x2, w2 = osh.gl(2 * L)
    s2 = osh.synthesize(f, L, x2, 4 * L)
    f2 = osh.analyze(s2 * s2, L, x2, w2)
    w_lm = osh.analyze(win, L, x, wq)
    score = so.volume(f, osh.analyze(g, L, x, wq)) * so.rden(so.volume(f, w_lm), so.volume(f2, w_lm), area) / gnorm
    assert abs(score[jb, ic, ia] - 1) < 1e-10
    assert np.unravel_index(np.argmax(score), score.shape) == (jb, ic, ia)
    for _ in range(4):
        jj, cc, aa = int(rng.integers(0, so.nb)), int(rng.integers(0, so.nc)), int(rng.integers(0, n))
        fq = rotated_master(op.zyz(2 * np.pi * aa / n, 2 * np.pi * jj / n, 2 * np.pi * cc / so.nc))
        mq = np.sum(win * fq * wt) / area
        ncc = np.sum(g * fq * wt) / np.sqrt(np.sum(win * (fq - mq) ** 2 * wt)) / gnorm
        assert abs(score[jj, cc, aa] - ncc) < 1e-10, (jj, cc, aa)


def test_single_cell_correlation_equals_the_volume():
    """So3.value_at (one cell from one row of b) against the whole volume, fold 1 and 4."""
    rng = np.random.default_rng(4)
    for L, fold in ((8, 1), (12, 4)):
        f, g = rand_coef(rng, L), rand_coef(rng, L)
        so = op.So3(L, fold)
        C = so.volume(f, g)
        part = op.So3(L, fold, rows=[1, 3])
        for _ in range(6):
            c, a = int(rng.integers(0, so.nc)), int(rng.integers(0, so.n))
            j = int(rng.integers(0, so.nb))
            assert abs(so.value_at(f, g, j, c, a) - C[j, c, a]) < 1e-9 * np.abs(C).max()
            assert abs(part.value_at(f, g, 3, c, a) - C[3, c, a]) < 1e-9 * np.abs(C).max()
