"""The product's host side (tables, masters, symmetry, geometry, settings) against the oracle and
closed forms; the reference's host tests restated (sh.zig:441-606, master.zig:755-850,
geometry.zig:110-150, sym.zig:1009-1059, settings.zig:492-519). No GPU."""
import numpy as np
import pytest

from oracle import pipeline as op
from oracle import sh as osh
from tests.conftest import GAN, NI
from zebsdsolve_amd import geometry, lambert, master, rot, sh, sym, tables
from zebsdsolve_amd.settings import SettingError, Settings


def test_gauss_legendre_matches_numpy():
    for n in (16, 64, 256):
        x, w = sh.gauss_legendre(n)
        xr, wr = osh.gl(n)
        assert np.abs(x - xr).max() < 1e-14 and np.abs(w - wr).max() < 1e-14


def test_legendre_recurrence_matches_scipy():
    x = np.linspace(-0.999, 0.999, 31)
    assert np.abs(sh.legendre_all(40, x) - osh.plm(40, x)).max() < 1e-12


def test_wigner_recurrence_matches_exact_and_eigen_routes():
    L = 24
    tab = sh.wigner_half_pi(L)
    for l in range(L):
        d = osh.wigner_d_matrix(l, np.pi / 2)
        for k in range(0, l + 1):
            for m in range(-l, l + 1):
                assert abs(sh.d_half_pi(tab, l, k, m) - d[k + l, m + l]) < 1e-12
    # at L = 64, a sample against the exact rational sum
    tab = sh.wigner_half_pi(64)
    rng = np.random.default_rng(0)
    for _ in range(40):
        l = int(rng.integers(0, 64))
        k = int(rng.integers(0, l + 1))
        m = int(rng.integers(-l, l + 1))
        assert abs(sh.d_half_pi(tab, l, k, m) - osh.wigner_exact(l, k, m)) < 1e-11


def test_grid_analysis_inverts_synthesis():
    rng = np.random.default_rng(7)
    L = 12
    ms, _ = sh.pairs_of(L)
    c = (rng.random(len(ms)) - 0.5) + 1j * np.where(ms == 0, 0, rng.random(len(ms)) - 0.5)
    g = sh.Grid.make(L, 16, 32)
    back = g.analyze(g.synthesize(c))
    assert np.abs(back - c).max() < 1e-12


def test_band_tables_layout():
    b = tables.BandHost(16)
    n = b.n
    assert n == 32 and b.ring.shape == (n, 2) and b.wplm.shape == (n // 2, b.pairs)
    # dq[m][l][k] = d^l_{k,m}(pi/2)
    assert b.dq[3, 7, 5] == np.float32(b.wigner[5, 3, 7])
    assert np.allclose(b.tw[256], [0.0, -1.0], atol=1e-7)
    # wplm: w_y P_l^m(x_y) 2 pi / n_phi
    p = sh.pair(16, 2, 9)
    assert abs(b.wplm[4, p] - b.grid.w[4] * b.grid.plm[p, 4] * 2 * np.pi / (2 * n)) < 1e-7


def test_volume_order_round_trip():
    n, nc = 16, 4
    h, nb = n // 2, n // 2 + 1
    nat = np.arange(nb * nc * n, dtype=np.float64).reshape(1, nb, nc, n)
    pos = np.empty((1, nb, h, nc, 2))
    for j in range(nb):
        for pp in range(h):
            for cp in range(nc):
                for k in range(2):
                    pos[0, j, pp, cp, k] = nat[0, j, tables.bitrev(cp, 2), 2 * tables.bitrev(pp, 3) + k]
    assert np.array_equal(tables.volume_natural(pos, n, nc), nat)


@pytest.fixture(scope="module")
def ni():
    return master.load(NI)


def test_nickel_master_expands_symmetric_and_normalized(ni):
    assert ni.side == 501 and ni.point_group == 32 and len(ni.ops) == 48 and len(ni.rotations) == 24
    assert abs(float(ni.raw.mean())) < 1e-4 and abs(float(ni.raw.std()) - 1) < 1e-3
    v = np.array([0.31, 0.52, 0.80])
    v /= np.linalg.norm(v)
    ref = lambert.at(ni.raw, ni.side, v)
    for g in sym.GROUPS["m-3m"]:
        w = rot.qapply(g, v)
        assert abs(lambert.at(ni.raw, ni.side, w) - ref) < 0.05
        assert abs(lambert.at(ni.raw, ni.side, -w) - ref) < 0.05


def test_gan_master_reduced_by_six_rotations_about_c():
    m = master.load(GAN)
    assert m.side == 301 and len(m.ops) == 12 and len(m.rotations) == 6
    for q in m.rotations:
        axis = q[1:] / max(np.linalg.norm(q[1:]), 1e-12)
        assert np.linalg.norm(q[1:]) < 1e-9 or abs(abs(axis[2]) - 1) < 1e-9
    rng = np.random.default_rng(7)
    z = 2 * rng.random(2000) - 1
    phi = 2 * np.pi * rng.random(2000)
    st = np.sqrt(1 - z * z)
    v = np.stack([st * np.cos(phi), st * np.sin(phi), z], 1)
    a = lambert.at(m.raw, m.side, v)
    for g in m.rotations:
        b = lambert.at(m.raw, m.side, rot.qapply(g, v))
        assert (a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()) > 0.97


def test_grid_symmetry_copies_equal_the_lookup():
    z = np.load(GAN, allow_pickle=False)
    kij = z["fundamental_kij"].astype(np.int64)
    vals = z["fundamental_sector"].reshape(-1)[: len(kij)].astype(np.float64)
    ops = z["pg_operators"].reshape(-1, 3, 3)
    nrm = z["fs_normals"]
    a = master.from_sector(vals, kij, 150, ops, nrm, by_grid=True)
    b = master.from_sector(vals, kij, 150, ops, nrm, by_grid=False)
    assert np.abs(a - b).max() < 1e-5


def test_harmonics_and_weights_match_the_oracle(ni):
    L = 16
    hm = master.harmonics(ni.side, ni.raw, L)
    f, f2, fold, inv = op.harmonics(ni.raw, ni.side, L)
    assert hm.rules.n_fold == 4 and hm.rules.mirror and hm.rules.inversion and fold == 4 and inv
    assert np.abs(hm.f - f).max() < 1e-9 * np.abs(f).max()
    assert np.abs(hm.f2 - f2).max() < 1e-9 * np.abs(f2).max()
    # W[mu, l, k] = i^{-m} f_lm d^l_{km}(pi/2), m = fold mu, from the oracle's d and coefficients
    tab = sh.wigner_half_pi(L)
    W = master.weights(L, hm.f, tab, 4)
    mu_max = (L - 1) // 4
    F = osh.signed(f, L)
    for mi in range(2 * mu_max + 1):
        m = 4 * (mi - mu_max)
        for l in range(L):
            d = osh.wigner_d_matrix(l, np.pi / 2)
            for k in range(L):
                ref = 0.0 if (l < max(k, abs(m))) else (1j ** ((-m) % 4)) * F[l, L - 1 + m] * d[k + l, m + l]
                assert abs(W[mi, l, k] - ref) < 1e-9


@pytest.mark.parametrize("path", [NI, GAN], ids=["Ni", "GaN"])
def test_gemm_table_reproduces_the_oracle_volume(path):
    """The matrix-core path's table, applied the way the MFMA applies it (row re times Re g plus
    row im times Im g, over the degree slots l >= m'), then the transform along a, equals the
    oracle's volume, whose b-axis rotation matrices come from the eigenvector route (not the
    factorization at pi/2 the table uses)."""
    m = master.load(path)
    L = 16
    band = tables.BandHost(L)
    ph = tables.PhaseHost(band, m.side, m.raw)
    vt = tables.VTableHost(band, ph, ph.hm.f)
    so = op.So3(L, ph.fold)
    rng = np.random.default_rng(4)
    g = rng.normal(size=band.pairs) + 1j * rng.normal(size=band.pairs)
    T = vt.table.astype(np.float64)                              # [j, m', s, 2, cols]
    nc, n = vt.nc, vt.n
    V = np.zeros((so.nb, L, nc), complex)
    for mp in range(L):
        for s_ in range(vt.slots):
            l = s_ * vt.l_step
            if l < mp:
                continue
            gv = g[sh.pair(L, mp, l)]
            out = gv.real * T[:, mp, s_, 0] + gv.imag * T[:, mp, s_, 1]   # [j, cols]
            V[:, mp] += out[:, 0:2 * nc:2] + 1j * out[:, 1:2 * nc:2]
    Ea = np.exp(-2j * np.pi * np.outer(np.arange(L), np.arange(n)) / n)
    k = np.where(np.arange(L) == 0, 1.0, 2.0)
    V[:, 0] = V[:, 0].real
    C = np.real(np.einsum("jpc,pa->jca", V * k[None, :, None], Ea))
    ref = so.volume(ph.hm.f, g)
    # float32 table entries: 1e-6 of the volume's scale
    assert np.abs(C - ref).max() < 1e-6 * np.abs(ref).max()
    # and in float64 (the table before rounding) the two routes agree to rounding
    assert vt.cols >= 2 * nc and vt.slots == (L + vt.l_step - 1) // vt.l_step


def test_bake_sigma_quantized_to_five_percent_steps():
    # 60 rows at pcz 0.55: a pixel is 1 / 33 rad; the 501-texel sheet spreads 2 pi sr over 501^2
    s = master.bake_sigma(501, 1.0 / 33.0, 0.3)
    px = (1.0 / 33.0) * 501 / np.sqrt(2 * np.pi)
    q = 1.05 ** round(np.log(px) / np.log(1.05))
    assert abs(s - 0.3 * q) < 1e-12 and abs(q / px - 1) < 0.026
    assert master.blur_passes(s) == [s]
    assert master.bake_sigma(501, 0.0, 0.3) == 0.0


def test_symmetry_groups_fundamental_zone_and_disorientation():
    for name, g in sym.GROUPS.items():
        assert len(g) == {"-1": 1, "2/m": 2, "mmm": 4, "4/m": 4, "4/mmm": 8, "-3": 3, "-3m": 6, "6/m": 6, "6/mmm": 12, "m-3": 12, "m-3m": 24}[name]
        for a in g:
            for b in g:
                c = rot.qstd(rot.qmul(a, b))
                assert np.any(np.abs(np.abs(g @ c) - 1) < 1e-9)
    rng = np.random.default_rng(3)
    ops = sym.GROUPS["m-3m"]
    q = rot.qstd(rng.normal(size=(50, 4)))
    f = sym.to_fz(q, ops)
    assert np.allclose(sym.to_fz(f, ops), f)
    # the cubic fundamental zone: w >= cos(62.8 deg / 2)
    assert np.all(f[:, 0] >= np.cos(np.radians(62.81) / 2) - 1e-4)
    g = ops[5]
    assert np.allclose(sym.disorientation(rot.qmul(np.broadcast_to(g, q.shape), q), q, ops), 0, atol=1e-7)
    r = rot.qstd(rng.normal(size=(50, 4)))
    assert np.allclose(sym.disorientation(q, r, ops), sym.disorientation(r, q, ops), atol=1e-9)
    assert np.all(sym.disorientation(q, r, ops) <= np.radians(62.81) + 1e-6)


def test_to_fz_is_the_largest_real_part_among_the_equivalents():
    """For every group: to_fz(q) is one of q's equivalents g . q (sign fixed, unit), its real
    part is at least every equivalent's (the smallest rotation angle), and the first operator
    reaching the maximum wins a tie; any leading shape, and none."""
    rng = np.random.default_rng(11)
    for name, ops in sym.GROUPS.items():
        q = rng.normal(size=(3, 17, 4)) * rng.uniform(0.5, 2.0, size=(3, 17, 1))
        f = sym.to_fz(q, ops)
        assert f.shape == q.shape
        cands = np.stack([rot.qstd(q)] + [rot.qstd(rot.qmul(np.broadcast_to(g, q.shape), q)) for g in ops])
        assert (f[..., 0] >= cands[..., 0].max(0) - 1e-12).all(), name
        assert (np.abs(cands - f[None]).max(-1).min(0) < 1e-12).all(), name
        assert np.allclose(np.linalg.norm(f, axis=-1), 1.0)
    ops = sym.GROUPS["m-3m"]
    # the identity and its equivalents tie at real part 1: the identity stays
    assert np.allclose(sym.to_fz(np.array([1.0, 0, 0, 0]), ops), [1, 0, 0, 0])
    # a quarter turn about z ties with the identity's equivalents too: the first operator that
    # reaches real part 1 (the identity) maps it
    qz = np.array([np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)])
    assert np.isclose(sym.to_fz(qz, ops)[0], 1.0)
    assert sym.to_fz(np.zeros((0, 4)), ops).shape == (0, 4)


def test_rotations_of_measured_operators():
    ops = np.array([rot.to_mat(q) for q in sym.GROUPS["6/m"]] + [-rot.to_mat(q) for q in sym.GROUPS["6/m"]])
    r = sym.rotations_of(ops)
    assert len(r) == 6 and np.allclose(r[0], [1, 0, 0, 0])
    # a set that is not a group reduces by nothing
    assert len(sym.rotations_of(ops[:5])) == 1


def test_pattern_centre_and_pose_are_two_views_of_one_thing():
    g = geometry.Geometry(pc=(0.42, 0.21, 0.5), rot=(0.01, -0.02, 0.005))
    R = g.rotation()
    t = g.translation(60, 80)
    assert np.allclose(geometry.Geometry.pc_of(R, t, 60, 80), g.pc, atol=1e-13)
    assert abs(R[:, 2] @ t - 0.5 * 60) < 1e-10
    g.rot = (0, 0, 0)
    t0 = g.translation(60, 60)
    a = np.radians(-70.0)
    assert abs(0.21 * 60 * np.sin(a) + 0.5 * 60 * np.cos(a) - np.sin(a) / 2 - t0[0]) < 1e-10
    assert abs(-60 * 0.42 + 0.5 - t0[1]) < 1e-10
    assert abs(0.21 * 60 * np.cos(a) - 0.5 * 60 * np.sin(a) - np.cos(a) / 2 - t0[2]) < 1e-10
    h = geometry.Geometry(rot=(0, np.radians(3.0), 0))
    assert np.allclose(h.rotation(), geometry.Geometry(detector_tilt_deg=3).rotation(), atol=1e-12)


def test_beam_travel_shifts_the_detector_as_kikuchipy_extrapolates():
    g = geometry.Geometry(pc=(0.42, 0.21, 0.5), px_um=70, step_x_um=1.5, step_y_um=1.5)
    R = g.rotation()
    t = g.translation(60, 60)
    s = g.scan_offset(1, 1, 0, 0, -1, -1, 8)
    pc = geometry.Geometry.pc_of(R, t - s, 60, 60)
    f = 1.5 / (70.0 * 8.0 * 60.0)
    alpha = np.radians(20.0)
    assert abs(pc[0] - (0.42 - f)) < 1e-12 and abs(pc[1] - (0.21 - f * np.cos(alpha))) < 1e-12 and abs(pc[2] - (0.5 + f * np.sin(alpha))) < 1e-12


def test_settings_reject_unknown_and_out_of_range_atomically():
    s = Settings()
    s.apply(bandwidth=88, refine_iters=5)
    assert s["bandwidth"] == 88
    with pytest.raises(SettingError):
        s.apply(nope=1)
    with pytest.raises(SettingError):
        s.apply(bandwidth=4, refine_iters=9)
    assert s["refine_iters"] == 5
    with pytest.raises(SettingError):
        s.apply(static_background="gaussian")
    with pytest.raises(SettingError):
        s.apply(dynamic_background="gaussian")
    with pytest.raises(SettingError):
        s.apply(nlpar=True)
    # CLAHE and the band pass: the reference's ranges (settings.zig:430, 456)
    s.apply(clahe=True, bandpass=True)
    for bad in ({"clahe_grid": 17}, {"clahe_bins": 1}, {"clahe_clip": 0.0}, {"bp_fine": 8.0}, {"bp_coarse": 0.5}):
        with pytest.raises(SettingError):
            s.apply(**bad)
    # the reference's defaults for the filters (settings.zig:138-159)
    d = Settings()
    assert d["static_background"] == "average" and d["dynamic_background"] == "subtract"
    assert d["static_patterns"] == 1024 and d["dynamic_sigma"] == 0.15
    assert (d["clahe"], d["clahe_grid"], d["clahe_bins"], d["clahe_clip"]) == (False, 4, 64, 40.0)
    assert (d["bandpass"], d["bp_fine"], d["bp_coarse"]) == (False, 1.0, 8.0)


def test_sweeps_climb_to_the_final_bandwidth():
    """settings.zig:511-519: three sweeps from 32 to 128 run at 32, 64, 128; one sweep at the
    bandwidth; the count and first bandwidth are range-checked."""
    s = Settings()
    s.apply(sweeps=3, sweep_bandwidth=32, bandwidth=128)
    assert [s.sweep_bandwidth_of(i) for i in range(3)] == [32, 64, 128]
    s.apply(sweeps=2, sweep_bandwidth=32, bandwidth=64)
    assert [s.sweep_bandwidth_of(i) for i in range(2)] == [32, 64]
    s.apply(sweeps=4, sweep_bandwidth=32, bandwidth=64)
    # equal ratios: 32 * 2^(1/3) = 40.3, 32 * 2^(2/3) = 50.8
    assert [s.sweep_bandwidth_of(i) for i in range(4)] == [32, 40, 51, 64]
    s.apply(sweeps=1)
    assert s.sweep_bandwidth_of(0) == 64
    for bad in ({"sweeps": 0}, {"sweeps": 7}, {"sweep_bandwidth": 4}, {"sweep_method": "pca"}):
        with pytest.raises(SettingError):
            s.apply(**bad)


def test_rotation_conversions_round_trip():
    rng = np.random.default_rng(1)
    q = rot.qstd(rng.normal(size=(100, 4)))
    assert np.allclose(rot.qstd(rot.from_bunge(rot.to_bunge(q))), q, atol=1e-10)
    for qq in q[:10]:
        assert np.allclose(rot.from_mat(rot.to_mat(qq)), qq, atol=1e-12)
    v = rng.normal(size=3)
    assert np.allclose(rot.to_mat(q[0]) @ v, rot.qapply(q[0], v))


def test_six_direction_local_fit_matches_the_oracle():
    """fit.local_of (batched, float64) against the oracle's localOf restatement, which its own
    test pins to central differences of the NCC."""
    from zebsdsolve_amd import fit
    rng = np.random.default_rng(8)
    # sums of a random set of pixels (M, e, d) so that the variances are positive
    batch = []
    for _ in range(5):
        M = rng.normal(size=500)
        e = rng.normal(size=500)
        d = rng.normal(size=(500, 6))
        s = [500, M.sum(), (M * M).sum(), (M * e).sum()] + list(d.sum(0)) + list((M[:, None] * d).sum(0)) + list((e[:, None] * d).sum(0))
        dd = d.T @ d
        s += [dd[i, j] for i in range(6) for j in range(i, 6)]
        batch.append(s)
    batch = np.asarray(batch)
    ok, ncc, g, H = fit.local_of(batch)
    for b in range(5):
        ok1, n1, g1, H1 = op.local6(batch[b])
        assert ok[b] and ok1
        assert abs(ncc[b] - n1) < 1e-12 and np.abs(g[b] - g1).max() < 1e-12 and np.abs(H[b] - H1).max() < 1e-12
    flat = batch[0].copy()
    flat[2] = flat[1] ** 2 / flat[0]
    assert not fit.local_of(flat[None])[0][0]


def test_left_jacobian_of_exp():
    from zebsdsolve_amd import fit
    rng = np.random.default_rng(9)
    for phi in [np.zeros(3), rng.normal(size=3) * 0.3, rng.normal(size=3) * 2.0]:
        J = fit.left_jacobian(phi[None])[0]
        h = 1e-6
        for i in range(3):
            dphi = np.zeros(3)
            dphi[i] = h
            q = rot.qmul(rot.so3_exp(phi + dphi), rot.qconj(rot.so3_exp(phi)))
            q = q * np.sign(q[0])
            # the rotation vector of exp(phi + h e_i) exp(phi)^-1 is J h e_i to first order
            assert np.abs(2 * q[1:] / h - J[:, i]).max() < 1e-5


def test_solver_surface_returns_the_reference_shapes():
    """The Solver calls a drop-in caller unpacks (python/zebsdsolve/__init__.py): reduce returns
    (quaternions (n, 4) float32, euler (n, 3) float64), each quaternion the largest-w equivalent
    with its Bunge angles; sweep_plan lists (bandwidth, dictionary grid, method) a sweep with
    equal ratios (src/settings.zig:465-483); events() drains the queue; cancel() makes the next
    index_step return 0 and index_begin clears it."""
    import zebsdsolve_amd as ze
    from zebsdsolve_amd import rot
    s = ze.Solver()
    assert s.add_phase(NI) == 0
    ev = s.events()
    assert ev and ev[0]["kind"] == "log" and ev[0]["text"].startswith("phase 0:")
    assert s.events() == []
    rng = np.random.default_rng(1)
    q = rot.qstd(rng.normal(size=(5, 4)))
    fz, eu = s.reduce(q)
    assert fz.dtype == np.float32 and fz.shape == (5, 4) and eu.dtype == np.float64 and eu.shape == (5, 3)
    rots = s._e.phases[0].master.rotations
    for k in range(5):
        # the representative is an equivalent of q with the largest real part
        cands = rot.qmul(np.asarray(rots, np.float64), np.repeat(q[k:k + 1], len(rots), 0))
        assert abs(abs(fz[k, 0]) - np.abs(cands[:, 0]).max()) < 1e-6
        assert np.allclose(rot.from_bunge(eu[k:k + 1])[0] * np.sign(fz[k, 0]), fz[k], atol=1e-5)
    q1, e1 = s.reduce(q[0])
    assert q1.shape == (1, 4) and e1.shape == (1, 3)
    s.configure(sweeps=3, sweep_bandwidth=32, bandwidth=64)
    assert s.sweep_plan == [(32, 0, "spherical"), (45, 0, "spherical"), (64, 0, "spherical")]
    s.configure(sweeps=1)
    assert s.sweep_plan == [(64, 0, "spherical")]
    s.set_patterns(np.zeros((2, 3, 60, 60), np.float32))
    s.index_begin()
    s.cancel()
    assert s.index_step() == 0 and s.index_status["done"] == 0
    s.index_begin()
    assert not s._cancel
    with pytest.raises(ze.Error):
        s.reduce(q, phase=3)


def test_correlation_launches_are_equal_whole_tiles():
    """A batch's correlation launches: the fewest that fit the workspace, of equal whole 128-item
    tiles (the last one the rest), never past the workspace's items."""
    from zebsdsolve_amd.engine import correlation_part
    for n in (1, 129, 1000, 7300, 7301, 7500, 8000, 14700, 16384):
        for chunk in (100, 1950, 7300, 16384):
            part = correlation_part(n, chunk)
            sizes = [min(part, n - c0) for c0 in range(0, n, part)]
            assert sum(sizes) == n and max(sizes) <= chunk
            assert len(sizes) == -(-n // chunk), (n, chunk, sizes)
            if chunk >= 128 and len(sizes) > 1:
                assert part % 128 == 0
    assert correlation_part(7500, 7300) == 3840  # not 7300 + 200
