"""The matrix-core correlation (zb_vcorr_*) against the float64 oracle and against the slab path
(zb_so3_*), for a centrosymmetric cubic master (fold 4, even l only) and a hexagonal one without
inversion (fold 2, every l), at L = 32 and at the benchmark's L = 64.

Tolerances: float32 GEMM and transforms against float64, stated at each assertion.
"""
import numpy as np
import pytest

from tests import cases
from tests.conftest import GAN, NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

# the bench's bandwidth, and the others of the two-phase workload at n = 128 (L = 48, 64)
CASES = [("Ni", NI, 32, 10), ("GaN", GAN, 32, 10), ("Ni", NI, 64, 4), ("Ni", NI, 48, 4), ("GaN", GAN, 48, 4), ("GaN", GAN, 64, 3)]


def _setup(path, L, N):
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import engine
    from zebsdsolve_amd.geometry import Geometry

    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")  # the path alone, as the oracle
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(path)
    e.settings.apply(bandwidth=L)
    e.set_patterns(np.zeros((1, N, 60, 60), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, truth = cases.simulate(host, side, N, seed=21)
    e.set_patterns(pats.reshape(1, N, 60, 60))
    fr, t = cases.oracle_frame(60, 60)
    torch.cuda.synchronize()
    return e, op, host, side, pats, truth, fr, t


def _coefficients(e, L, pats):
    import torch
    from zebsdsolve_amd import _native as nat
    from zebsdsolve_amd.engine import stream_ptr
    lib = nat.load()
    N = pats.shape[0]
    band = e.band(L)
    raw = torch.from_numpy(pats).cuda()
    pn = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, N, 60, 60, mask, pn.data_ptr(), stream_ptr()), "normalize")
    t = e.positions(np.arange(N))
    items = torch.from_numpy(np.concatenate([t, np.arange(N, dtype=np.float32)[:, None]], 1)).cuda()
    alm = torch.empty((N, band.host.pairs, 2), dtype=torch.float32, device="cuda")
    stats = torch.empty((N, 4), dtype=torch.float32, device="cuda")
    ws = torch.empty((lib.zb_sh_workspace(band.c, N),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_sh_coefficients(band.c, e.frame(), pn.data_ptr(), items.data_ptr(), N, 3, ws.data_ptr(), ws.numel(), alm.data_ptr(),
                                     stats.data_ptr(), stream_ptr()), "coefficients")
    return alm, pn.cpu().numpy()


@pytest.mark.parametrize("name,path,L,N", CASES, ids=[f"{c[0]}-L{c[2]}" for c in CASES])
def test_vcorr_against_oracle_and_slab_path(name, path, L, N):
    import torch
    from zebsdsolve_amd import _native as nat, sym
    from zebsdsolve_amd.engine import stream_ptr
    print(f"vcorr case {name} L={L} N={N}", flush=True)
    e, op, host, side, pats, truth, fr, t = _setup(path, L, N)
    lib = nat.load()
    ph = e.phase_band(0, L)
    band = e.band(L)
    f, f2, fold, inv = op.harmonics(host, side, L)
    assert fold == ph.host.fold and (ph.host.l_step == 2) == inv
    so = op.So3(L, fold)
    wa, wst = op.coefficients(None, fr, t, L, center=False, norm=False, window=True)
    rden_ref = so.rden(so.volume(f, wa), so.volume(f2, wa), wst[0])          # [j][c][a]
    rden = e.vrden(0, L).cpu().numpy().reshape(so.n // 2 + 1, so.nc, so.n)  # [j][c][a]
    big = rden_ref > 0
    # the same cells are defined; reciprocal denominators agree to float32 GEMM and transforms
    assert np.mean((rden > 0) == big) > 0.999
    rel = np.abs(rden[big] - rden_ref[big]) / rden_ref[big]
    assert np.median(rel) < 1e-4 and np.percentile(rel, 99) < 2e-3, (np.median(rel), np.percentile(rel, 99))

    alm, pats_n = _coefficients(e, L, pats)
    (vt, _), _ = ph.vtables()
    peaks = torch.empty((N, 8), dtype=torch.float32, device="cuda")
    ws = torch.empty((lib.zb_vcorr_workspace(vt, N, 0),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_vcorr_correlate(vt, band.tw.data_ptr(), alm.data_ptr(), N, e.vrden(0, L).data_ptr(), N, ws.data_ptr(), ws.numel(),
                                     peaks.data_ptr(), stream_ptr()), "vcorr")
    pk = peaks.cpu().numpy()
    # the slab path on the same coefficients
    peaks2 = torch.empty((N, 8), dtype=torch.float32, device="cuda")
    ws2 = torch.empty((lib.zb_so3_workspace(band.c, ph.c, N),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_so3_correlate(band.c, ph.c, alm.data_ptr(), N, e.rden(0, L).data_ptr(), N, ws2.data_ptr(), ws2.numel(), peaks2.data_ptr(),
                                   stream_ptr()), "so3")
    pk2 = peaks2.cpu().numpy()
    rots = e.phases[0].master.rotations
    step = 360.0 / so.n
    for b in range(N):
        ref_alm, _ = op.coefficients(pats_n[b].astype(np.float64), fr, t, L)
        C = so.volume(f, ref_alm)
        q, v0, ang, idx = so.peak(C, rden_ref)
        # the oracle's peak: score to float32 accuracy, orientation within 1/100 of a grid step
        assert abs(pk[b, 4] - v0) < 2e-4 * abs(v0), (b, pk[b, 4], v0)
        d = np.degrees(sym.disorientation(pk[b, :4].astype(np.float64), q, rots))
        assert d < 0.01 * step, (b, d)
        # the slab path: the same cell, score and sub-cell offsets to float32 rounding
        assert abs(pk[b, 4] - pk2[b, 4]) < 2e-5 * abs(pk2[b, 4]), (b, pk[b, 4], pk2[b, 4])
        d2 = np.degrees(sym.disorientation(pk[b, :4].astype(np.float64), pk2[b, :4].astype(np.float64), rots))
        assert d2 < 0.005 * step, (b, d2)


def test_vcorr_argument_checks():
    import torch
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    bad = nat.VTable(128, 512, 64, 2, 64, 128, 0)  # n = 512: past the register transform
    assert lib.zb_vcorr_workspace(bad, 4, 0) < 0
    assert "n <= 256" in nat.last_error()
    x = torch.zeros(16, device="cuda")
    good = nat.VTable(32, 64, 16, 2, 16, 32, x.data_ptr())
    need = lib.zb_vcorr_workspace(good, 4, 0)
    assert need > 0
    rc = lib.zb_vcorr_correlate(good, x.data_ptr(), x.data_ptr(), 4, x.data_ptr(), 4, x.data_ptr(), need - 1, x.data_ptr(), None)
    assert rc == -2 and "workspace" in nat.last_error()


@pytest.mark.parametrize("L,N", [(32, 300), (64, 200), (64, 1500), (96, 301)])
def test_vcorr_matches_slab_path_over_many_tiles(L, N):
    """Several item tiles and a ragged last one (N not a multiple of 128), with many blocks in
    flight at once: random unit-norm coefficients through both correlators, and the matrix-core
    one run twice gives the same bits."""
    import torch
    from zebsdsolve_amd import _native as nat, engine, sym
    from zebsdsolve_amd.geometry import Geometry
    from zebsdsolve_amd.engine import stream_ptr
    print(f"vcorr tiles L={L} N={N}", flush=True)
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")  # the path alone, as the oracle
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.settings.apply(bandwidth=L)
    e.set_patterns(np.zeros((1, 1, 60, 60), np.float32))
    lib = nat.load()
    ph = e.phase_band(0, L)
    band = e.band(L)
    (vt, _), _ = ph.vtables()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    g = torch.randn((N, band.host.pairs, 2), device="cuda", generator=gen)
    g /= g.norm(dim=(1, 2), keepdim=True)
    p1 = torch.empty((N, 8), device="cuda")
    p2 = torch.empty((N, 8), device="cuda")
    ws = torch.empty((lib.zb_vcorr_workspace(vt, N, 0),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_vcorr_correlate(vt, band.tw.data_ptr(), g.data_ptr(), N, e.vrden(0, L).data_ptr(), N, ws.data_ptr(), ws.numel(),
                                     p1.data_ptr(), stream_ptr()), "vcorr")
    ws2 = torch.empty((lib.zb_so3_workspace(band.c, ph.c, N),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_so3_correlate(band.c, ph.c, g.data_ptr(), N, e.rden(0, L).data_ptr(), N, ws2.data_ptr(), ws2.numel(), p2.data_ptr(),
                                   stream_ptr()), "so3")
    a, b = p1.cpu().numpy(), p2.cpu().numpy()
    p3 = torch.empty_like(p1)
    nat.check(lib.zb_vcorr_correlate(vt, band.tw.data_ptr(), g.data_ptr(), N, e.vrden(0, L).data_ptr(), N, ws.data_ptr(), ws.numel(),
                                     p3.data_ptr(), stream_ptr()), "vcorr again")
    assert np.array_equal(a, p3.cpu().numpy())
    # the best score of every item to float32 rounding of the two summation orders
    assert np.all(np.abs(a[:, 4] - b[:, 4]) <= 2e-5 * np.abs(b[:, 4])), np.abs(a[:, 4] - b[:, 4]).max()
    # the same orientation, except where two cells score within rounding of each other
    d = np.degrees(sym.disorientation(a[:, :4].astype(np.float64), b[:, :4].astype(np.float64), e.phases[0].master.rotations))
    assert np.mean(d < 0.01) > 0.98, np.sort(d)[-5:]


@pytest.mark.parametrize("L,precision", [(48, "fp32"), (64, "fp32"), (64, "fp16"), (96, "fp32")])
def test_half_b_search_matches_the_whole_range(L, precision):
    """Nickel has a 2-fold axis along y, so its correlation obeys C(a, b, c) = C(a + pi, pi - b,
    pi - c): searching j <= n / 4 (b_half) finds, for every pattern, the whole range's peak or
    its partner under that axis. On 48 simulated patterns (some of them with their peak past
    b = pi / 2 over the whole range): the same orientation up to the master's symmetry
    within 1/200 of a grid step, the same score to float32 rounding, every peak at b <= pi / 2
    plus half a step; and the oracle's peak, as for the whole range.""" 
    import torch
    from zebsdsolve_amd import _native as nat, sym
    from zebsdsolve_amd.engine import stream_ptr
    print(f"half b L={L} {precision}", flush=True)
    N = 48
    e, op, host, side, pats, truth, fr, t = _setup(NI, L, N)
    e.settings.apply(correlation_precision=precision)
    lib = nat.load()
    ph = e.phase_band(0, L)
    band = e.band(L)
    assert ph.y2
    alm, pats_n = _coefficients(e, L, pats)
    out = {}
    for half in (0, 1):
        e.settings.apply(correlation_symmetry=bool(half))
        vt = e.corr_vtable(ph)
        assert vt.b_half == half
        pk = torch.empty((N, 8), dtype=torch.float32, device="cuda")
        ws = torch.empty((lib.zb_vcorr_workspace(vt, N, 0),), dtype=torch.uint8, device="cuda")
        nat.check(lib.zb_vcorr_correlate(vt, band.tw.data_ptr(), alm.data_ptr(), N, e.vrden(0, L).data_ptr(), N, ws.data_ptr(), ws.numel(),
                                         pk.data_ptr(), stream_ptr()), "vcorr")
        out[half] = pk.cpu().numpy()
    full, half = out[0], out[1]
    rots = e.phases[0].master.rotations
    n = band.host.n
    step = 360.0 / n
    # over the whole range the lowest index wins a tie, so most peaks already sit at b <= pi / 2;
    # those the whole range puts past it (its float32 rounding favoured the partner) are here
    past = full[:, 6] > np.pi / 2 + 0.5 * 2 * np.pi / n
    if precision == "fp32":
        assert past.sum() >= 2, past.sum()
    assert (half[:, 6] <= np.pi / 2 + 0.5 * 2 * np.pi / n + 1e-6).all()
    d = np.degrees(sym.disorientation(half[:, :4].astype(np.float64), full[:, :4].astype(np.float64), rots))
    if precision == "fp32":
        assert d.max() < 0.005 * step, d.max()
        assert np.all(np.abs(half[:, 4] - full[:, 4]) <= 2e-5 * np.abs(full[:, 4])), np.abs(half[:, 4] - full[:, 4]).max()
    else:
        # a partner cell's score comes from other rows of the 16-bit V (rounded apart by up to
        # ~1e-4), which can reorder a near-tie
        assert np.mean(d < 0.005 * step) >= 0.95, d
        assert np.all(np.abs(half[:, 4] - full[:, 4]) <= 3e-4 * np.abs(full[:, 4])), np.abs(half[:, 4] - full[:, 4]).max()
    if precision == "fp32" and L <= 64:
        f, f2, fold, _ = op.harmonics(host, side, L)
        so = op.So3(L, fold)
        wa, wst = op.coefficients(None, fr, t, L, center=False, norm=False, window=True)
        rden_ref = so.rden(so.volume(f, wa), so.volume(f2, wa), wst[0])
        for b in np.flatnonzero(past)[:3]:
            ref_alm, _ = op.coefficients(pats_n[b].astype(np.float64), fr, t, L)
            q, v0, _, _ = so.peak(so.volume(f, ref_alm), rden_ref)
            assert abs(half[b, 4] - v0) < 2e-4 * abs(v0), (b, half[b, 4], v0)
            assert np.degrees(sym.disorientation(half[b, :4].astype(np.float64), q, rots)) < 0.01 * step
