"""The HIP path, stage by stage, against the float64 oracle on the same seeded inputs, and end to
end against the known truth (thresholds of the reference's own synthetic test,
tests/test_api.py:206-219: median < 0.15 deg, max < 1 deg; peak alone median < 1.5 deg).

Tolerances: the kernels compute in float32 and the oracle in float64, so every comparison has
one, stated at the assertion.
"""
import numpy as np
import pytest

from tests import cases
from tests.conftest import NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

L = 32
N = 12


@pytest.fixture(scope="module")
def case():
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import engine, sym
    from zebsdsolve_amd.geometry import Geometry

    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")  # the path alone, as the oracle
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.set_patterns(np.zeros((1, N, 60, 60), np.float32))
    key, (sheets_d, grad_d, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, truth = cases.simulate(host, side, N, seed=11)
    e.set_patterns(pats.reshape(1, N, 60, 60))
    fr, t = cases.oracle_frame(60, 60)
    torch.cuda.synchronize()
    return {"e": e, "op": op, "sym": sym, "torch": torch, "sheets": host, "side": side, "pats": pats, "truth": truth, "fr": fr, "t": t,
            "grad": grad_d.cpu().numpy().reshape(-1, 4), "raw_sheets": e.phases[0].master.raw}


def test_normalize(case):
    import torch
    from zebsdsolve_amd import _native as nat
    e, op = case["e"], case["op"]
    lib = nat.load()
    raw = torch.from_numpy(case["pats"]).cuda()
    out = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, N, 60, 60, mask, out.data_ptr(), engine_stream()), "normalize")
    got = out.cpu().numpy()
    ref = np.stack([op.normalize(p.astype(np.float64), case["fr"]["mask"]) for p in case["pats"]])
    # unit-norm patterns of 2828 pixels: values ~ 0.02; float32 sums: 1e-6 absolute
    assert np.abs(got - ref).max() < 2e-6


def engine_stream():
    from zebsdsolve_amd import engine
    return engine.stream_ptr()


def test_master_blur_and_gradient(case):
    import torch
    from zebsdsolve_amd import _native as nat, master as mm
    op, side = case["op"], case["side"]
    lib = nat.load()
    raw = case["raw_sheets"]
    e = case["e"]
    sigma = mm.bake_sigma(side, e.pixel_angle(), 0.3)
    assert sigma > 0.05
    px = np.sqrt(2 * np.pi) / side
    src = torch.from_numpy(np.ascontiguousarray(raw)).cuda().reshape(-1)
    dst = torch.empty_like(src)
    s0 = mm.blur_passes(sigma)[0]
    nat.check(lib.zb_master_blur(src.data_ptr(), dst.data_ptr(), side, float(s0 * px), engine_stream()), "blur")
    got = dst.cpu().numpy()
    rng = np.random.default_rng(1)
    tex = rng.choice(2 * side * side, 3000, replace=False)
    ref = op.blur(raw, side, float(np.float32(s0 * px)), tex)
    # unit-variance master, float32 trigonometry on the lattice: 2e-4 absolute
    assert np.abs(got[tex] - ref).max() < 2e-4
    g = op.gradient(case["sheets"], side, tex)
    gg = case["grad"][tex]
    # a difference quotient one texel (0.005 rad) apart: float32 rounding of the two lookups
    # (1e-7 relative of values ~3) over 0.01 rad gives ~1e-4; allow 2e-3 of the field's scale
    scale = np.abs(g[:, :3]).max()
    assert np.abs(gg[:, :3] - g[:, :3]).max() < 2e-3 * scale
    assert np.abs(gg[:, 3] - g[:, 3]).max() == 0.0


def test_master_pack_rounds_the_field_to_half(case):
    """zb_master_pack: every value of the gradient field rounded to the nearest half (ties to
    even), bit for bit as numpy's float16 conversion, each texel followed by the one below it
    (the last row of a sheet by itself)."""
    e = case["e"]
    g = case["grad"].reshape(2, -1, 4)
    side = int(round(np.sqrt(g.shape[1])))
    gh = e.grad_half(0).cpu().numpy().view(np.uint16).reshape(2, side, side, 2, 4)
    ref = g.reshape(2, side, side, 4).astype(np.float16).view(np.uint16)
    below = np.concatenate([ref[:, 1:], ref[:, -1:]], axis=1)
    assert np.array_equal(gh[:, :, :, 0], ref) and np.array_equal(gh[:, :, :, 1], below)


def _gpu_coefficients(case, flags=3, window=False):
    import torch
    from zebsdsolve_amd import _native as nat
    e = case["e"]
    lib = nat.load()
    band = e.band(L)
    dev = torch.device("cuda")
    raw = torch.from_numpy(case["pats"]).cuda()
    pats = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, N, 60, 60, mask, pats.data_ptr(), engine_stream()), "normalize")
    t = e.positions(np.arange(N))
    items = torch.from_numpy(np.concatenate([t, np.arange(N, dtype=np.float32)[:, None]], 1)).to(dev)
    alm = torch.empty((N, band.host.pairs, 2), dtype=torch.float32, device=dev)
    stats = torch.empty((N, 4), dtype=torch.float32, device=dev)
    ws = torch.empty((lib.zb_sh_workspace(band.c, N),), dtype=torch.uint8, device=dev)
    nat.check(lib.zb_sh_coefficients(band.c, e.frame(), pats.data_ptr(), items.data_ptr(), N, flags, ws.data_ptr(), ws.numel(),
                                     alm.data_ptr(), stats.data_ptr(), engine_stream()), "coefficients")
    a = alm.cpu().numpy()
    return a[..., 0] + 1j * a[..., 1], stats.cpu().numpy(), pats.cpu().numpy()


def test_coefficients(case):
    op = case["op"]
    alm, stats, pats_n = _gpu_coefficients(case)
    for b in range(N):
        ref, st = op.coefficients(pats_n[b].astype(np.float64), case["fr"], case["t"], L)
        err = np.linalg.norm(alm[b] - ref) / np.linalg.norm(ref)
        # unit-norm coefficients from 8192 float32 samples through a float32 FFT and sums
        assert err < 2e-5, (b, err)
        assert abs(stats[b, 0] - st[0]) < 1e-5 * st[0]
        assert abs(stats[b, 2] - st[2]) < 1e-4 * st[2]


@pytest.mark.parametrize("Lx", [128, 160])
def test_coefficients_at_large_bandwidths(case, Lx):
    """The coefficients at bandwidths past the unrolled forms of the sphere stage: L = 128 (2 x 256
    rings of 512 points; the unit norm over a thread's coefficients read twice, past the forms
    that keep them in registers) and L = 160 (2 x 512 rings of 1 024 points, the ring FFT at a
    line length given at run time), against the oracle."""
    import torch
    from zebsdsolve_amd import _native as nat
    op, e = case["op"], case["e"]
    print(f"coefficients at L = {Lx}", flush=True)
    lib = nat.load()
    band = e.band(Lx)
    dev = torch.device("cuda")
    k = 2
    raw = torch.from_numpy(case["pats"][:k]).cuda()
    pats = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, k, 60, 60, mask, pats.data_ptr(), engine_stream()), "normalize")
    t = e.positions(np.arange(k))
    items = torch.from_numpy(np.concatenate([t, np.arange(k, dtype=np.float32)[:, None]], 1)).to(dev)
    alm = torch.empty((k, band.host.pairs, 2), dtype=torch.float32, device=dev)
    stats = torch.empty((k, 4), dtype=torch.float32, device=dev)
    ws = torch.empty((lib.zb_sh_workspace(band.c, k),), dtype=torch.uint8, device=dev)
    nat.check(lib.zb_sh_coefficients(band.c, e.frame(), pats.data_ptr(), items.data_ptr(), k, 3, ws.data_ptr(), ws.numel(), alm.data_ptr(),
                                     stats.data_ptr(), engine_stream()), "coefficients")
    a = alm.cpu().numpy()
    got = a[..., 0] + 1j * a[..., 1]
    pn = pats.cpu().numpy()
    for b in range(k):
        ref, st = op.coefficients(pn[b].astype(np.float64), case["fr"], case["t"], Lx)
        err = np.linalg.norm(got[b] - ref) / np.linalg.norm(ref)
        print(f"L = {Lx}, pattern {b}: relative error {err:.2e}", flush=True)
        assert err < 5e-5, (Lx, b, err)
        assert abs(stats[b, 2] - st[2]) < 1e-4 * st[2]


def test_denominators_and_correlation(case):
    from zebsdsolve_amd.tables import volume_natural
    op, e = case["op"], case["e"]
    ph = e.phase_band(0, L)
    f, f2, fold, inv = op.harmonics(case["sheets"], case["side"], L)
    assert fold == ph.host.fold == 4 and inv
    so = op.So3(L, fold)
    wa, wst = op.coefficients(None, case["fr"], case["t"], L, center=False, norm=False, window=True)
    rden_ref = so.rden(so.volume(f, wa), so.volume(f2, wa), wst[0])
    rden = volume_natural(e.rden(0, L).cpu().numpy(), so.n, so.nc)[0]
    big = rden_ref > 0
    # the same cells are defined, and the reciprocal denominators agree to float32 transforms
    assert np.mean((rden > 0) == big) > 0.999
    rel = np.abs(rden[big] - rden_ref[big]) / rden_ref[big]
    assert np.median(rel) < 1e-4 and np.percentile(rel, 99) < 2e-3
    # the correlation peaks
    import torch
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    alm, _, pats_n = _gpu_coefficients(case)
    band = e.band(L)
    a = torch.from_numpy(np.stack([alm.real, alm.imag], -1).astype(np.float32)).cuda()
    peaks = torch.empty((N, 8), dtype=torch.float32, device="cuda")
    ws = torch.empty((lib.zb_so3_workspace(band.c, ph.c, N),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_so3_correlate(band.c, ph.c, a.data_ptr(), N, e.rden(0, L).data_ptr(), N, ws.data_ptr(), ws.numel(), peaks.data_ptr(),
                                   engine_stream()), "correlate")
    pk = peaks.cpu().numpy()
    sym = case["sym"]
    rots = e.phases[0].master.rotations
    for b in range(N):
        ref_alm, _ = op.coefficients(pats_n[b].astype(np.float64), case["fr"], case["t"], L)
        C = so.volume(f, ref_alm)
        q, v0, ang, idx = so.peak(C, rden_ref)
        # the same peak: its score to float32 accuracy, its orientation within a hundredth of a
        # grid step (360/64 deg)
        assert abs(pk[b, 4] - v0) < 2e-4 * abs(v0), (b, pk[b, 4], v0)
        d = np.degrees(sym.disorientation(pk[b, :4].astype(np.float64), q, rots))
        assert d < 0.06, (b, d)


def test_end_to_end_and_refinement(case):
    op, e, sym = case["op"], case["e"], case["sym"]
    rots = e.phases[0].master.rotations
    e.settings.apply(bandwidth=L)
    # peak alone (refine off)
    e.settings.apply(refine=False)
    r0 = e.index_cells(np.arange(N))[0]
    d0 = np.degrees(sym.disorientation(r0[:, :4].astype(np.float64), case["truth"], rots))
    assert np.median(d0) < 1.5 and np.percentile(d0, 95) < 3.0, d0
    e.settings.apply(refine=True)
    rec = e.index_cells(np.arange(N))[0]
    d = np.degrees(sym.disorientation(rec[:, :4].astype(np.float64), case["truth"], rots))
    assert np.median(d) < 0.15 and d.max() < 1.0, d
    # the refinement and NMI against the oracle from the same seeds (the GPU's peaks)
    _, _, pats_n = _gpu_coefficients(case)
    grad = case["grad"]
    for b in range(N):
        q, ncc, step = op.refine(grad, case["side"], case["fr"], pats_n[b].astype(np.float64), case["t"], r0[b, :4].astype(np.float64),
                                 iters=8)
        dq = np.degrees(sym.disorientation(rec[b, :4].astype(np.float64), q, rots))
        assert dq < 0.02, (b, dq)
        assert abs(rec[b, 4] - ncc) < 1e-3, (b, rec[b, 4], ncc)
        nm = op.nmi(case["sheets"], case["side"], case["fr"], pats_n[b].astype(np.float64), case["t"], rec[b, :4].astype(np.float64))
        assert abs(rec[b, 5] - nm) < 2e-3, (b, rec[b, 5], nm)


@pytest.mark.parametrize("h,w,mask_on", [(40, 56, False), (80, 80, True)], ids=["40x56-nomask", "80x80-mask"])
def test_refinement_on_other_frames(case, h, w, mask_on):
    """zb_refine on a non-square frame with the mask off, and on a frame whose masked pixels are
    more than a lane keeps for the NMI (5024 > 48 x 64), against the oracle from the same seeds
    (the truth turned by about 1 degree)."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine, sym
    from zebsdsolve_amd.geometry import Geometry
    print(f"refine frame {h}x{w} mask={mask_on}", flush=True)
    n = 6
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")  # the path alone, as the oracle
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.settings.apply(mask=mask_on)
    e.set_patterns(np.zeros((1, n, h, w), np.float32))
    _, (sheets_d, grad_d, host) = e.baked(0)
    side = e.phases[0].master.side
    R = op.bruker_R(70.0, 0.0)
    fr = op.frame(R, h, w, mask_on=mask_on)
    t = op.translation(R, cases.PC, float(h), float(w))
    rng = np.random.default_rng(3)
    truth = cases.random_quaternions(rng, n)
    pats = np.stack([op.project(host, side, fr, t, q) for q in truth]).astype(np.float32)
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    half = np.radians(1.0) / 2
    turn = np.concatenate([np.full((n, 1), np.cos(half)), axis * np.sin(half)], 1)
    seeds = np.stack([op.qmul(turn[i], truth[i]) for i in range(n)])
    lib = nat.load()
    raw = torch.from_numpy(pats).cuda()
    pn = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    st = engine.stream_ptr()
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, n, h, w, mask, pn.data_ptr(), st), "normalize")
    items = torch.from_numpy(np.concatenate([np.tile(t.astype(np.float32), (n, 1)), np.arange(n, dtype=np.float32)[:, None]], 1)).cuda()
    sd = torch.from_numpy(seeds.astype(np.float32)).cuda()
    opts = nat.RefineOpts(4, float(np.radians(2.0)), 1.0, 0.02, 1e4, 16, 9)
    rec = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    nat.check(lib.zb_refine(sheets_d.data_ptr(), e.grad_half(0).data_ptr(), side, e.frame(), pn.data_ptr(), items.data_ptr(), sd.data_ptr(), 4, n, opts, rec.data_ptr(), 0, st),
              "refine")
    r = rec.cpu().numpy()
    assert np.all(r[:, 7] == 9.0)
    pnh = pn.cpu().numpy()
    g = grad_d.cpu().numpy().reshape(-1, 4)
    rots = e.phases[0].master.rotations
    for b in range(n):
        q, ncc, _ = op.refine(g, side, fr, pnh[b].astype(np.float64), t, seeds[b], iters=4)
        # float32 sums over the frame against float64: the same iterate to 0.02 deg, NCC to 1e-3
        assert np.degrees(sym.disorientation(r[b, :4].astype(np.float64), q, rots)) < 0.02, b
        assert abs(r[b, 4] - ncc) < 1e-3, (b, r[b, 4], ncc)
        nm = op.nmi(host, side, fr, pnh[b].astype(np.float64), t, r[b, :4].astype(np.float64))
        assert abs(r[b, 5] - nm) < 2e-3, (b, r[b, 5], nm)


def test_batches_on_two_streams_match_one_stream(case):
    """Batches spread over two stream lanes (each with its own buffers) give the records one
    stream gives: every item's result depends only on its own pattern."""
    import torch
    e = case["e"]
    e.settings.apply(bandwidth=L, refine=True)
    raw = torch.from_numpy(case["pats"]).cuda()
    t = e.positions(np.arange(N))
    items = torch.from_numpy(np.concatenate([t, np.arange(N, dtype=np.float32)[:, None]], 1)).cuda()
    one, _ = e.run_batch(raw, items, L)
    torch.cuda.synchronize()
    lanes = e.lanes(2)
    half = N // 2
    it2 = items[half:].clone()
    it2[:, 3] -= half
    a, _ = e.run_batch(raw[:half], items[:half].contiguous(), L, lane=lanes[0])
    b, _ = e.run_batch(raw[half:].contiguous(), it2, L, lane=lanes[1])
    torch.cuda.synchronize()
    got = torch.cat([a, b], 1).cpu().numpy()
    ref = one.cpu().numpy()
    assert np.array_equal(got[..., :7], ref[..., :7])
