"""zb_prep, the pattern filters fused with the normalization (static background subtracted, the
dynamic background subtracted or divided by, zero mean and unit norm under the mask), against
the float64 oracle (oracle/pipeline.py prep: prep.wgsl's direct and box-sum blurs) on the same
patterns, for both blur paths (sigma below and above 12 pixels), non-square frames, u16 input."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _patterns(n, h, w, seed):
    """EBSD-like: a broad background hump, bands, Poisson-like noise, around 400 counts."""
    rng = np.random.default_rng(seed)
    r, c = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    hump = 300 * np.exp(-((r - 0.4 * h) ** 2 + (c - 0.6 * w) ** 2) / (2 * (0.5 * w) ** 2))
    out = []
    for _ in range(n):
        a = rng.uniform(0, np.pi, 3)
        bands = sum(60 * np.cos(0.3 * (np.cos(t) * r + np.sin(t) * c) + rng.uniform(0, 6)) for t in a)
        out.append(np.maximum(hump + bands + 200 + rng.normal(0, 8, (h, w)), 0))
    return np.stack(out)


CASES = [  # h, w, static, dynamic, dtype
    (60, 60, True, "none", np.float32),
    (60, 60, False, "subtract", np.float32),
    (60, 60, True, "divide", np.float32),
    (45, 70, True, "subtract", np.uint16),
    (80, 80, True, "subtract", np.float32),    # sigma 12: boxes of 2
    (120, 120, False, "subtract", np.float32),  # sigma 18: boxes of 3
]


@pytest.mark.parametrize("h,w,static,dynamic,dtype", CASES)
def test_prep_matches_the_oracle(h, w, static, dynamic, dtype):
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    print(f"prep {h}x{w} static={static} dynamic={dynamic} {np.dtype(dtype).name}", flush=True)
    n = 5
    raw = _patterns(n, h, w, seed=h + w).astype(dtype)
    bg = raw.astype(np.float64).mean(0).astype(np.float32) if static else None
    fr = op.frame(np.eye(3), h, w)
    lib = nat.load()
    dev = torch.device("cuda")
    t_raw = torch.from_numpy(raw).to(dev)
    out = torch.empty((n, h, w), dtype=torch.float32, device=dev)
    t_bg = torch.from_numpy(bg).to(dev) if static else None
    o = nat.PrepOpts()
    o.background = t_bg.data_ptr() if static else None
    o.dynamic = {"none": 0, "subtract": 1, "divide": 2}[dynamic]
    o.dynamic_sigma = 0.15 * w
    mask = (nat.ctypes.c_float * 3)(*[float(x) for x in fr["mask"]])
    kind = {np.float32: 2, np.uint16: 1, np.uint8: 0}[dtype]
    nat.check(lib.zb_prep(t_raw.data_ptr(), kind, n, h, w, mask, nat.ctypes.byref(o), out.data_ptr(), engine.stream_ptr()), "prep")
    got = out.cpu().numpy()
    for b in range(n):
        ref = op.prep(raw[b], fr["mask"], background=bg.astype(np.float64) if static else None, dynamic=dynamic, dynamic_sigma=0.15)
        # float32 sums of up to ~80 taps and a difference of values ~400 against float64: 2e-6
        # of the unit-norm pattern's largest value
        err = np.abs(got[b] - ref).max() / np.abs(ref).max()
        assert err < 2e-5, (b, err)
        assert abs(np.sum(got[b] ** 2) - 1) < 1e-4


def test_prep_without_filters_is_normalize():
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    n, h, w = 4, 60, 60
    raw = _patterns(n, h, w, seed=1).astype(np.float32)
    fr = op.frame(np.eye(3), h, w)
    lib = nat.load()
    t_raw = torch.from_numpy(raw).cuda()
    a = torch.empty((n, h, w), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    o = nat.PrepOpts()
    mask = (nat.ctypes.c_float * 3)(*[float(x) for x in fr["mask"]])
    st = engine.stream_ptr()
    nat.check(lib.zb_prep(t_raw.data_ptr(), 2, n, h, w, mask, nat.ctypes.byref(o), a.data_ptr(), st), "prep")
    nat.check(lib.zb_normalize(t_raw.data_ptr(), 2, n, h, w, mask, b.data_ptr(), st), "normalize")
    assert torch.allclose(a, b, atol=1e-7, rtol=0)


def test_toy_scan_through_the_default_filters():
    """The reference's defaults (static background: the map's average; dynamic background:
    subtracted) on a toy scan: the orientations are recovered (tests/test_api.py:236-240 checks
    the average alone)."""
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.set_geometry(pc=(0.44, 0.24, 0.53), sample_tilt_deg=70, detector_tilt_deg=0, px_um=70 * 8, step_x_um=20, step_y_um=20)
    s.add_phase(__import__("tests.conftest", fromlist=["NI"]).NI)
    truth = s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=3)
    tq = truth["quaternions"].reshape(-1, 4)
    assert s.settings["static_background"] == "average" and s.settings["dynamic_background"] == "subtract"
    for kw in ({"dynamic_background": "none"}, {"dynamic_background": "subtract"}):
        s.configure(**kw)
        r = s.index()
        d = np.degrees(s.disorientation(r["quaternions"].reshape(-1, 4), tq))
        assert np.median(d) < 0.15 and d.max() < 1.0, (kw, np.median(d), d.max())


def test_global_clahe_matches_the_oracle():
    """zb_clahe (the master's CLAHE, in global memory) on two images of odd size against the
    oracle's restatement of prep.wgsl: the stretch and the levels in float32 as there, so only
    the blends' last bits differ."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    lib = nat.load()
    rng = np.random.default_rng(9)
    h, w = 101, 77
    imgs = (_patterns(2, h, w, seed=4) + rng.normal(0, 20, (2, h, w))).astype(np.float32)
    for grid, bins, clip in ((8, 64, 4.0), (3, 256, 40.0), (16, 16, 1.5)):
        x = torch.from_numpy(imgs).cuda()
        ws = torch.empty((lib.zb_clahe_workspace(2, grid, bins),), dtype=torch.uint8, device="cuda")
        nat.check(lib.zb_clahe(x.data_ptr(), 2, h, w, grid, bins, clip, ws.data_ptr(), ws.numel(), engine.stream_ptr()), "clahe")
        got = x.cpu().numpy()
        for b in range(2):
            ref = op.clahe(op.stretch(imgs[b]), grid, bins, clip)
            assert np.abs(got[b] - ref).max() < 1e-5, (grid, bins, clip, b, np.abs(got[b] - ref).max())
    x = torch.from_numpy(imgs).cuda()
    assert lib.zb_clahe(x.data_ptr(), 2, h, w, 4, 64, 40.0, x.data_ptr(), 10, engine.stream_ptr()) < 0
    assert "workspace" in nat.last_error()


def test_master_symmetrize_is_the_point_group_mean():
    """zb_master_symmetrize against the oracle's bilinear lookup averaged over the master's
    operators at 3000 texels; and the nickel master, which obeys its point group, moves only by
    interpolation (0.069 of its unit deviation, rms, unsmoothed)."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine, master as mm
    from tests.conftest import NI
    lib = nat.load()
    m = mm.load(NI)
    side = m.side
    rng = np.random.default_rng(2)
    smooth = np.asarray(m.raw, np.float32)
    src = torch.from_numpy(smooth.reshape(-1)).cuda()
    dst = torch.empty_like(src)
    ops = torch.from_numpy(np.asarray(m.ops, np.float32).reshape(-1)).cuda()
    nat.check(lib.zb_master_symmetrize(src.data_ptr(), dst.data_ptr(), side, ops.data_ptr(), len(m.ops), engine.stream_ptr()), "symmetrize")
    got = dst.cpu().numpy()
    tid = rng.integers(0, 2 * side * side, 3000)
    pix = tid % (side * side)
    x = 2.0 * (pix % side) / (side - 1) - 1.0
    y = 2.0 * (pix // side) / (side - 1) - 1.0
    d = op.square_to_sphere(x, y, tid >= side * side)
    want = np.mean([op.sheet_at(smooth.reshape(2, side, side).astype(np.float64), side, d @ np.asarray(g).T) for g in m.ops], axis=0)
    assert np.abs(got[tid] - want).max() < 1e-4, np.abs(got[tid] - want).max()
    assert np.sqrt(np.mean((got - smooth.reshape(-1)) ** 2)) < 0.1
    assert lib.zb_master_symmetrize(src.data_ptr(), src.data_ptr(), side, ops.data_ptr(), len(m.ops), None) < 0


def _prep_call(lib, nat, t_raw, kind, n, h, w, mask, o, out, work=True):
    from zebsdsolve_amd import engine
    import torch
    need = int(lib.zb_prep_workspace(n, h, w, nat.ctypes.byref(o)))
    assert need >= 0, nat.last_error()
    ws = torch.empty((max(need, 1),), dtype=torch.uint8, device="cuda") if work else None
    nat.check(lib.zb_prep_work(t_raw.data_ptr(), kind, n, h, w, mask, nat.ctypes.byref(o), ws.data_ptr() if work else None,
                               need if work else 0, out.data_ptr(), engine.stream_ptr()), "prep work")
    return need


@pytest.mark.parametrize("h,w", [(60, 60), (120, 120), (45, 70)])
def test_prep_through_global_memory_equals_the_lds_form(h, w):
    """zb_prep_work with the filters forced through global memory (knob prep_global) gives every
    pixel of zb_prep's LDS form, bit for bit, for every filter: static and dynamic backgrounds
    (both blur paths), CLAHE, the band pass with and without its fine Gaussian."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    print(f"prep global {h}x{w}", flush=True)
    n = 6
    raw = _patterns(n, h, w, seed=7 + h).astype(np.float32)
    bg = torch.from_numpy(raw.astype(np.float64).mean(0).astype(np.float32)).cuda()
    fr = op.frame(np.eye(3), h, w)
    mask = (nat.ctypes.c_float * 3)(*[float(x) for x in fr["mask"]])
    lib = nat.load()
    t_raw = torch.from_numpy(raw).cuda()
    combos = [dict(dynamic=1, dynamic_sigma=0.15 * w), dict(dynamic=2, dynamic_sigma=0.15 * w, clahe=1, clahe_grid=4, clahe_bins=64, clahe_clip=2.0),
              dict(dynamic=1, dynamic_sigma=4.0, bandpass=1, bp_fine=1.5, bp_coarse=0.2 * w), dict(bandpass=1, bp_fine=0.0, bp_coarse=13.0)]
    prev = lib.zb_knob_set(b"prep_global", 0)
    tested = 0
    try:
        for kw in combos:
            o = nat.PrepOpts()
            o.background = bg.data_ptr()
            for k, v in kw.items():
                setattr(o, k, v)
            a = torch.empty((n, h, w), dtype=torch.float32, device="cuda")
            b = torch.full((n, h, w), np.nan, dtype=torch.float32, device="cuda")
            if int(lib.zb_prep_workspace(n, h, w, nat.ctypes.byref(o))) > 0:
                continue  # past one block's LDS with these filters: no LDS form to hold it to
            tested += 1
            nat.check(lib.zb_prep(t_raw.data_ptr(), 2, n, h, w, mask, nat.ctypes.byref(o), a.data_ptr(), engine.stream_ptr()), "prep")
            lib.zb_knob_set(b"prep_global", 1)
            need = _prep_call(lib, nat, t_raw, 2, n, h, w, mask, o, b)
            lib.zb_knob_set(b"prep_global", 0)
            assert need > 0
            assert torch.equal(a, b), (kw, (a - b).abs().max().item())
    finally:
        lib.zb_knob_set(b"prep_global", prev)
    assert tested >= 3


def test_prep_of_240x320_patterns_matches_the_oracle():
    """240 x 320 patterns (the reference's large frames; 300 KB an image, past one block's LDS)
    through the default filters (the scan's average subtracted, the dynamic background of
    sigma 48 px subtracted by the box path) and the band pass: zb_prep refuses them, zb_prep_work
    takes them through global memory and matches the float64 oracle; a workspace of one slot
    (one block) gives the same pixels, and less than a slot is refused."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    print("prep 240x320", flush=True)
    n, h, w = 3, 240, 320
    raw = _patterns(n, h, w, seed=5).astype(np.uint16)
    bg64 = raw.astype(np.float64).mean(0)
    bg = torch.from_numpy(bg64.astype(np.float32)).cuda()
    fr = op.frame(np.eye(3), h, w)
    mask = (nat.ctypes.c_float * 3)(*[float(x) for x in fr["mask"]])
    lib = nat.load()
    t_raw = torch.from_numpy(raw.astype(np.int32)).cuda().to(torch.uint16)
    o = nat.PrepOpts()
    o.background = bg.data_ptr()
    o.dynamic, o.dynamic_sigma = 1, 0.15 * w
    out = torch.empty((n, h, w), dtype=torch.float32, device="cuda")
    assert lib.zb_prep(t_raw.data_ptr(), 1, n, h, w, mask, nat.ctypes.byref(o), out.data_ptr(), engine.stream_ptr()) != 0
    assert "zb_prep_work" in nat.last_error()
    _prep_call(lib, nat, t_raw, 1, n, h, w, mask, o, out)
    got = out.cpu().numpy()
    for b in range(n):
        ref = op.prep(raw[b], fr["mask"], background=bg64.astype(np.float32).astype(np.float64), dynamic="subtract", dynamic_sigma=0.15)
        err = np.abs(got[b] - ref).max() / np.abs(ref).max()
        assert err < 2e-5, (b, err)
    # one slot: one block takes the patterns in turn
    one = torch.full_like(out, np.nan)
    need = int(lib.zb_prep_workspace(n, h, w, nat.ctypes.byref(o)))
    slot = need // n
    ws = torch.empty((slot,), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_prep_work(t_raw.data_ptr(), 1, n, h, w, mask, nat.ctypes.byref(o), ws.data_ptr(), slot, one.data_ptr(),
                               engine.stream_ptr()), "prep one slot")
    assert torch.equal(one, out)
    assert lib.zb_prep_work(t_raw.data_ptr(), 1, n, h, w, mask, nat.ctypes.byref(o), ws.data_ptr(), slot - 256, one.data_ptr(),
                            engine.stream_ptr()) == -2
    # the band pass at this size (its coarse Gaussian by boxes of 6 pixels); CLAHE's global form
    # is held to the LDS form's pixels above
    o.bandpass, o.bp_fine, o.bp_coarse = 1, 2.0, 40.0
    _prep_call(lib, nat, t_raw, 1, n, h, w, mask, o, out)
    got = out.cpu().numpy()
    for b in range(n):
        ref = op.prep(raw[b], fr["mask"], background=bg64.astype(np.float32).astype(np.float64), dynamic="subtract", dynamic_sigma=0.15,
                      bandpass=(2.0, 40.0))
        err = np.abs(got[b] - ref).max() / np.abs(ref).max()
        assert err < 2e-5, (b, err)


def test_240x320_toy_scan_through_the_default_filters():
    """A toy scan of 240 x 320 patterns with the reference's default filters (the map's average
    and the dynamic background subtracted, through global memory) is recovered to the
    reference's thresholds (median < 0.15 deg, max < 1 deg)."""
    import zebsdsolve_amd as ze
    print("240 x 320 toy scan, default filters", flush=True)
    s = ze.Solver()
    s.set_geometry(pc=(0.44, 0.24, 0.53), sample_tilt_deg=70, detector_tilt_deg=0, px_um=70 * 8, step_x_um=20, step_y_um=20)
    s.add_phase(__import__("tests.conftest", fromlist=["NI"]).NI)
    truth = s.synthesize(toy_rows=4, toy_cols=6, toy_grains=24, toy_counts=400, toy_seed=17, toy_height=240, toy_width=320)
    assert s.settings["static_background"] == "average" and s.settings["dynamic_background"] == "subtract"
    r = s.index()
    d = np.degrees(s.disorientation(r["quaternions"].reshape(-1, 4), truth["quaternions"].reshape(-1, 4)))
    assert np.median(d) < 0.15 and d.max() < 1.0, (np.median(d), d.max())


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_normalize_of_every_pixel_type_and_alignment_matches_the_oracle(dtype):
    """zb_normalize for u8, u16 and f32 pixels: four pixels a load where the frame's pixel count
    is a multiple of 4 and the buffers are aligned (60 x 60, 120 x 120), one pixel a load
    otherwise (45 x 70, 121 x 130, and a raw buffer one pixel off its alignment), all against the
    float64 oracle."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    lib = nat.load()
    kind = {np.float32: 2, np.uint16: 1, np.uint8: 0}[dtype]
    for h, w, skew in ((60, 60, 0), (120, 120, 0), (45, 70, 0), (121, 130, 0), (60, 60, 1)):
        print(f"normalize {np.dtype(dtype).name} {h}x{w} skew {skew}", flush=True)
        n = 3
        raw = _patterns(n, h, w, seed=h * w + kind).astype(dtype)
        fr = op.frame(np.eye(3), h, w)
        mask = (nat.ctypes.c_float * 3)(*[float(x) for x in fr["mask"]])
        flat = np.concatenate([np.zeros(skew, dtype), raw.reshape(-1)])
        t_raw = torch.from_numpy(flat).cuda()
        out = torch.empty((n, h, w), dtype=torch.float32, device="cuda")
        ptr = t_raw.data_ptr() + skew * np.dtype(dtype).itemsize
        nat.check(lib.zb_normalize(ptr, kind, n, h, w, mask, out.data_ptr(), engine.stream_ptr()), "normalize")
        got = out.cpu().numpy()
        for b in range(n):
            ref = op.normalize(raw[b].astype(np.float64), fr["mask"])
            err = np.abs(got[b] - ref).max() / np.abs(ref).max()
            assert err < 2e-5, (h, w, skew, b, err)
