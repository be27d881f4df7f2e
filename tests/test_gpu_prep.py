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
