"""The renderer (zb_project, detector.wgsl:254-263; src/engine.zig:2753-2769) against the oracle's
project at 60 x 60 and 120 x 120, and the matrix-core correlation of GaN at L = 96 on the
orientations whose peak lies at b in [160, 180] degrees, the rows of its 2.4 GB table past
2 GiB, against the VALU slab path."""
import numpy as np
import pytest

from tests import cases
from tests.conftest import GAN, NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


@pytest.mark.parametrize("size", [60, 120])
def test_project_matches_the_oracle(size):
    """Each pixel the bilinear master at q . normalize(ray) in float32, against float64: within
    1e-4 of the master's unit spread (the Lambert map's arithmetic and float32 rays), on the
    baked nickel master at 8 random orientations, the rays from per-pattern positions."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine
    from zebsdsolve_amd.engine import stream_ptr
    from zebsdsolve_amd.geometry import Geometry
    print(f"project {size}x{size}", flush=True)
    n = 8
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.set_patterns(np.zeros((1, n, size, size), np.float32))
    _, (sheets, _, host) = e.baked(0)
    side = e.phases[0].master.side
    rng = np.random.default_rng(size)
    q = cases.random_quaternions(rng, n)
    # positions jittered per pattern, as the bench's PC jitter
    pcs = np.array(cases.PC)[None, :] + rng.uniform(-0.02, 0.02, size=(n, 3))
    R = op.bruker_R(70.0, 0.0)
    ts = np.stack([op.translation(R, pc, float(size), float(size)) for pc in pcs])
    poses = np.concatenate([ts, np.zeros((n, 1)), q], 1).astype(np.float32)
    out = torch.empty((n, size, size), dtype=torch.float32, device="cuda")
    lib = nat.load()
    pd = torch.from_numpy(poses).cuda()
    nat.check(lib.zb_project(sheets.data_ptr(), side, e.frame(), pd.data_ptr(), n, out.data_ptr(), stream_ptr()), "project")
    got = out.cpu().numpy()
    fr = op.frame(R, size, size)
    for b in range(n):
        want = op.project(host, side, fr, poses[b, :3].astype(np.float64), poses[b, 4:].astype(np.float64))
        err = np.abs(got[b] - want).max()
        assert err < 1e-4, (b, err)
    assert np.isfinite(got).all() and got.std() > 0.1


def test_gan_l96_peaks_at_b_past_160_degrees_read_the_whole_table():
    """Orientations with b in [162, 179.5] degrees put the GaN correlation's peak in the table's
    rows j >= 115 of 129 (the 2.4 GB table's last 0.3 GB, past 2 GiB): the matrix-core scores and
    cells equal the slab path's, and the whole path recovers the truth."""
    import torch
    from oracle import pipeline as op
    from tests.test_gpu_vcorr import _coefficients
    from zebsdsolve_amd import _native as nat, engine, sym
    from zebsdsolve_amd.engine import stream_ptr
    from zebsdsolve_amd.geometry import Geometry
    print("GaN L=96 b > 160 deg", flush=True)
    L, n = 96, 8
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(GAN)
    e.settings.apply(bandwidth=L)
    e.set_patterns(np.zeros((1, n, 60, 60), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    rng = np.random.default_rng(41)
    A, C = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    B = np.radians(rng.uniform(162.0, 179.5, n))
    # the peak's orientation is conj(Rz(a) Ry(b) Rz(c)); rotations about the master's c axis keep b
    truth = np.stack([op.qconj(op.zyz(a, b, c)) for a, b, c in zip(A, B, C)])
    fr, t = cases.oracle_frame(60, 60)
    pats = np.stack([op.project(host, side, fr, t, qq) for qq in truth]).astype(np.float32)
    e.set_patterns(pats.reshape(1, n, 60, 60))
    ph = e.phase_band(0, L)
    assert e.use_mfma(L) and ph.big_tables()
    alm, _ = _coefficients(e, L, pats)
    lib = nat.load()
    band = e.band(L)
    vt, _ = ph.vtable_of(0)
    p1 = torch.empty((n, 8), device="cuda")
    ws = torch.empty((lib.zb_vcorr_workspace(vt, n, 0),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_vcorr_correlate(vt, band.tw.data_ptr(), alm.data_ptr(), n, e.vrden(0, L).data_ptr(), n, ws.data_ptr(), ws.numel(),
                                     p1.data_ptr(), stream_ptr()), "vcorr")
    p2 = torch.empty((n, 8), device="cuda")
    ws2 = torch.empty((lib.zb_so3_workspace(band.c, ph.c, n),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_so3_correlate(band.c, ph.c, alm.data_ptr(), n, e.rden(0, L).data_ptr(), n, ws2.data_ptr(), ws2.numel(), p2.data_ptr(),
                                   stream_ptr()), "so3")
    a, b = p1.cpu().numpy(), p2.cpu().numpy()
    # the peak cells are in the far rows: b (slot 6, radians) past 160 degrees
    assert (np.degrees(b[:, 6]) > 160).all(), np.degrees(b[:, 6])
    assert np.all(np.abs(a[:, 4] - b[:, 4]) <= 2e-5 * np.abs(b[:, 4])), np.abs(a[:, 4] - b[:, 4]).max()
    rots = e.phases[0].master.rotations
    d = np.degrees(sym.disorientation(a[:, :4].astype(np.float64), b[:, :4].astype(np.float64), rots))
    assert d.max() < 0.01, d
    rec = e.index_cells(np.arange(n), L)[0]
    dt = np.degrees(sym.disorientation(rec[:, :4].astype(np.float64), truth, rots))
    assert np.median(dt) < 0.15 and dt.max() < 1.0, dt
