"""Bandwidth 96 (n = 256): the matrix-core correlation (zb_vcorr_*, its transform along a with
two lanes per column) and the VALU correlation (zb_so3_*) against the float64 oracle, and the
whole path against the truth.

Tolerances as at L = 32 and 64 (tests/test_gpu_pipeline.py), stated at each assertion. The
oracle's rotation tables at this size take ~30 s to build.
"""
import numpy as np
import pytest

from tests import cases
from tests.conftest import NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

L = 96
N = 6


def test_l96_correlation_against_the_oracle_and_recovery():
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import _native as nat, engine, sym
    from zebsdsolve_amd.engine import stream_ptr
    from zebsdsolve_amd.geometry import Geometry
    print("L=96 slab path", flush=True)
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")  # the path alone, as the oracle
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.settings.apply(bandwidth=L)
    e.set_patterns(np.zeros((1, N, 60, 60), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, truth = cases.simulate(host, side, N, seed=31)
    e.set_patterns(pats.reshape(1, N, 60, 60))
    fr, t = cases.oracle_frame(60, 60)
    assert e.use_mfma(L)  # nickel at n = 256: the matrix cores
    # the whole path against the truth (the reference's thresholds)
    rec = e.index_cells(np.arange(N), L)[0]
    rots = e.phases[0].master.rotations
    d = np.degrees(sym.disorientation(rec[:, :4].astype(np.float64), truth, rots))
    assert np.median(d) < 0.15 and d.max() < 1.0, d
    # the correlation's peaks for two of them against the oracle
    lib = nat.load()
    band = e.band(L)
    ph = e.phase_band(0, L)
    f, f2, fold, inv = op.harmonics(host, side, L)
    so = op.So3(L, fold)
    wa, wst = op.coefficients(None, fr, t, L, center=False, norm=False, window=True)
    rden_ref = so.rden(so.volume(f, wa), so.volume(f2, wa), wst[0])
    raw = torch.from_numpy(pats[:2]).cuda()
    pn = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, 2, 60, 60, mask, pn.data_ptr(), stream_ptr()), "normalize")
    items = torch.from_numpy(np.concatenate([e.positions(np.arange(2)), np.arange(2, dtype=np.float32)[:, None]], 1)).cuda()
    alm = torch.empty((2, band.host.pairs, 2), dtype=torch.float32, device="cuda")
    stats = torch.empty((2, 4), dtype=torch.float32, device="cuda")
    ws = torch.empty((lib.zb_sh_workspace(band.c, 2),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_sh_coefficients(band.c, e.frame(), pn.data_ptr(), items.data_ptr(), 2, 3, ws.data_ptr(), ws.numel(), alm.data_ptr(),
                                     stats.data_ptr(), stream_ptr()), "coefficients")
    peaks = torch.empty((2, 8), dtype=torch.float32, device="cuda")
    ws2 = torch.empty((lib.zb_so3_workspace(band.c, ph.c, 2),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_so3_correlate(band.c, ph.c, alm.data_ptr(), 2, e.rden(0, L).data_ptr(), 2, ws2.data_ptr(), ws2.numel(), peaks.data_ptr(),
                                   stream_ptr()), "so3")
    (vt, _), _ = ph.vtables()
    peaks2 = torch.empty((2, 8), dtype=torch.float32, device="cuda")
    ws3 = torch.empty((lib.zb_vcorr_workspace(vt, 2, 0),), dtype=torch.uint8, device="cuda")
    nat.check(lib.zb_vcorr_correlate(vt, band.tw.data_ptr(), alm.data_ptr(), 2, e.vrden(0, L).data_ptr(), 2, ws3.data_ptr(), ws3.numel(),
                                     peaks2.data_ptr(), stream_ptr()), "vcorr")
    pnh = pn.cpu().numpy()
    for pk in (peaks.cpu().numpy(), peaks2.cpu().numpy()):
        for b in range(2):
            ref_alm, _ = op.coefficients(pnh[b].astype(np.float64), fr, t, L)
            q, v0, _, _ = so.peak(so.volume(f, ref_alm), rden_ref)
            # the score to float32 accuracy, the orientation within 1/100 of a grid step (1.4 deg)
            assert abs(pk[b, 4] - v0) < 2e-4 * abs(v0), (b, pk[b, 4], v0)
            assert np.degrees(sym.disorientation(pk[b, :4].astype(np.float64), q, rots)) < 0.014, b


def test_gan_l96_matrix_cores_match_the_slab_path():
    """GaN at L = 96 (fold 2, every l: 96 degree slots, nc = 128): the matrix-core correlation
    with its 2.4 GB table and the denominators taken from the slab path in natural order,
    against the VALU slab path on the same coefficients (scores to float32 rounding of the two
    summation orders, the same cell and sub-cell peak), and the whole path against the truth."""
    import torch
    from tests.conftest import GAN
    from tests.test_gpu_vcorr import _coefficients
    from zebsdsolve_amd import _native as nat, engine, sym
    from zebsdsolve_amd.engine import stream_ptr
    from zebsdsolve_amd.geometry import Geometry
    print("GaN L=96 matrix cores", flush=True)
    n = 8
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(GAN)
    e.settings.apply(bandwidth=L)
    e.set_patterns(np.zeros((1, n, 60, 60), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, truth = cases.simulate(host, side, n, seed=17)
    e.set_patterns(pats.reshape(1, n, 60, 60))
    assert e.use_mfma(L)
    ph = e.phase_band(0, L)
    assert ph.big_tables() and ph.host.nc == 128
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
    rots = e.phases[0].master.rotations
    # the scores to float32 rounding of the two summation orders, the same peak
    assert np.all(np.abs(a[:, 4] - b[:, 4]) <= 2e-5 * np.abs(b[:, 4])), np.abs(a[:, 4] - b[:, 4]).max()
    d = np.degrees(sym.disorientation(a[:, :4].astype(np.float64), b[:, :4].astype(np.float64), rots))
    assert d.max() < 0.01, d
    # the whole path against the truth (the reference's thresholds)
    rec = e.index_cells(np.arange(n), L)[0]
    dt = np.degrees(sym.disorientation(rec[:, :4].astype(np.float64), truth, rots))
    assert np.median(dt) < 0.15 and dt.max() < 1.0, dt
