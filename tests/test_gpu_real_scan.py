"""The real nickel scan (kikuchipy's nickel_ebsd_large, the reference's testdata/nickel) indexed
with the reference's default settings (static background: the map's average; dynamic
background: subtracted; L = 64; 8 Gauss-Newton iterations) against the orientations kikuchipy
refined for it (reference_quaternions.f32), with the reference's own thresholds
(tests/test_api.py:717-731): median < 0.6 deg, > 97 % within 3 deg, mean NCC > 0.45, and nearly
every match accepted (> 98 % at z >= 15, median z > 30)."""
import os

import numpy as np
import pytest

from tests.conftest import NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

DIR = os.path.join(os.path.dirname(__file__), "golden", "nickel_scan")


def test_nickel_scan_indexes_as_kikuchipy_refined_it():
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.open(os.path.join(DIR, "patterns_v2.h5"))
    s.add_phase(NI)
    q_ref = np.fromfile(os.path.join(DIR, "reference_quaternions.f32"), "<f4").reshape(-1, 4)
    seen = []
    r = s.index(lambda st: seen.append((st["done"], st["total"])))
    dis = np.degrees(s.disorientation(r["quaternions"].reshape(-1, 4), q_ref))
    print(f"median {np.median(dis):.3f} deg, within 3 deg {100 * (dis < 3).mean():.1f} %, mean NCC {r['ncc'].mean():.3f}, "
          f"accepted {100 * (r['sweep'] >= 0).mean():.1f} %, median z {np.median(r['z']):.1f}", flush=True)
    assert np.median(dis) < 0.6 and (dis < 3).mean() > 0.97 and r["ncc"].mean() > 0.45
    assert seen[-1] == (4125, 4125)
    assert (r["sweep"] >= 0).mean() > 0.98 and np.median(r["z"]) > 30
    # the stored background, subtracted instead of the map's average, indexes as well
    s.configure(static_background="file")
    r2 = s.index()
    d2 = np.degrees(s.disorientation(r2["quaternions"].reshape(-1, 4), q_ref))
    assert np.median(d2) < 0.6 and (d2 < 3).mean() > 0.97


def _unit(a, mask=None):
    m = np.ones(a.shape, bool) if mask is None else mask
    out = np.where(m, a - a[m].mean(), 0).astype(np.float64)
    return out / np.sqrt((out ** 2).sum())


def test_clahe_on_the_scan_is_the_references():
    """The reference's CLAHE check on the scan (tests/test_api.py:698-706): pattern 7 after the
    default backgrounds, then CLAHE at three grid / level / clip settings, against the oracle's
    restatement of prep.wgsl applied to the pattern without it (a stretch first, so the
    normalization before does not matter), within 2e-4 (pixels are ~1/60); also at 16 x 16 tiles
    of 256 levels, whose histograms (256 KB) pass one block's LDS, so that the filters run through
    global memory (zb_prep_work); then with the mask, which comes last."""
    from oracle import pipeline as op
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.open(os.path.join(DIR, "patterns_v2.h5"))
    s.configure(mask=False)
    before = s.pattern(7)
    worst = 0.0
    for grid, bins, clip in ((4, 64, 40.0), (7, 32, 2.0), (8, 128, 40.0), (16, 256, 40.0)):
        s.configure(clahe=True, clahe_grid=grid, clahe_bins=bins, clahe_clip=clip)
        worst = max(worst, float(np.abs(s.pattern(7) - _unit(op.clahe(op.stretch(before), grid, bins, clip))).max()))
    print(f"CLAHE: largest difference from the oracle {worst:.1e}", flush=True)
    assert worst < 2e-4
    s.configure(mask=True, clahe_grid=4, clahe_bins=64, clahe_clip=40)
    inside = (np.arange(60)[:, None] - 29.5) ** 2 + (np.arange(60)[None, :] - 29.5) ** 2 <= 30 ** 2
    got = s.pattern(7)
    assert np.abs(got - _unit(op.clahe(op.stretch(before), 4, 64, 40.0), inside)).max() < 2e-4 and (got[~inside] == 0).all()


def test_bandpass_on_the_scan_is_the_oracles():
    """The band pass after the default backgrounds: the fine Gaussian (1 pixel) minus the coarse
    one (8; and 14, the box path) of the pattern without it, as the oracle computes them."""
    from oracle import pipeline as op
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.open(os.path.join(DIR, "patterns_v2.h5"))
    s.configure(mask=False)
    # the backgrounds' output, unnormalized, from the oracle on the stored pattern
    raw = s.pattern(7, processed=False).astype(np.float64)
    bg = s.pattern(0, "background").astype(np.float64)
    base = op.prep(raw, None, background=bg, dynamic="subtract", dynamic_sigma=0.15, norm=False)
    for fine, coarse in ((1.0, 8.0), (0.0, 14.0)):
        s.configure(bandpass=True, bp_fine=fine, bp_coarse=coarse)
        want = _unit(op.prep(base, None, bandpass=(fine, coarse), norm=False))
        err = float(np.abs(s.pattern(7) - want).max())
        print(f"band pass {fine}/{coarse}: largest difference from the oracle {err:.1e}", flush=True)
        assert err < 2e-4


def test_clahe_and_bandpass_on_the_master_keep_the_scan_indexable():
    """The reference's check (tests/test_api.py:870-874): with CLAHE and the band pass on, the
    master filtered to match, the scan still indexes (> 97 % within 3 deg of kikuchipy's) and the
    mean NCC rises above the default filters'."""
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.open(os.path.join(DIR, "patterns_v2.h5"))
    s.add_phase(NI)
    q_ref = np.fromfile(os.path.join(DIR, "reference_quaternions.f32"), "<f4").reshape(-1, 4)
    r = s.index()
    s.configure(clahe=True, bandpass=True)
    rf = s.index()
    disf = np.degrees(s.disorientation(rf["quaternions"].reshape(-1, 4), q_ref))
    print(f"CLAHE + band pass: within 3 deg {100 * (disf < 3).mean():.1f} %, mean NCC {rf['ncc'].mean():.3f} "
          f"(default filters {r['ncc'].mean():.3f})", flush=True)
    assert (disf < 3).mean() > 0.97 and rf["ncc"].mean() > r["ncc"].mean()
