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
