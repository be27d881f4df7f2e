"""The HDF5 reader and Solver.open on the reference's real nickel scan (kikuchipy's
nickel_ebsd_large, testdata/nickel/patterns_v2.h5, kept under tests/golden/nickel_scan with its
licence), against the numbers the reference's own reader test holds (src/io/hdf5.zig:1236-1288:
sums of stored patterns checked there against h5py) and the header values
(tests/test_api.py:666-670)."""
import os

import numpy as np
import pytest

SCAN = os.path.join(os.path.dirname(__file__), "golden", "nickel_scan", "patterns_v2.h5")


def test_reader_matches_the_reference_reader_test():
    from zebsdsolve_amd import h5
    with h5.File(SCAN) as f:
        root = f["/"]
        assert len(root.keys()) == 3
        assert isinstance(f["Scan 1/EBSD"], h5.Group)
        hdr = f["Scan 1/EBSD/Header"]
        assert float(hdr["n_rows"].read()[0]) == 55 and float(hdr["step_x"].read()[0]) == 1.5
        with pytest.raises(KeyError):
            f["Scan 1/EBSD/Nope"]
        ds = f["Scan 1/EBSD/Data/patterns"]
        assert ds.shape == (4125, 60, 60) and ds._chunk[0] == 277 and ds.dtype == np.uint8
        assert int(ds.read_rows(0, 1).astype(np.int64).sum()) == 531330
        three = ds.read_rows(276, 279)  # across a chunk boundary
        assert int(three.astype(np.int64).sum()) == 1571367
        assert int(ds.read_rows(4124, 4125).astype(np.int64).sum()) == 523040
        allp = ds.read()
        assert np.array_equal(allp[276:279], three)
        pcx = hdr["pcx"].read().reshape(-1)
        assert pcx.size == 4125 and abs(pcx[0] - 0.42771687591335866) < 1e-15


def test_open_reads_the_scan_and_its_geometry():
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.open(SCAN)
    info = s.scan
    assert info["rows"] == 55 and info["cols"] == 75 and info["pattern_shape"] == (60, 60)
    assert info["format"] == "kikuchipy h5ebsd" and info["has_background"] and info["has_pc"]
    g = s.geometry
    assert np.allclose(g["pc"], (0.42326, 0.21363, 0.50207), atol=1e-4) and g["sample_tilt_deg"] == 70
    assert g["px_um"] == 560 and g["step_x_um"] == 1.5 and g["step_y_um"] == 1.5
    raw = s.pattern(0, processed=False)
    assert raw.shape == (60, 60) and 0 <= raw.min() and raw.max() <= 255 and int(raw.sum()) == 531330
