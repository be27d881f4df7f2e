"""The CPU restatement (oracle/cpu/zbcpu.cpp, float32 C++ / OpenMP, the bench's CPU baseline)
against the float64 oracle on the same patterns: the window's denominators, the correlation
peak's score and cell, the refined NCC and orientation, the NMI. No GPU."""
import numpy as np
import pytest

from oracle import pipeline as op
from tests import cases
from tests.conftest import NI


@pytest.fixture(scope="module")
def port_case():
    from oracle import cpu_port
    from zebsdsolve_amd import master, tables
    cpu_port.build()
    m = master.load(NI)
    L = 32
    band = tables.BandHost(L)
    ph = tables.PhaseHost(band, m.side, m.raw)
    grad = op.gradient(m.raw, m.side)
    fr, t = cases.oracle_frame(60, 60)
    port = cpu_port.Port(band, ph, fr, np.asarray(m.raw, np.float32), grad.astype(np.float32), m.side, iters=5)
    return m, L, grad, fr, t, port


def test_cpu_restatement_matches_the_oracle(port_case):
    from zebsdsolve_amd import sym
    m, L, grad, fr, t, port = port_case
    n = 4
    pats, truth = cases.simulate(m.raw, m.side, n, seed=3, counts=400)
    rden = port.rden(t)
    f, f2, fold, _ = op.harmonics(m.raw, m.side, L)
    so = op.So3(L, fold)
    wa, wst = op.coefficients(None, fr, t, L, center=False, norm=False, window=True)
    rden_o = so.rden(so.volume(f, wa), so.volume(f2, wa), wst[0])
    # float32 arithmetic of the same volumes
    assert np.abs(rden - rden_o).max() < 2e-6 * np.abs(rden_o).max()
    items = np.concatenate([np.repeat(np.asarray(t, np.float32)[None], n, 0), np.arange(n, dtype=np.float32)[:, None]], 1)
    rec, pk = port.index(pats, items, rden, threads=2)
    for b in range(n):
        pn = op.normalize(pats[b].astype(np.float64), fr["mask"])
        alm, _ = op.coefficients(pn, fr, t, L)
        q, v0, _, _ = so.peak(so.volume(f, alm), rden_o)
        assert abs(pk[b, 4] - v0) < 2e-5 * abs(v0), (b, pk[b, 4], v0)
        assert np.degrees(sym.disorientation(pk[b:b + 1, :4].astype(np.float64), q[None], m.rotations))[0] < 1e-3
        qr, ncc, _ = op.refine(grad, m.side, fr, pn, t, q, iters=5)
        assert abs(rec[b, 4] - ncc) < 1e-5, (b, rec[b, 4], ncc)
        assert np.degrees(sym.disorientation(rec[b:b + 1, :4].astype(np.float64), qr[None], m.rotations))[0] < 2e-3
        # NMI bins at each side's own extremes: float32 rounding moves a few pixels across a bin edge
        assert abs(rec[b, 5] - op.nmi(m.raw, m.side, fr, pn, t, qr)) < 5e-3
    d = np.degrees(sym.disorientation(rec[:, :4].astype(np.float64), truth, m.rotations))
    assert np.median(d) < 0.1
    # one thread and many give the same records
    rec1, _ = port.index(pats, items, rden, threads=1)
    assert np.array_equal(rec, rec1)
