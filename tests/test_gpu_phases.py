"""Several phases' refinements in one launch (zb_refine_phases).
- Through the C ABI: for two and three phases (different masters and sides, different seeds),
  with and without each phase's item order, at a record offset, and in each launch form (a block
  per work item; the persistent grid on device-side queues, keeping or looking up again the NMI's
  master values), every record is bit-identical to zb_refine's for that phase alone.
- Through the engine: a two-phase scan's records (one refinement launch for both phases) equal
  those of one-phase solvers on the same patterns, phase by phase.
The reference refines phase after phase, one dispatch each (src/engine.zig:2085-2119); the
records do not depend on how the work is launched."""
import ctypes

import numpy as np
import pytest

from tests.conftest import GAN, NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _solver(*phases, patterns=None):
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.set_geometry(pc=(0.44, 0.24, 0.53), sample_tilt_deg=70, detector_tilt_deg=0, px_um=70 * 8, step_x_um=20, step_y_um=20)
    s.configure(static_background="none", dynamic_background="none")
    for p in phases:
        s.add_phase(p)
    if patterns is not None:
        s.set_patterns(patterns)
    return s


def _seeds(truth, n, sigma, seed):
    rng = np.random.default_rng(seed)
    q = truth + rng.normal(scale=sigma, size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return np.concatenate([q, np.zeros((n, 4))], 1).astype(np.float32)


@pytest.mark.parametrize("queue", [0, 1, 2])
def test_one_launch_matches_per_phase_refinements(queue):
    import torch
    from zebsdsolve_amd import _native as nat
    from zebsdsolve_amd import fit
    from zebsdsolve_amd.engine import stream_ptr
    print(f"zb_refine_phases, refine_queue {queue}", flush=True)
    s = _solver(NI, GAN)
    truth = s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=21)
    n = 96
    e = s._e
    lib = nat.load()
    st = stream_ptr()
    pats = fit.PoseField.for_cells(e, 0, np.arange(n)).pats
    items = torch.from_numpy(np.concatenate([e.positions(np.arange(n)), np.arange(n, dtype=np.float32)[:, None]], 1)).cuda()
    tq = truth["quaternions"].reshape(-1, 4).astype(np.float64)
    # phases: Ni, GaN, and Ni again with other seeds (three entries, two masters, two sides)
    which = [0, 1, 0]
    seeds = [torch.from_numpy(_seeds(tq, n, 0.01, 5 + k)).cuda() for k in range(3)]
    out0 = 7
    opts = e.refine_opts()
    fr = e.frame()
    prev = lib.zb_knob_set(b"refine_queue", queue)
    assert prev >= 0
    try:
        for use_order in (False, True):
            orders = []
            for k in range(3):
                o = torch.empty((n,), dtype=torch.int32, device="cuda")
                nat.check(lib.zb_refine_order(fr, items.data_ptr(), seeds[k].data_ptr(), 8, n, o.data_ptr(), st), "order")
                orders.append(o)
            # each phase alone
            alone = []
            for k, pi in enumerate(which):
                rec = torch.full((n + out0, 8), np.nan, dtype=torch.float32, device="cuda")
                opts.order = orders[k].data_ptr() if use_order else None
                nat.check(lib.zb_refine(e.baked(pi)[1][0].data_ptr(), e.grad_half(pi).data_ptr(), e.phases[pi].master.side, fr,
                                        pats.data_ptr(), items.data_ptr(), seeds[k].data_ptr(), 8, n, opts, rec.data_ptr(), out0, st), "refine")
                alone.append(rec)
            opts.order = None
            for count in (2, 3):
                recs = [torch.full((n + out0, 8), np.nan, dtype=torch.float32, device="cuda") for _ in range(count)]
                arr = (nat.RefinePhase * count)(*[
                    nat.RefinePhase(e.baked(which[k])[1][0].data_ptr(), e.grad_half(which[k]).data_ptr(), e.phases[which[k]].master.side,
                                    seeds[k].data_ptr(), 8, orders[k].data_ptr() if use_order else None, recs[k].data_ptr())
                    for k in range(count)])
                work = torch.empty((nat.REFINE_QUEUE_BYTES,), dtype=torch.uint8, device="cuda")
                nat.check(lib.zb_refine_phases(arr, count, fr, pats.data_ptr(), items.data_ptr(), n, opts, out0, work.data_ptr(),
                                               work.numel(), st), "refine phases")
                torch.cuda.synchronize()
                for k in range(count):
                    a, b = alone[k].cpu().numpy(), recs[k].cpu().numpy()
                    assert np.isnan(b[:out0]).all(), (count, k)
                    assert np.array_equal(a[out0:], b[out0:]), (use_order, count, k, np.abs(a - b)[out0:].max())
    finally:
        lib.zb_knob_set(b"refine_queue", prev)


def test_two_phase_scan_records_equal_one_phase_solvers():
    print("two-phase scan against one-phase solvers", flush=True)
    s = _solver(NI, GAN)
    s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=9)
    cells = np.arange(96)
    both = s._e.index_cells(cells)
    pats = np.asarray(s._e.patterns)
    for pi, path in enumerate((NI, GAN)):
        one = _solver(path, patterns=pats)._e.index_cells(cells)
        assert np.array_equal(both[pi][:, :7], one[0][:, :7]), pi


def test_five_phases_go_in_two_launches_with_the_same_records():
    """Past four phases the engine launches the refinement in groups of four: five phases (the
    two masters alternating) give each phase the records a one-phase solver gives it."""
    print("five phases", flush=True)
    s = _solver(NI, GAN, NI, GAN, NI)
    s.synthesize(toy_rows=4, toy_cols=6, toy_grains=24, toy_counts=400, toy_seed=31)
    cells = np.arange(24)
    five = s._e.index_cells(cells)
    assert five.shape[0] == 5
    pats = np.asarray(s._e.patterns)
    ni = _solver(NI, patterns=pats)._e.index_cells(cells)[0]
    gan = _solver(GAN, patterns=pats)._e.index_cells(cells)[0]
    for pi in range(5):
        want = ni if pi % 2 == 0 else gan
        assert np.array_equal(five[pi][:, :7], want[:, :7]), pi
