"""Edge cases of the indexing path on the GPU.
- Batches of every size across the kernels' tile boundaries (32 items a correlation tile, 128 a
  GEMM block, 256 a launch block of the per-item kernels) give each cell the record it gets in
  one large batch, in both product precisions of the correlation (the reference's batch-size
  independence, tests/test_api.py:511-513).
- Degenerate patterns (constant, all zero, one bright pixel) among good ones give finite
  records, are not accepted (their NMI is at or near 0), and leave their neighbours' records
  unchanged.
- Zero-count calls through the C ABI return success and write nothing."""
import ctypes

import numpy as np
import pytest

from tests.conftest import GAN, NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _solver(**cfg):
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.set_geometry(pc=(0.44, 0.24, 0.53), sample_tilt_deg=70, detector_tilt_deg=0, px_um=70 * 8, step_x_um=20, step_y_um=20)
    s.configure(static_background="none", dynamic_background="none", **cfg)
    s.add_phase(NI)
    return s


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_ragged_batches_match_one_batch(precision):
    s = _solver(correlation_precision=precision)
    s.synthesize(toy_rows=16, toy_cols=24, toy_grains=200, toy_counts=400, toy_seed=11)
    n = 16 * 24
    one = s.index_cells(np.arange(n))
    start = 0
    for size in (1, 31, 33, 127, 129, 63):
        cells = np.arange(start, start + size)
        part = s.index_cells(cells)
        print(f"{precision}: batch of {size} from cell {start}", flush=True)
        for k in ("quaternions", "ncc", "nmi"):
            assert np.allclose(part[k][0], one[k][0, cells], atol=1e-6, rtol=0), (size, k, np.abs(part[k][0] - one[k][0, cells]).max())
        start += size
    # a batch of scattered cells in reverse order
    cells = np.arange(n - 1, -1, -3)
    part = s.index_cells(cells)
    assert np.allclose(part["quaternions"][0], one["quaternions"][0, cells], atol=1e-6, rtol=0)
    assert s.index_cells(np.arange(0))["quaternions"].shape == (1, 0, 4)


@pytest.mark.parametrize("two", [False, True])
def test_uneven_chunks_in_one_batch_match_each_chunk_alone(two):
    """Engine.run_batch over a list of uneven chunks (each with its own patterns and local
    indices; u16 and f32 pixels) gives each item, bit for bit, the record that chunk gets alone,
    with one and with two phases, and on a lane's stream."""
    import torch
    print(f"uneven chunks, two phases {two}", flush=True)
    s = _solver()
    if two:
        s.add_phase(GAN)
    s.synthesize(toy_rows=16, toy_cols=24, toy_grains=200, toy_counts=400, toy_seed=23)
    e = s._e
    bounds = [0, 1, 34, 35, 200, 331, 384]
    for dtype in (torch.float32, torch.uint16):
        chunks, items = [], []
        for a, b in zip(bounds[:-1], bounds[1:]):
            cells = np.arange(a, b)
            raw = e._raw_for(cells).to(dtype) if dtype != torch.float32 else e._raw_for(cells).float()
            chunks.append(raw.contiguous())
            items.append(torch.from_numpy(np.concatenate([e.positions(cells), np.arange(b - a, dtype=np.float32)[:, None]], 1)).cuda())
        alone = [e.run_batch(c, it)[0][..., :7].cpu().numpy() for c, it in zip(chunks, items)]
        together = e.run_batch(chunks, items)[0][..., :7].cpu().numpy()
        assert together.shape[1] == 384
        assert np.array_equal(np.concatenate(alone, 1), together), dtype
        lane = e.lanes(2)[1]
        r, _ = e.run_batch(chunks[1:], items[1:], lane=lane)
        torch.cuda.synchronize()
        assert np.array_equal(r[..., :7].cpu().numpy(), together[:, 1:]), dtype


def test_degenerate_patterns_are_finite_and_not_accepted():
    s = _solver()
    s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=5)
    pats = s._e.patterns.reshape(96, 60, 60).copy()
    base = s.index_cells(np.arange(96))
    odd = {3: np.full((60, 60), 400.0, np.float32),  # constant
           17: np.zeros((60, 60), np.float32),      # all zero
           40: np.zeros((60, 60), np.float32)}      # one bright pixel
    odd[40][30, 30] = 1000.0
    for c, p in odd.items():
        pats[c] = p
    s.set_patterns(pats.reshape(8, 12, 60, 60))
    r = s.index()
    got = s.index_cells(np.arange(96))
    bad = np.array(sorted(odd))
    good = np.setdiff1d(np.arange(96), bad)
    for k in ("quaternions", "ncc", "nmi", "step"):
        assert np.isfinite(got[k]).all(), k
    print("degenerate ncc", got["ncc"][0, bad], "nmi", got["nmi"][0, bad], "z", got["z"][0, bad], flush=True)
    # the zero pattern has nothing to correlate, and one value in its histogram
    assert abs(got["ncc"][0, 17]) < 1e-6 and abs(got["nmi"][0, 17]) < 1e-6
    assert (got["nmi"][0, bad] < 0.05).all()
    assert (r["sweep"].reshape(-1)[bad] < 0).all()
    assert (r["sweep"].reshape(-1)[good] >= 0).all()
    for k in ("quaternions", "ncc", "nmi"):
        assert np.allclose(got[k][0, good], base[k][0, good], atol=1e-6, rtol=0), k


def test_zero_count_calls_return_success():
    import torch

    from zebsdsolve_amd import _native as nat
    from zebsdsolve_amd.engine import stream_ptr
    lib = nat.load()
    st = stream_ptr()
    sentinel = torch.full((64,), 7.0, device="cuda")
    p = sentinel.data_ptr()
    mask = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    opts = nat.PrepOpts()
    assert lib.zb_normalize(p, 2, 0, 8, 8, mask, p, st) == 0
    assert lib.zb_prep(p, 2, 0, 8, 8, mask, ctypes.byref(opts), p, st) == 0
    assert lib.zb_fit6_propose(p, p, p, 0, 6, 0.03, 2.0, p, p, p, st) == 0
    assert lib.zb_fit6_accept(p, p, 0, 1e-7, p, p, p, p, p, st) == 0
    torch.cuda.synchronize()
    assert (sentinel == 7.0).all()
    # and a negative count is an error, with a message
    assert lib.zb_fit6_propose(p, p, p, -1, 6, 0.03, 2.0, p, p, p, st) != 0
    assert "bad arguments" in nat.last_error()


@pytest.mark.parametrize("global_filters", [False, True])
def test_index_cells_over_two_streams_matches_one(global_filters):
    """A request large enough to be cut over two streams (Engine.index_lanes) gives every cell
    the record one stream gives it; also with the dynamic background through the filters'
    global-memory workspace (knob prep_global), which each stream holds for itself."""
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    print(f"two streams, filters through global memory {global_filters}", flush=True)
    s = _solver()
    s.synthesize(toy_rows=48, toy_cols=96, toy_grains=400, toy_counts=400, toy_seed=13)
    n = 48 * 96
    prev = lib.zb_knob_set(b"prep_global", int(global_filters))
    try:
        if global_filters:
            s.configure(dynamic_background="subtract")
        s._e.index_lanes = 1
        one = s.index_cells(np.arange(n))
        s._e.index_lanes = 2
        two = s.index_cells(np.arange(n))
        s._e.index_lanes = 1
    finally:
        lib.zb_knob_set(b"prep_global", prev)
    assert len(s._e._lanes) >= 2
    if global_filters:
        assert all("prep_ws" in lane.bufs for lane in s._e._lanes[:2])
    for k in ("quaternions", "ncc", "nmi", "step"):
        assert np.array_equal(one[k], two[k]), (k, np.abs(one[k] - two[k]).max())


def test_refinement_order_is_a_permutation_and_leaves_the_records_alone():
    """zb_refine_order sorts a batch's items by the master region they look up; zb_refine with
    that order gives every item the record it gets in batch order (bit for bit), and the order is
    a permutation of the items grouped by that region."""
    import torch
    from zebsdsolve_amd import _native as nat
    from zebsdsolve_amd.engine import stream_ptr
    s = _solver()
    s.synthesize(toy_rows=40, toy_cols=60, toy_grains=2400, toy_counts=400, toy_seed=13)
    n = 40 * 60
    s.configure(refine_order=False)
    plain = s.index_cells(np.arange(n))
    s.configure(refine_order=True)
    ordered = s.index_cells(np.arange(n))
    for k in ("quaternions", "ncc", "nmi", "step"):
        assert np.array_equal(plain[k], ordered[k]), k
    # the permutation itself, on random seeds
    e = s._e
    lib = nat.load()
    rng = np.random.default_rng(3)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    seeds = torch.from_numpy(np.concatenate([q, np.zeros((n, 4))], 1).astype(np.float32)).cuda()
    items = torch.from_numpy(np.concatenate([e.positions(np.arange(n)), np.arange(n, dtype=np.float32)[:, None]], 1)).cuda()
    order = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    nat.check(lib.zb_refine_order(e.frame(), items.data_ptr(), seeds.data_ptr(), 8, n, order.data_ptr(), stream_ptr()), "order")
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(n))
    # neighbours in the order look up nearby master regions: the detector centre's direction at
    # consecutive items is far closer than at random pairs
    fr = e.frame()
    h, w = e.pattern_hw()
    t = items[:, :3].cpu().numpy().astype(np.float64)
    u = np.asarray(fr.c0) * (h - 1) / 2 + np.asarray(fr.c1) * (w - 1) / 2 + t
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    # q . u (q acting as q u q*)
    tt = 2 * np.cross(q[:, 1:], u)
    v = u + q[:, :1] * tt + np.cross(q[:, 1:], tt)
    near = np.mean(np.sum(v[o[1:]] * v[o[:-1]], 1))
    rand = np.mean(np.sum(v[1:] * v[:-1], 1))
    assert near > 0.95 and near > rand + 0.5, (near, rand)


def test_frames_past_65536_pixels_index_and_refine_as_the_oracle():
    """240 x 320 patterns (76 800 pixels: the refinement's pixel list no longer fits LDS, so the
    whole frame is walked with the mask tested per pixel): a toy scan is recovered to the
    reference's thresholds (median < 0.15 deg, max < 1 deg), the refinement's NCC from the same
    seed matches the float64 oracle's, and zb_sums matches the oracle's sums."""
    from oracle import pipeline as op
    from zebsdsolve_amd import fit
    print("240 x 320 frames", flush=True)
    s = _solver()
    truth = s.synthesize(toy_rows=4, toy_cols=6, toy_grains=24, toy_counts=400, toy_seed=17, toy_height=240, toy_width=320)
    r = s.index()
    d = np.degrees(s.disorientation(r["quaternions"].reshape(-1, 4), truth["quaternions"].reshape(-1, 4)))
    assert np.median(d) < 0.15 and d.max() < 1.0, d
    e = s._e
    rec = e.index_cells(np.arange(2))[0]
    # the oracle's NCC at the GPU's refined orientation (iters 0: the score there)
    from zebsdsolve_amd.geometry import frame_of
    f = frame_of(e.geo, 240, 320, True, 1.0)
    _, (_, grad_d, host) = e.baked(0)
    side = e.phases[0].master.side
    g = grad_d.cpu().numpy().reshape(-1, 4)
    t = e.positions(np.arange(2))
    raw = np.asarray(e.patterns).reshape(-1, 240, 320)[:2].astype(np.float64)
    for b in range(2):
        pn = op.normalize(raw[b], f["mask"])
        _, ncc, _ = op.refine(g, side, f, pn, t[b].astype(np.float64), rec[b, :4].astype(np.float64), iters=0)
        assert abs(ncc - rec[b, 4]) < 1e-3, (b, ncc, rec[b, 4])
    field = fit.PoseField.for_cells(e, 0, np.arange(2))
    poses = np.concatenate([t, np.arange(2, dtype=np.float32)[:, None], rec[:, :4]], 1).astype(np.float32)
    got = field.sums(poses).cpu().numpy().astype(np.float64)
    pats_n = field.pats.cpu().numpy()
    for b in range(2):
        ref = op.sums6(g, side, f, pats_n[b].astype(np.float64), poses[b, :3].astype(np.float64), poses[b, 4:8].astype(np.float64))
        assert got[b, 0] == ref[0] and np.abs(got[b] - ref).max() < 1e-4 * np.abs(ref).max()
