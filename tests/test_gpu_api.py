"""The Solver / index surface on the GPU against ground truth, with the reference's own thresholds
(tests/test_api.py): toy scans recovered to median < 0.15 deg and max < 1 deg, every cell indexed
in the fundamental zone; noise cells far below chance + 15 sigma and patterns far above; two
candidate phases told apart; results independent of the batch size."""
import numpy as np
import pytest

from tests.conftest import GAN, NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _solver():
    import zebsdsolve_amd as ze
    s = ze.Solver()
    s.set_geometry(pc=(0.44, 0.24, 0.53), sample_tilt_deg=70, detector_tilt_deg=0, px_um=70 * 8, step_x_um=20, step_y_um=20)
    # as the reference's test (tests/test_api.py:199): the synthetic scans without filters
    s.configure(static_background="none", dynamic_background="none")
    return s


def test_toy_scan_is_recovered_in_the_fundamental_zone():
    import zebsdsolve_amd as ze
    s = _solver()
    s.add_phase(NI)
    truth = s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=3)
    r = s.index()
    q = r["quaternions"].reshape(-1, 4)
    d = np.degrees(s.disorientation(q, truth["quaternions"].reshape(-1, 4)))
    assert np.median(d) < 0.15 and d.max() < 1.0, (np.median(d), d.max())
    assert (r["phase"] == 0).all()
    assert (r["quaternions"][..., 0] >= np.cos(np.radians(62.81) / 2) - 1e-4).all()
    assert r["euler"].shape == (8, 12, 3) and r["ncc"].shape == (8, 12)
    # the peak alone is within the grid's reach
    s.configure(refine=False)
    r0 = s.index()
    d0 = np.degrees(s.disorientation(r0["quaternions"].reshape(-1, 4), truth["quaternions"].reshape(-1, 4)))
    assert np.median(d0) < 1.5 and np.percentile(d0, 95) < 3.0
    s.configure(refine=True)
    # batch-size independence (test_api.py:511-513)
    one = s.index_cells(np.arange(96))
    part = s.index_cells(np.arange(40, 56))
    assert np.allclose(one["quaternions"][0, 40:56], part["quaternions"][0], atol=1e-6)
    with pytest.raises(ze.Error):
        s.configure(no_such_option=1)
    with pytest.raises(ze.Error):
        s.configure(bandwidth=3)
    assert s.settings["bandwidth"] == 64


def test_two_sweeps_accept_most_at_the_first_bandwidth():
    """sweeps = 2: the first pass at sweep_bandwidth (32) accepts the well-indexed cells, the
    second (at 64) takes only the rest (src/engine.zig:2426-2465); every cell keeps its best
    match, and the map is recovered as well as by one sweep at 64."""
    s = _solver()
    s.add_phase(NI)
    truth = s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=7)
    tq = truth["quaternions"].reshape(-1, 4)
    s.configure(sweeps=2, sweep_bandwidth=32, bandwidth=64)
    seen = []
    r = s.index(progress=lambda st: seen.append((st["sweep"], st["bandwidth"], st["total"])))
    assert all(bw == (32 if sw == 0 else 64) for sw, bw, _ in seen), seen
    sweep = r["sweep"].reshape(-1)
    assert set(np.unique(sweep)) <= {-1, 0, 1}
    assert (sweep == 0).sum() >= 0.8 * 96, np.bincount(sweep + 1)
    if (sweep == 1).any():
        # the second sweep visited only what the first left
        assert seen[-1][2] == 96 - (sweep == 0).sum()
    d = np.degrees(s.disorientation(r["quaternions"].reshape(-1, 4), tq))
    assert np.median(d) < 0.15 and d.max() < 1.0, (np.median(d), d.max())
    assert (r["phase"] == 0).all()
    st = s.index_status
    assert st["sweeps"] == 2 and st["region"] == 96


def test_noise_is_not_accepted_and_patterns_are():
    s = _solver()
    s.add_phase(NI)
    truth = s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=5)
    rng = np.random.default_rng(0)
    pats = s._e.patterns.reshape(96, 60, 60).copy()
    noise = rng.choice(96, 32, replace=False)
    pats[noise] = rng.normal(400, 20, size=(32, 60, 60)).astype(np.float32)
    s.set_patterns(pats.reshape(8, 12, 60, 60))
    r = s.index()
    z = r["z"].reshape(-1)
    is_noise = np.zeros(96, bool)
    is_noise[noise] = True
    assert z[~is_noise].min() > 15 and z[is_noise].max() < 8, (z[~is_noise].min(), z[is_noise].max())
    assert ((r["sweep"].reshape(-1) >= 0) == ~is_noise).all()
    one = s.index_cells(np.arange(96))
    assert np.degrees(one["step"][0][~is_noise]).max() < 0.2


def test_two_phases_are_told_apart():
    s = _solver()
    s.add_phase(NI)
    s.add_phase(GAN)
    truth = s.synthesize(toy_rows=8, toy_cols=12, toy_grains=96, toy_counts=400, toy_seed=9)
    r = s.index()
    ph = r["phase"].reshape(-1)
    tp = truth["phase"].reshape(-1)
    assert (ph == tp).mean() > 0.97
    ok = ph == tp
    d = []
    for p in (0, 1):
        sel = ok & (tp == p)
        if sel.any():
            d.append(np.degrees(s.disorientation(r["quaternions"].reshape(-1, 4)[sel], truth["quaternions"].reshape(-1, 4)[sel], phase=p)))
    d = np.concatenate(d)
    assert np.median(d) < 0.3


def test_index_function_returns_a_result():
    import zebsdsolve_amd as ze
    s = _solver()
    s.add_phase(NI)
    truth = s.synthesize(toy_rows=4, toy_cols=5, toy_grains=20, toy_counts=400, toy_seed=2)
    pats = s._e.patterns
    res = ze.index(pats, [NI], pc=(0.44, 0.24, 0.53), px_um=560, step_x_um=20, step_y_um=20, bandwidth=48)
    assert res.quaternions.shape == (4, 5, 4) and res.phase.shape == (4, 5)
    d = np.degrees(s.disorientation(res.quaternions.reshape(-1, 4), truth["quaternions"].reshape(-1, 4)))
    assert np.median(d) < 0.15


def test_graph_replays_match_eager_batches():
    """A batch captured as a HIP graph and replayed three times on other patterns, with eager
    batches and synchronizations between the replays: every replay's records equal the eager
    run's on the same inputs (the kernels are deterministic: integer atomics and a max-key
    argmax only)."""
    import torch
    from tests import cases
    from tests.conftest import NI
    from zebsdsolve_amd import engine
    from zebsdsolve_amd.geometry import Geometry
    print("graph replays", flush=True)
    n = 96
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none", bandwidth=32)
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.set_patterns(np.zeros((1, n, 60, 60), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    gb = e.graph(n, 60, 60)
    t = torch.from_numpy(e.positions(np.arange(n))).cuda()
    items = torch.cat([t, torch.arange(n, dtype=torch.float32, device="cuda")[:, None]], 1)
    for seed in (1, 2, 3):
        pats, _ = cases.simulate(host, side, n, seed=seed, counts=400.0)
        raw = torch.from_numpy(pats).cuda()
        got = gb(raw, items).clone()
        torch.cuda.synchronize()
        want, stamp = e.run_batch(raw, items)
        torch.cuda.synchronize()
        g, w = got.cpu().numpy(), want.cpu().numpy()
        assert np.array_equal(g[..., :7], w[..., :7]), seed
        assert np.all(g[..., 7] == np.float32(gb.stamp)) and np.all(w[..., 7] == np.float32(stamp))


def test_tune_climbs_the_knobs_and_keeps_the_results():
    """Solver.tune (python/zebsdsolve/__init__.py tune): it times a sample, leaves knobs within
    their ranges, reports before/after and a summary, stops when progress says so, and the map
    indexes to the same records afterwards, bit for bit (the knobs cut the work, nothing else).
    The library's knobs are process-wide: put back to their defaults at the end."""
    s = _solver()
    s.add_phase(NI)
    s.synthesize(toy_rows=16, toy_cols=24, toy_grains=96, toy_counts=400, toy_seed=5)
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    r0 = s.index()
    calls = []
    out = s.tune(lambda k: calls.append(k) or len(calls) < 5)
    keys = {"step_cells", "spec_budget_mb", "index_lanes", "batch_items", "refine_queue", "rings_group"}
    assert set(out["after"]) == keys
    assert len(calls) <= 5 and "tuned on" in out["summary"]
    try:
        # a whole run: every knob tried, the library's launch-shape knobs included
        out = s.tune()
        a = out["after"]
        assert set(a) == keys
        assert 256 <= a["step_cells"] <= 1 << 20 and 64 <= a["spec_budget_mb"] <= 1 << 16 and 1 <= a["index_lanes"] <= 4
        assert 1024 <= a["batch_items"] <= 1 << 16
        assert a["refine_queue"] in (0, 1, 2) and a["rings_group"] in (0, 4, 8, 16, 32)
        assert lib.zb_knob_get(b"refine_queue") == a["refine_queue"]
        r1 = s.index()
        for k in ("quaternions", "ncc", "nmi", "phase"):
            assert np.array_equal(r0[k], r1[k]), k
    finally:
        lib.zb_knob_set(b"refine_queue", 0)
        lib.zb_knob_set(b"rings_group", 0)


def test_tune_schedule_keeps_a_schedule_that_agrees_with_one_sweep():
    """Solver.tune_schedule times two-sweep bandwidth schedules against the one sweep on a sample
    and keeps the fastest that agrees with it: the chosen schedule is in the settings, every
    schedule is reported with its time, and indexing the map with the chosen schedule accepts
    as many cells as the one sweep, in the same phase and within 0.1 deg of it."""
    s = _solver()
    s.add_phase(NI)
    s.synthesize(toy_rows=32, toy_cols=48, toy_grains=300, toy_counts=400, toy_seed=19)
    ref = s.index()
    sb0 = s.settings["sweep_bandwidth"]
    out = s.tune_schedule(bandwidths=(24, 32, 48))
    print(out["summary"], flush=True)
    sw = [(a, b) for a, b, _, _ in out["schedules"]]
    assert sw[0] == (1, sb0) and set(sw[1:]) == {(2, 24), (2, 32), (2, 48)}
    assert all(t > 0 for _, _, t, _ in out["schedules"]) and out["schedules"][0][3]
    assert (s.settings["sweeps"], s.settings["sweep_bandwidth"]) == out["chosen"]
    assert "chosen" in out["summary"]
    got = s.index()
    acc_r, acc_g = ref["sweep"].reshape(-1) >= 0, got["sweep"].reshape(-1) >= 0
    assert acc_g.sum() >= acc_r.sum() - 0.005 * acc_r.size
    both = acc_r & acc_g
    d = np.degrees(s.disorientation(got["quaternions"].reshape(-1, 4)[both], ref["quaternions"].reshape(-1, 4)[both]))
    assert np.mean(d > 0.1) <= 0.005
    s.configure(sweeps=1)


def test_steps_launched_ahead_match_one_step_and_follow_configure():
    """index_step launches the sweep's next step before landing the current one: many small
    steps give what one step over the whole map gives, and a configure between steps drops the
    step launched ahead (the next step follows the new settings)."""
    s = _solver()
    s.add_phase(NI)
    s.synthesize(toy_rows=16, toy_cols=24, toy_grains=200, toy_counts=400, toy_seed=9)
    s._step_cells = 1 << 20
    one = s.index()
    s._step_cells = 64
    s.configure(batch=1)  # steps of 64 cells: six of them
    many = s.index()
    for k in ("quaternions", "ncc", "nmi", "phase", "sweep"):
        assert np.array_equal(one[k], many[k]), k
    s.index_begin()
    s.index_step()
    assert s._pending is not None
    s.configure(refine=False)
    assert s._pending is None
    s.index_step()
    got = s._res["ncc"][64:128]
    want = s.index_cells(np.arange(64, 128))["ncc"][0]
    assert np.array_equal(got, want)
    s.configure(refine=True)
    with_refine = s.index_cells(np.arange(64, 128))["ncc"][0]
    assert not np.array_equal(got, with_refine)


def test_chance_is_measured_again_after_what_it_depends_on_changes():
    """The chance calibration (median and spread of NMI at random orientations) is dropped by
    every setter: a filter toggled, or a phase swapped for another, gives a new measurement (the
    reference keys it on every setting, the geometry and the scan's and phases' serials,
    src/engine.zig:2202-2240). Index progress and the chance arrive as events; profile_report is
    the reference's text."""
    s = _solver()
    s.add_phase(NI)
    s.synthesize(toy_rows=6, toy_cols=8, toy_grains=12, toy_counts=400, toy_seed=2)
    s.events()
    s.profile(True)
    s.index()
    ev = s.events()
    assert any(e["kind"] == "progress" and e["stage"] == "index" and e["done"] == e["total"] == 48 for e in ev)
    assert any(e["kind"] == "log" and e["text"].startswith("chance, phase 0:") for e in ev)
    rep = s.profile_report()
    s.profile(False)
    lines = rep.splitlines()
    assert lines[0].endswith("dispatches  total") and any("k_refine" in ln for ln in lines[1:])
    ms = [float(ln.split()[0]) for ln in lines[1:]]
    assert ms == sorted(ms, reverse=True)
    # zb_profile_select: only the named kernels are timed, then every kernel again
    from zebsdsolve_amd import _native as nat
    nat.profile_select(["k_refine", "k_vgemm"])
    nat.profile(True)
    s.index_cells(np.arange(48))
    only = nat.profile_report()
    nat.profile(False)
    nat.profile_select(None)
    assert set(only) == {"k_refine", "k_vgemm"}, only
    nat.profile(True)
    s.index_cells(np.arange(48))
    every = nat.profile_report()
    nat.profile(False)
    assert {"k_refine", "k_vgemm", "k_apeak", "k_rings"} <= set(every)
    c0 = s.chance(0)
    s.configure(dynamic_background="subtract")
    assert s._chance is None
    c1 = s.chance(0)
    assert c1 != c0
    s.configure(dynamic_background="none")
    assert s.chance(0) == c0
    s.clear_phases()
    s.add_phase(GAN)
    c2 = s.chance(0)
    assert c2 != c0
