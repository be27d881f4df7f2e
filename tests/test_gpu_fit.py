"""The six-direction fit's sums (zb_sums) and the differentiable NCC built on them, against the
float64 oracle (whose gradient its own test pins to central differences of the NCC).

Tolerances: float32 sums over 2828 pixels against float64, stated at each assertion.
"""
import numpy as np
import pytest

from tests import cases
from tests.conftest import NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

N = 8


@pytest.fixture(scope="module")
def fitcase():
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import engine, fit
    from zebsdsolve_amd.geometry import Geometry
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")  # the path alone, as the oracle
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.set_patterns(np.zeros((1, N, 60, 60), np.float32))
    _, (_, grad_d, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, truth = cases.simulate(host, side, N, seed=17, counts=400.0)
    e.set_patterns(pats.reshape(1, N, 60, 60))
    fr, t = cases.oracle_frame(60, 60)
    field = fit.PoseField.for_cells(e, 0, np.arange(N))
    rng = np.random.default_rng(5)
    # poses near the truth: the detector position off by up to a pixel, the crystal by ~0.5 deg
    tt = t[None, :] + rng.uniform(-1, 1, size=(N, 3))
    qq = np.stack([op.qmul(op.qexp(rng.normal(size=3) * 0.005), truth[b]) for b in range(N)])
    poses = np.concatenate([tt, np.arange(N)[:, None], qq], 1).astype(np.float32)
    torch.cuda.synchronize()
    return {"e": e, "op": op, "fit": fit, "field": field, "fr": fr, "poses": poses, "side": side,
            "grad": grad_d.cpu().numpy().reshape(-1, 4), "pats_n": field.pats.cpu().numpy()}


def test_sums_match_the_oracle(fitcase):
    op, field = fitcase["op"], fitcase["field"]
    print("zb_sums vs oracle", flush=True)
    got = field.sums(fitcase["poses"]).cpu().numpy().astype(np.float64)
    for b in range(N):
        p = fitcase["poses"][b].astype(np.float64)
        ref = op.sums6(fitcase["grad"], fitcase["side"], fitcase["fr"], fitcase["pats_n"][b].astype(np.float64), p[:3], p[4:8])
        # each sum to float32 accumulation over 2828 terms: 1e-4 of the largest sum of its kind
        scale = np.abs(ref).max()
        assert got[b, 0] == ref[0]
        assert np.abs(got[b] - ref).max() < 1e-4 * scale, (b, np.abs(got[b] - ref).max(), scale)
        ok, n1, g1, H1 = op.local6(ref)
        ok2, n2, g2, H2 = fitcase["fit"].local_of(got[b])
        assert ok and ok2
        # the NCC to 1e-5, its gradient to 1e-3 of its size
        assert abs(n1 - n2) < 1e-5
        assert np.abs(g1 - g2).max() < 1e-3 * np.abs(g1).max()


def test_rotation_only_zeroes_the_shift_sums(fitcase):
    field = fitcase["field"]
    a = field.sums(fitcase["poses"]).cpu().numpy()
    b = field.sums(fitcase["poses"], flags=fitcase["fit"].ROTATION_ONLY).cpu().numpy()
    shift = np.array([(4 <= i < 7) or (10 <= i < 13) or (16 <= i < 19) or (22 <= i < 37) for i in range(43)])
    assert np.all(b[:, shift] == 0) and np.array_equal(a[:, ~shift], b[:, ~shift])


def test_differentiable_ncc_backward_is_the_oracle_gradient(fitcase):
    import torch
    op, fit = fitcase["op"], fitcase["fit"]
    rng = np.random.default_rng(6)
    delta0 = np.concatenate([rng.uniform(-0.3, 0.3, size=(N, 3)), rng.normal(size=(N, 3)) * 0.002], 1)
    delta = torch.tensor(delta0, dtype=torch.float64, requires_grad=True)
    ncc = fit.PoseNcc.apply(delta, fitcase["field"], fitcase["poses"])
    ncc.sum().backward()
    gd = delta.grad.numpy()
    moved = fit.moved(fitcase["poses"], delta0)
    for b in range(N):
        pb = fitcase["pats_n"][b].astype(np.float64)

        def f(dd):
            p = fit.moved(fitcase["poses"][b:b + 1], dd[None])[0]
            return op.local6(op.sums6(fitcase["grad"], fitcase["side"], fitcase["fr"], pb, p[:3], p[4:8]))[1]

        assert abs(ncc[b].item() - f(delta0[b])) < 1e-5
        # the float64 gradient at the moved pose, carried to delta through exp's left Jacobian
        _, _, g, _ = op.local6(op.sums6(fitcase["grad"], fitcase["side"], fitcase["fr"], pb, moved[b, :3], moved[b, 4:8]))
        J = fit.left_jacobian(delta0[b, 3:][None])[0]
        want = np.concatenate([g[:3], J.T @ g[3:]])
        assert np.abs(gd[b] - want).max() < 1e-3 * np.abs(want).max(), (b, gd[b], want)
        # and central differences of the float64 NCC in delta (1e-5 of the pattern height in
        # pixels, 1e-4 rad): the bilinear master is only piecewise smooth, so 10 %
        fd = np.zeros(6)
        for i in range(6):
            h = 60 * 1e-5 if i < 3 else 1e-4
            e = np.zeros(6)
            e[i] = h
            fd[i] = (f(delta0[b] + e) - f(delta0[b] - e)) / (2 * h)
        assert np.abs(fd[:3] - gd[b, :3]).max() < 0.1 * np.abs(fd[:3]).max(), (b, fd, gd[b])
        assert np.abs(fd[3:] - gd[b, 3:]).max() < 0.1 * np.abs(fd[3:]).max(), (b, fd, gd[b])
    assert moved.shape == (N, 8)


def _qangle(a, b):
    """Degrees between unit quaternions (either sign), from the vector part of a* b (well
    conditioned near zero, unlike the arccos of a dot product of float32 quaternions)."""
    a = np.asarray(a, np.float64) / np.linalg.norm(a, axis=-1, keepdims=True)
    b = np.asarray(b, np.float64) / np.linalg.norm(b, axis=-1, keepdims=True)
    w = np.sum(a * b, -1)
    v = a[..., :1] * b[..., 1:] - b[..., :1] * a[..., 1:] - np.cross(a[..., 1:], b[..., 1:])
    return np.degrees(2 * np.arctan2(np.linalg.norm(v, axis=-1), np.abs(w)))


def test_fit6_matches_the_oracles_fit(fitcase):
    """The device fit (zb_fit6_propose / zb_sums / zb_fit6_accept) and the oracle's, from the
    same poses on the same noisy patterns: both climb to the same NCC maximum (the NCC to 1e-5,
    the orientation to 0.01 deg, the detector position to 0.02 px)."""
    op, fit, field = fitcase["op"], fitcase["fit"], fitcase["field"]
    print("fit6 vs oracle", flush=True)
    poses, ncc, state = fit.fit6(field, fitcase["poses"], iters=20)
    got, ncc = poses.cpu().numpy().astype(np.float64), ncc.cpu().numpy()
    for b in range(3):
        p = fitcase["poses"][b].astype(np.float64)
        t, q, n = op.fit6(fitcase["grad"], fitcase["side"], fitcase["fr"], fitcase["pats_n"][b].astype(np.float64), p[:3], p[4:8], iters=20)
        d = _qangle(q, got[b, 4:8])
        assert abs(ncc[b] - n) < 1e-5 and d < 0.01 and np.abs(got[b, :3] - t).max() < 0.02, (b, ncc[b] - n, d, got[b, :3] - t)
    assert (state.cpu().numpy() != 2).all()


def test_fit6_to_convergence_from_config4_starts():
    """The differentiation form's starts (80 x 80 patterns; the crystal 0.5 - 2 deg off, uniform
    angle and axis; the pattern centre up to 0.02 off per coordinate): on noise-free patterns the
    NCC maximum is the truth, and the fit reaches it (NCC > 0.999, orientation within 0.02 deg,
    the detector position within 0.05 px) for every pattern."""
    import torch
    from oracle import pipeline as op
    from zebsdsolve_amd import engine, fit
    from zebsdsolve_amd.geometry import Geometry
    print("fit6 config 4 starts", flush=True)
    n, S = 16, 80
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    e.set_patterns(np.zeros((1, n, S, S), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, truth = cases.simulate(host, side, n, h=S, w=S, seed=29)
    e.set_patterns(pats.reshape(1, n, S, S))
    _, t = cases.oracle_frame(S, S)
    field = fit.PoseField.for_cells(e, 0, np.arange(n))
    rng = np.random.default_rng(6)
    ang = np.radians(rng.uniform(0.5, 2.0, n))
    ax = rng.normal(size=(n, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    q0 = np.stack([op.qmul(op.qexp(ax[b] * ang[b]), truth[b]) for b in range(n)])
    dpc = rng.uniform(-0.02, 0.02, size=(n, 3))
    R = op.bruker_R(70.0, 0.0)
    t0 = np.stack([R @ np.array([-dpc[b, 1] * S, -dpc[b, 0] * S, dpc[b, 2] * S]) for b in range(n)]) + t[None, :]
    poses0 = np.concatenate([t0, np.arange(n)[:, None], q0], 1).astype(np.float32)
    poses, ncc, state = fit.fit6(field, poses0, iters=40)
    torch.cuda.synchronize()
    got, ncc = poses.cpu().numpy().astype(np.float64), ncc.cpu().numpy()
    d = _qangle(got[:, 4:8], truth)
    dt = np.abs(got[:, :3] - t[None, :]).max(1)
    print(f"fit6: NCC min {ncc.min():.5f}, orientation max {d.max():.4f} deg, position max {dt.max():.3f} px", flush=True)
    assert ncc.min() > 0.999 and d.max() < 0.02 and dt.max() < 0.05, (ncc, d, dt)


def test_ncc_in_pattern_centre_coordinates_against_differences_and_float64(fitcase):
    """PoseNccPC: the NCC as a function of the crystal rotation (tangent phi at q0) and the
    pattern centre pc, t(pc) = R (0.5 - pcy h, 0.5 - pcx w, pcz h) (src/geometry.zig:71-75). Its
    backward pass against central differences of its own forward pass (h = 1e-4 rad, 1e-5 of
    the pattern centre: the bilinear master is piecewise smooth, so 10 % of the gradient's size)
    and against the float64 gradient (the oracle's sums at the same pose, carried to phi and pc
    by the same chain rule written out here: 1e-3 of its size)."""
    import torch
    op, fit = fitcase["op"], fitcase["fit"]
    print("PoseNccPC", flush=True)
    H = W = 60
    R = op.bruker_R(70.0, 0.0)
    poses = fitcase["poses"].astype(np.float64)
    # the pattern centre of each pose's detector position: t = R v, v = (0.5 - pcy H, 0.5 - pcx W, pcz H)
    v = poses[:, :3] @ R
    pc0 = np.stack([(0.5 - v[:, 1]) / W, (0.5 - v[:, 0]) / H, v[:, 2] / H], 1)
    rng = np.random.default_rng(9)
    phi0 = rng.normal(size=(N, 3)) * 0.002
    q0 = poses[:, 4:8]
    idx = np.arange(N, dtype=np.float64)
    phi = torch.tensor(phi0, dtype=torch.float64, requires_grad=True)
    pc = torch.tensor(pc0, dtype=torch.float64, requires_grad=True)
    ncc = fit.PoseNccPC.apply(phi, pc, fitcase["field"], q0, R, H, W, idx)
    ncc.sum().backward()
    g_phi, g_pc = phi.grad.numpy(), pc.grad.numpy()

    def f(ph, c):
        with torch.no_grad():
            return fit.PoseNccPC.apply(torch.tensor(ph), torch.tensor(c), fitcase["field"], q0, R, H, W, idx).numpy()

    for i in range(3):
        e = np.zeros((N, 3))
        e[:, i] = 1e-4
        fd = (f(phi0 + e, pc0) - f(phi0 - e, pc0)) / 2e-4
        assert np.all(np.abs(fd - g_phi[:, i]) < 0.1 * np.abs(g_phi).max(1)), (i, fd, g_phi[:, i])
        e = np.zeros((N, 3))
        e[:, i] = 1e-5
        fd = (f(phi0, pc0 + e) - f(phi0, pc0 - e)) / 2e-5
        assert np.all(np.abs(fd - g_pc[:, i]) < 0.1 * np.abs(g_pc).max(1)), (i, fd, g_pc[:, i])
    for b in range(N):
        pb = fitcase["pats_n"][b].astype(np.float64)
        t = R @ np.array([0.5 - pc0[b, 1] * H, 0.5 - pc0[b, 0] * W, pc0[b, 2] * H])
        q = op.qmul(op.qexp(phi0[b]), q0[b])
        q = q / np.linalg.norm(q)
        _, n64, g, _ = op.local6(op.sums6(fitcase["grad"], fitcase["side"], fitcase["fr"], pb, t, q))
        assert abs(ncc[b].item() - n64) < 1e-5
        J = fit.left_jacobian(phi0[b][None])[0]
        want_phi = J.T @ g[3:]
        dt = np.stack([-W * R[:, 1], -H * R[:, 0], H * R[:, 2]])  # d t / d (pcx, pcy, pcz), rows
        want_pc = dt @ g[:3]
        assert np.abs(g_phi[b] - want_phi).max() < 1e-3 * np.abs(want_phi).max(), (b, g_phi[b], want_phi)
        assert np.abs(g_pc[b] - want_pc).max() < 1e-3 * np.abs(want_pc).max(), (b, g_pc[b], want_pc)
