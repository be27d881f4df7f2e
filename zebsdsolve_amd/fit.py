"""The six-direction local fit of a pattern's pose: the NCC of each pattern with a phase's master
at a pose (detector position t, crystal orientation q), its gradient and Gauss-Newton matrix
over a shift of the rays' origin (pixels, along the sample's x, y, z) and a crystal rotation
(radians, about x, y, z), from the 43 sums of `zb_sums` (detector.wgsl sums; planefit.zig
177-201 localOf, which the reference's pattern-centre and plane fits step with).

`PoseNcc` makes the NCC a torch function of those six directions, and `PoseNccPC` of the crystal
rotation and the pattern centre (the detector position t(PC) of src/geometry.zig:71-75): each
forward pass is zb_sums (the adjoint of the projection: the derivative of every masked pixel's
master value along each direction, summed against the pattern) and zb_local6 (NCC, gradient,
Gauss-Newton matrix) on the device, its backward pass the gradient from the same sums, chained
through the left Jacobian of exp and dt/dPC. `fit6` takes its Levenberg-Marquardt steps with the
gradient of a `PoseNcc` backward pass and the matrix of the same forward pass.
"""
import numpy as np
import torch

from . import _native as nat
from . import rot

N_SUMS = 43
ROTATION_ONLY = 2
# the local model's floats per item: NCC, ok, g[6], the upper triangle of H [21]
LOCAL = 29


def local_of(sums):
    """(ok, ncc, g [..., 6], H [..., 6, 6]) in float64 from sums [..., 43] (planefit.zig:177-201):
    ok is False where the master is flat under the mask."""
    s = np.asarray(sums, np.float64)
    n = np.maximum(s[..., 0], 1.0)
    mean = s[..., 1] / n
    var_n = s[..., 2] - n * mean * mean
    ok = var_n > 1e-12
    inv = np.where(ok, 1.0 / np.sqrt(np.where(ok, var_n, 1.0)), 0.0)
    ncc = s[..., 3] * inv
    sd = s[..., 4:10]
    au = (s[..., 10:16] - mean[..., None] * sd) * inv[..., None]
    g = (s[..., 16:22] - ncc[..., None] * au) * inv[..., None]
    H = np.zeros(s.shape[:-1] + (6, 6))
    o = 22
    for i in range(6):
        for j in range(i, 6):
            v = (s[..., o] - sd[..., i] * sd[..., j] / n - au[..., i] * au[..., j]) * inv * inv
            H[..., i, j] = v
            H[..., j, i] = v
            o += 1
    ncc = np.where(ok, ncc, 0.0)
    return ok, ncc, g, H


def left_jacobian(phi):
    """J with exp(phi + d) = exp(J d) exp(phi) to first order in d (SO(3)), [..., 3, 3]."""
    phi = np.asarray(phi, np.float64)
    th = np.linalg.norm(phi, axis=-1)
    K = np.zeros(phi.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2] = -phi[..., 2], phi[..., 1]
    K[..., 1, 0], K[..., 1, 2] = phi[..., 2], -phi[..., 0]
    K[..., 2, 0], K[..., 2, 1] = -phi[..., 1], phi[..., 0]
    small = th < 1e-6
    ts = np.where(small, 1.0, th)
    a = np.where(small, 0.5 - th * th / 24, (1 - np.cos(ts)) / (ts * ts))
    b = np.where(small, 1.0 / 6 - th * th / 120, (ts - np.sin(ts)) / (ts ** 3))
    return np.eye(3) + a[..., None, None] * K + b[..., None, None] * (K @ K)


class PoseField:
    """Phase pi's master against given normalized patterns: the sums at any poses. `pats` is a
    device tensor [n, h, w] of patterns filtered and normalized as the engine's settings say
    (Engine.prep_into)."""

    def __init__(self, engine, pi, pats):
        self.e = engine
        self.pi = pi
        self.pats = pats.contiguous()
        _, (_, self.grad, _) = engine.baked(pi)
        self.side = engine.phases[pi].master.side
        self.frame = engine.frame()

    @classmethod
    def for_cells(cls, engine, pi, cells):
        """The field of the patterns of scan cells, normalized on the device."""
        from .engine import stream_ptr
        dev = engine._device()
        cells = np.asarray(cells, np.int64).reshape(-1)
        raw = engine._raw_for(cells)
        h, w = engine.pattern_hw()
        pats = torch.empty((len(cells), h, w), dtype=torch.float32, device=dev)
        engine.prep_into(raw, len(cells), h, w, pats, stream_ptr())
        return cls(engine, pi, pats)

    def sums(self, poses, flags=0, state=None):
        """Sums [n, 43] (device, float32) at poses [n, 8] = (t, pattern index, q) (device or
        host, float32). With `state` (int32 [n], the fit's), only the items whose state is 0
        are summed (zb_sums_active); the others' rows are zero."""
        from .engine import stream_ptr
        lib = nat.load()
        p = torch.as_tensor(poses, dtype=torch.float32, device=self.pats.device).contiguous()
        n = p.shape[0]
        if state is None:
            out = torch.empty((n, N_SUMS), dtype=torch.float32, device=self.pats.device)
            nat.check(lib.zb_sums(self.grad.data_ptr(), self.side, self.frame, self.pats.data_ptr(), p.data_ptr(), n, int(flags),
                                  out.data_ptr(), stream_ptr()), "sums")
        else:
            out = torch.zeros((n, N_SUMS), dtype=torch.float32, device=self.pats.device)
            nat.check(lib.zb_sums_active(self.grad.data_ptr(), self.side, self.frame, self.pats.data_ptr(), p.data_ptr(), n, int(flags),
                                         state.data_ptr(), out.data_ptr(), stream_ptr()), "sums")
        return out

    def local(self, poses, flags=0):
        return local_of(self.sums(poses, flags).cpu().numpy())

    def local6(self, poses, flags=0, state=None):
        """The local model at poses on the device: [n, 29] float64 = (NCC, ok, g[6], the
        Gauss-Newton matrix's upper triangle [21]) (zb_sums, then zb_local6); with `state`, of
        the items still moving (the others' rows say flat)."""
        from .engine import stream_ptr
        lib = nat.load()
        sums = self.sums(poses, flags, state)
        n = sums.shape[0]
        out = torch.empty((n, LOCAL), dtype=torch.float64, device=self.pats.device)
        nat.check(lib.zb_local6(sums.data_ptr(), n, out.data_ptr(), stream_ptr()), "local6")
        return out


def moved(poses, delta):
    """Poses [n, 8] moved by delta [n, 6]: t + delta[:3], exp(delta[3:]) q."""
    p = np.array(poses, np.float64, copy=True)
    d = np.asarray(delta, np.float64)
    p[:, :3] += d[:, :3]
    p[:, 4:8] = rot.qstd(rot.qmul(rot.so3_exp(d[:, 3:]), p[:, 4:8]))
    return p


def _qmul_t(a, b):
    w1, x1, y1, z1 = a.unbind(-1)
    w2, x2, y2, z2 = b.unbind(-1)
    return torch.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], -1)


def _exp_t(phi):
    """Unit quaternions exp(phi) [..., 4] of rotation vectors phi [..., 3] (float64)."""
    th = torch.linalg.norm(phi, dim=-1, keepdim=True)
    small = th < 1e-6
    k = torch.where(small, 0.5 - th * th / 48, torch.sin(0.5 * th) / torch.where(small, torch.ones_like(th), th))
    return torch.cat([torch.cos(0.5 * th), k * phi], -1)


def _left_jacobian_t(phi):
    """J [..., 3, 3] with exp(phi + d) = exp(J d) exp(phi) to first order (float64, device)."""
    th = torch.linalg.norm(phi, dim=-1)
    z = torch.zeros_like(th)
    K = torch.stack([torch.stack([z, -phi[..., 2], phi[..., 1]], -1), torch.stack([phi[..., 2], z, -phi[..., 0]], -1),
                     torch.stack([-phi[..., 1], phi[..., 0], z], -1)], -2)
    small = th < 1e-6
    ts = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 0.5 - th * th / 24, (1 - torch.cos(ts)) / (ts * ts))
    b = torch.where(small, 1.0 / 6 - th * th / 120, (ts - torch.sin(ts)) / ts ** 3)
    eye = torch.eye(3, dtype=phi.dtype, device=phi.device).expand(K.shape)
    return eye + a[..., None, None] * K + b[..., None, None] * (K @ K)


def _poses_t(t, index, q):
    """Poses [n, 8] float32 (t, pattern index, q with w >= 0) on the device."""
    q = q / torch.linalg.norm(q, dim=-1, keepdim=True)
    q = torch.where(q[:, :1] < 0, -q, q)
    return torch.cat([t, index[:, None], q], 1).to(torch.float32).contiguous()


class PoseNcc(torch.autograd.Function):
    """NCC [n] (float64) of each pattern at poses moved by delta [n, 6] (shift of the rays'
    origin in pixels, then crystal rotation in radians) from `poses0` [n, 8], differentiable in
    delta. Forward: zb_sums + zb_local6 on the device (the local model is kept as
    field.last_local for a Gauss-Newton driver); backward: the gradient from the same sums, the
    rotation part through the left Jacobian of exp at delta."""

    @staticmethod
    def forward(ctx, delta, field, poses0, flags=0, state=None):
        dev = field.pats.device  # the work is on the field's device, whatever delta's is
        p0 = torch.as_tensor(poses0).to(device=dev, dtype=torch.float64)
        d = delta.detach().to(device=dev, dtype=torch.float64)
        poses = _poses_t(p0[:, :3] + d[:, :3], p0[:, 3], _qmul_t(_exp_t(d[:, 3:]), p0[:, 4:8]))
        local = field.local6(poses, flags, state)
        field.last_local = local
        g = local[:, 2:8]
        J = _left_jacobian_t(d[:, 3:])
        gd = torch.cat([g[:, :3], torch.einsum("nij,ni->nj", J, g[:, 3:])], 1)
        ctx.save_for_backward(gd.to(delta.device))
        return local[:, 0].to(device=delta.device, dtype=delta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (gd,) = ctx.saved_tensors
        return (grad_out[:, None] * gd).to(grad_out.dtype), None, None, None, None


class PoseNccPC(torch.autograd.Function):
    """NCC [n] (float64) of each pattern at crystal rotation exp(phi) q0 and pattern centre pc
    (phi, pc: [n, 3], differentiable): the detector position t(pc) = R (0.5 - pcy h, 0.5 - pcx w,
    pcz h) in pixels (src/geometry.zig:71-75, R the detector's frame in the sample), the sums at
    that pose on the device. Backward: dNCC/dphi = J_l(phi)' g_rot and dNCC/dpc = (dt/dpc)' g_t,
    dt/dpcx = -w R[:, 1], dt/dpcy = -h R[:, 0], dt/dpcz = h R[:, 2]."""

    @staticmethod
    def forward(ctx, phi, pc, field, q0, R, h, w, index):
        dev = field.pats.device
        p = phi.detach().to(device=dev, dtype=torch.float64)
        c = pc.detach().to(device=dev, dtype=torch.float64)
        R = torch.as_tensor(np.asarray(R, np.float64), device=dev)
        v = torch.stack([0.5 - c[:, 1] * h, 0.5 - c[:, 0] * w, c[:, 2] * h], 1)
        t = v @ R.T
        q0 = torch.as_tensor(q0, device=dev).to(torch.float64)
        index = torch.as_tensor(index, device=dev).to(torch.float64)
        local = field.local6(_poses_t(t, index, _qmul_t(_exp_t(p), q0)))
        field.last_local = local
        g = local[:, 2:8]
        g_phi = torch.einsum("nij,ni->nj", _left_jacobian_t(p), g[:, 3:])
        dt = torch.stack([-w * R[:, 1], -h * R[:, 0], h * R[:, 2]], 0)  # [pc component, t component]
        g_pc = g[:, :3] @ dt.T
        ctx.save_for_backward(g_phi.to(phi.device), g_pc.to(phi.device))
        return local[:, 0].to(device=phi.device, dtype=phi.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        g_phi, g_pc = ctx.saved_tensors
        go = grad_out[:, None]
        return (go * g_phi).to(grad_out.dtype), (go * g_pc).to(grad_out.dtype), None, None, None, None, None, None


def local_with_autograd(field, poses, flags=0, state=None):
    """The local model [n, 29] at poses with its gradient taken from a PoseNcc backward pass at
    delta = 0 (the matrix and NCC from the same forward pass); with `state`, of the items still
    moving."""
    z = torch.zeros((poses.shape[0], 6), dtype=torch.float64, device=field.pats.device, requires_grad=True)
    with torch.enable_grad():
        ncc = PoseNcc.apply(z, field, poses, flags, state)
        (g,) = torch.autograd.grad(ncc.sum(), z)
    local = field.last_local.clone()
    local[:, 2:8] = g
    return local


def fit6(field, poses0, iters=20, unknowns=6, lambda0=1e-3, max_rot=np.radians(2.0), max_shift=2.0, tol=1e-7, with_ncc=True, counts=None):
    """Each pose fitted to its pattern by Levenberg-Marquardt steps over the six directions (or
    the rotation alone, unknowns = 3), on the device: per step each item's damped Newton step
    from its local model (zb_fit6_step, FP64: the gradient from a PoseNcc backward pass, the
    Gauss-Newton matrix from the same forward pass), the model at the moved poses (PoseNcc
    again, for the items still moving) deciding which are kept (zb_fit6_keep); an item whose
    step fell below tol has converged and stays put. Returns (poses [n, 8] float32, NCC [n]
    float64 (None without with_ncc), state [n]: 0 still moving after `iters` steps, 1
    converged (the last step below tol), 2 flat), all on the device. `counts` (a dict), if
    given, receives "scored": a device tensor of the poses scored (the start and every trial of
    an item still moving)."""
    from .engine import stream_ptr
    lib = nat.load()
    dev = field.pats.device
    poses = torch.as_tensor(poses0, dtype=torch.float32, device=dev).clone().contiguous()
    n = poses.shape[0]
    flags = ROTATION_ONLY if unknowns == 3 else 0
    local = local_with_autograd(field, poses, flags)
    scored = torch.full((), n, dtype=torch.int64, device=dev)
    lam = torch.full((n,), float(lambda0), dtype=torch.float64, device=dev)
    cur = torch.zeros((n, 2), dtype=torch.float64, device=dev)
    state = torch.zeros((n,), dtype=torch.int32, device=dev)
    trial = torch.empty_like(poses)
    st = stream_ptr()
    for _ in range(iters):
        nat.check(lib.zb_fit6_step(poses.data_ptr(), local.data_ptr(), lam.data_ptr(), n, unknowns, float(max_rot), float(max_shift),
                                   trial.data_ptr(), cur.data_ptr(), state.data_ptr(), st), "fit6 step")
        # only the items still moving are scored (converged ones stay put)
        tlocal = local_with_autograd(field, trial, flags, state)
        scored += (state == 0).sum()
        nat.check(lib.zb_fit6_keep(trial.data_ptr(), tlocal.data_ptr(), n, float(tol), cur.data_ptr(), poses.data_ptr(), local.data_ptr(),
                                   lam.data_ptr(), state.data_ptr(), st), "fit6 keep")
    if counts is not None:
        counts["scored"] = scored
    if not with_ncc:
        return poses, None, state
    return poses, torch.where(local[:, 1] > 0.5, local[:, 0], torch.zeros_like(local[:, 0])), state
