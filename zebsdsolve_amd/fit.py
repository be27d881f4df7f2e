"""The six-direction local fit of a pattern's pose: the NCC of each pattern with a phase's master
at a pose (detector position t, crystal orientation q), its gradient and Gauss-Newton matrix
over a shift of the rays' origin (pixels, along the sample's x, y, z) and a crystal rotation
(radians, about x, y, z), from the 43 sums of `zb_sums` (detector.wgsl sums; planefit.zig
177-201 localOf, which the reference's pattern-centre and plane fits step with).

`PoseNcc` makes the NCC a torch function of those six directions, with the gradient from the
same sums in its backward pass.
"""
import numpy as np
import torch

from . import _native as nat
from . import rot

N_SUMS = 43
ROTATION_ONLY = 2


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

    def sums(self, poses, flags=0):
        """Sums [n, 43] (device, float32) at poses [n, 8] = (t, pattern index, q) (device or
        host, float32)."""
        from .engine import stream_ptr
        lib = nat.load()
        p = torch.as_tensor(poses, dtype=torch.float32, device=self.pats.device).contiguous()
        n = p.shape[0]
        out = torch.empty((n, N_SUMS), dtype=torch.float32, device=self.pats.device)
        nat.check(lib.zb_sums(self.grad.data_ptr(), self.side, self.frame, self.pats.data_ptr(), p.data_ptr(), n, int(flags), out.data_ptr(),
                              stream_ptr()), "sums")
        return out

    def local(self, poses, flags=0):
        return local_of(self.sums(poses, flags).cpu().numpy())


def moved(poses, delta):
    """Poses [n, 8] moved by delta [n, 6]: t + delta[:3], exp(delta[3:]) q."""
    p = np.array(poses, np.float64, copy=True)
    d = np.asarray(delta, np.float64)
    p[:, :3] += d[:, :3]
    p[:, 4:8] = rot.qstd(rot.qmul(rot.so3_exp(d[:, 3:]), p[:, 4:8]))
    return p


class PoseNcc(torch.autograd.Function):
    """NCC [n] of each pattern at poses moved by delta [n, 6] (shift in pixels, then rotation
    in radians) from `poses0`, differentiable in delta: the backward pass uses the gradient
    from the same sums (the rotation part through the left Jacobian of exp at delta)."""

    @staticmethod
    def forward(ctx, delta, field, poses0):
        d = delta.detach().cpu().double().numpy()
        p = moved(np.asarray(poses0, np.float64), d)
        ok, ncc, g, _ = field.local(p.astype(np.float32))
        J = left_jacobian(d[:, 3:])
        gd = np.concatenate([g[:, :3], np.einsum("nij,ni->nj", J, g[:, 3:])], 1)
        ctx.save_for_backward(torch.from_numpy(gd).to(delta.dtype).to(delta.device))
        return torch.from_numpy(ncc).to(delta.dtype).to(delta.device)

    @staticmethod
    def backward(ctx, grad_out):
        (gd,) = ctx.saved_tensors
        return grad_out[:, None] * gd, None, None


def fit6(field, poses0, iters=20, unknowns=6, lambda0=1e-3, max_rot=np.radians(2.0), max_shift=2.0, tol=1e-7, with_ncc=True):
    """Each pose fitted to its pattern by Levenberg-Marquardt steps over the six directions (or
    the rotation alone, unknowns = 3), on the device: per step the sums at the poses give each
    item's damped Newton step (zb_fit6_propose, FP64), the sums at the moved poses decide which
    are kept (zb_fit6_accept). Returns (poses [n, 8] float32, NCC [n] float64 (None without
    with_ncc), state [n]: 0 still moving after `iters` steps, 1 converged, 2 flat), all on the
    device."""
    from .engine import stream_ptr
    lib = nat.load()
    dev = field.pats.device
    poses = torch.as_tensor(poses0, dtype=torch.float32, device=dev).clone().contiguous()
    n = poses.shape[0]
    flags = ROTATION_ONLY if unknowns == 3 else 0
    sums = field.sums(poses, flags)
    lam = torch.full((n,), float(lambda0), dtype=torch.float64, device=dev)
    cur = torch.zeros((n, 2), dtype=torch.float64, device=dev)
    state = torch.zeros((n,), dtype=torch.int32, device=dev)
    trial = torch.empty_like(poses)
    st = stream_ptr()
    for _ in range(iters):
        nat.check(lib.zb_fit6_propose(poses.data_ptr(), sums.data_ptr(), lam.data_ptr(), n, unknowns, float(max_rot), float(max_shift),
                                      trial.data_ptr(), cur.data_ptr(), state.data_ptr(), st), "fit6 propose")
        tsums = field.sums(trial, flags)
        nat.check(lib.zb_fit6_accept(trial.data_ptr(), tsums.data_ptr(), n, float(tol), cur.data_ptr(), poses.data_ptr(), sums.data_ptr(),
                                     lam.data_ptr(), state.data_ptr(), st), "fit6 accept")
    if not with_ncc:
        return poses, None, state
    ok, ncc, _, _ = local_of(sums.cpu().numpy())
    return poses, torch.from_numpy(np.where(ok, ncc, 0.0)).to(dev), state
