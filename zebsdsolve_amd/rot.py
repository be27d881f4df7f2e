"""Rotations in float64 on the host, vectorised over leading axes.

Conventions (src/rot.zig:1-7): a quaternion is (w, x, y, z), unit, acting on a vector as
q v q*; an orientation q maps sample-frame directions to crystal-frame directions, so crystal
symmetry multiplies on the left (g . q).
"""
import numpy as np


def qmul(a, b):
    """Hamilton product (src/rot.zig:45-52)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    w1, x1, y1, z1 = np.moveaxis(a, -1, 0)
    w2, x2, y2, z2 = np.moveaxis(b, -1, 0)
    return np.stack([
        w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
        w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
        w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
    ], -1)


def qconj(q):
    q = np.asarray(q, np.float64)
    return q * np.array([1.0, -1.0, -1.0, -1.0])


def qstd(q):
    """Unit norm with a non-negative real part (src/rot.zig:60-64)."""
    q = np.asarray(q, np.float64)
    n = np.linalg.norm(q, axis=-1, keepdims=True)
    n = np.where(n == 0, 1.0, n)
    s = np.where(q[..., :1] < 0, -1.0, 1.0)
    return q * s / n


def qapply(q, v):
    """q v q* (src/rot.zig:67-71)."""
    q = np.asarray(q, np.float64)
    v = np.asarray(v, np.float64)
    u = q[..., 1:]
    t = 2.0 * np.cross(u, v)
    return v + q[..., :1] * t + np.cross(u, t)


def to_mat(q):
    """The matrix of v -> q v q* (src/rot.zig:74-85)."""
    w, x, y, z = np.moveaxis(np.asarray(q, np.float64), -1, 0)
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def from_mat(m):
    """Quaternion of a proper rotation matrix (Shepperd's method, src/rot.zig:88-105)."""
    m = np.asarray(m, np.float64)
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        q = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s]
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        q = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s]
    elif m[1, 1] > m[2, 2]:
        s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        q = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s]
    else:
        s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        q = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s]
    return qstd(np.array(q))


def from_zyz(a, b, c):
    """Rz(a) Ry(b) Rz(c), active, as a quaternion (src/rot.zig:176-182)."""
    a, b, c = (np.asarray(x, np.float64) for x in (a, b, c))
    z = np.zeros(np.broadcast(a, b, c).shape)
    qa = np.stack([np.cos(0.5 * a) + z, z, z, np.sin(0.5 * a) + z], -1)
    qb = np.stack([np.cos(0.5 * b) + z, z, np.sin(0.5 * b) + z, z], -1)
    qc = np.stack([np.cos(0.5 * c) + z, z, z, np.sin(0.5 * c) + z], -1)
    return qmul(qa, qmul(qb, qc))


def so3_exp(phi):
    """exp of a rotation vector (src/rot.zig:192-198)."""
    phi = np.asarray(phi, np.float64)
    th2 = np.sum(phi * phi, -1, keepdims=True)
    th = np.sqrt(th2)
    k = np.where(th < 1e-6, 0.5 - th2 / 48.0, np.sin(0.5 * th) / np.where(th < 1e-6, 1.0, th))
    return np.concatenate([np.cos(0.5 * th), k * phi], -1)


def from_bunge(e):
    """Bunge (phi1, Phi, phi2), radians, passive ZXZ -> quaternion (src/rot.zig:131-138)."""
    e = np.asarray(e, np.float64)
    sigma = 0.5 * (e[..., 0] + e[..., 2])
    delta = 0.5 * (e[..., 0] - e[..., 2])
    c = np.cos(0.5 * e[..., 1])
    s = np.sin(0.5 * e[..., 1])
    return qstd(np.stack([c * np.cos(sigma), -s * np.cos(delta), -s * np.sin(delta), -c * np.sin(sigma)], -1))


def to_bunge(q0):
    """Quaternion -> Bunge (phi1, Phi, phi2); phi1, phi2 in [0, 2 pi) (src/rot.zig:141-164)."""
    q = qstd(q0)
    w, x, y, z = np.moveaxis(q, -1, 0)
    q03 = w * w + z * z
    q12 = x * x + y * y
    chi = np.sqrt(q03 * q12)
    safe = np.where(chi < 1e-14, 1.0, chi)
    e0 = np.arctan2((x * z - w * y) / safe, (-w * x - y * z) / safe)
    e1 = np.arctan2(2 * chi, q03 - q12)
    e2 = np.arctan2((w * y + x * z) / safe, (y * z - w * x) / safe)
    deg_lo = q12 <= q03
    e0 = np.where(chi < 1e-14, np.where(deg_lo, np.arctan2(-2 * w * z, w * w - z * z), np.arctan2(2 * x * y, x * x - y * y)), e0)
    e1 = np.where(chi < 1e-14, np.where(deg_lo, 0.0, np.pi), e1)
    e2 = np.where(chi < 1e-14, 0.0, e2)
    e0 = np.where(e0 < 0, e0 + 2 * np.pi, e0)
    e2 = np.where(e2 < 0, e2 + 2 * np.pi, e2)
    return np.stack([e0, e1, e2], -1)
