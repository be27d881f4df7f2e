"""The equal-area (Rosca-Lambert) map between a square and a hemisphere that master patterns are
stored on (src/lambert.zig:1-64): a side x side image per hemisphere, pixel (row, col) at square
coordinates (x, y) = (2 col / (side - 1) - 1, 2 row / (side - 1) - 1). Float64, vectorised.
"""
import numpy as np

_SC = np.sqrt(np.pi / 2.0)


def square_to_sphere(x, y, southern):
    """Square coordinates in [-1, 1]^2 -> unit vector on the northern (or southern) hemisphere."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    southern = np.asarray(southern, bool)
    xa = np.abs(x) * _SC
    ya = np.abs(y) * _SC
    swap = xa >= ya
    xn = np.where(swap, xa, ya)
    yn = np.where(swap, ya, xa)
    safe = np.where(xn == 0, 1.0, xn)
    r = (2 * safe / np.pi) * np.sqrt(np.maximum(np.pi - safe * safe, 0.0))
    ang = np.pi * yn / (4 * safe)
    xh = r * np.cos(ang)
    yh = r * np.sin(ang)
    z = 1 - 2 * safe * safe / np.pi
    hx = np.where(swap, xh, yh) * np.where(x < 0, -1.0, 1.0)
    hy = np.where(swap, yh, xh) * np.where(y < 0, -1.0, 1.0)
    v = np.stack([hx, hy, np.where(southern, -z, z)], -1)
    v = v / np.linalg.norm(v, axis=-1, keepdims=True)
    pole = np.stack(np.broadcast_arrays(0.0, 0.0, np.where(southern, -1.0, 1.0)), -1)
    return np.where((xn == 0)[..., None], pole, v)


def sphere_to_square(v):
    """Unit vector -> square coordinates of its own hemisphere (|z|)."""
    v = np.asarray(v, np.float64)
    kz = np.abs(v[..., 2])
    xn = np.sqrt(np.maximum(np.pi * (1 - kz) / 2, 0.0))
    ax = np.abs(v[..., 0])
    ay = np.abs(v[..., 1])
    big = ax >= ay
    u = np.where(big, xn, (4 * xn / np.pi) * np.arctan2(ax, np.maximum(ay, 1e-300)))
    w = np.where(big, (4 * xn / np.pi) * np.arctan2(ay, np.maximum(ax, 1e-300)), xn)
    u = (u / _SC) * np.where(v[..., 0] < 0, -1.0, 1.0)
    w = (w / _SC) * np.where(v[..., 1] < 0, -1.0, 1.0)
    pole = kz >= 1 - 1e-14
    return np.where(pole, 0.0, u), np.where(pole, 0.0, w)


def sample(sheet, side, x, y):
    """Bilinear sample of a side x side sheet at square coordinates (x, y)."""
    last = side - 1.0
    fj = np.clip((np.asarray(x) + 1) * 0.5 * last, 0.0, last)
    fi = np.clip((np.asarray(y) + 1) * 0.5 * last, 0.0, last)
    c0 = np.floor(fj).astype(np.int64)
    r0 = np.floor(fi).astype(np.int64)
    c1 = np.minimum(c0 + 1, side - 1)
    r1 = np.minimum(r0 + 1, side - 1)
    tj = fj - c0
    ti = fi - r0
    s = np.asarray(sheet, np.float64).reshape(-1)
    a = s[r0 * side + c0]
    b = s[r0 * side + c1]
    c = s[r1 * side + c0]
    d = s[r1 * side + c1]
    return (a * (1 - tj) + b * tj) * (1 - ti) + (c * (1 - tj) + d * tj) * ti


def at(sheets, side, v):
    """Intensity of [north | south] sheets in direction v (src/master.zig:152-156)."""
    v = np.asarray(v, np.float64)
    sheets = np.asarray(sheets).reshape(2, side * side)
    u, w = sphere_to_square(v)
    north = sample(sheets[0], side, u, w)
    south = sample(sheets[1], side, u, w)
    return np.where(v[..., 2] >= 0, north, south)
