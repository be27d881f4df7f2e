"""Spherical harmonics on the host, in float64: Gauss-Legendre grids, normalised associated
Legendre functions, analysis and synthesis, Wigner d at pi/2 (src/sh.zig:1-302).

Conventions (src/sh.zig:6-13): Y_lm = P_l^m(cos theta) e^{i m phi}, orthonormal, Condon-Shortley
phase; coefficients a_lm = int f conj(Y_lm) are stored for m >= 0 at pair(L, m, l) (m-major,
l = m..L-1); a grid has n_theta Gauss-Legendre rings (theta ascending, nodes descending) of
n_phi points phi_j = 2 pi j / n_phi.
"""
import math
from dataclasses import dataclass

import numpy as np


def pair_count(L):
    return L * (L + 1) // 2


def pair(L, m, l):
    """Index of (m, l), 0 <= m <= l < L (src/sh.zig:27-29)."""
    return m * L - (m * (m - 1)) // 2 + (l - m)


def pairs_of(L):
    """(m, l) of every stored pair, in storage order."""
    ms, ls = [], []
    for m in range(L):
        for l in range(m, L):
            ms.append(m)
            ls.append(l)
    return np.array(ms), np.array(ls)


def fft_size(L):
    """Smallest power of two >= max(16, 2L - 1) (src/spherical.zig:100-104)."""
    n = 16
    while n < 2 * L - 1:
        n *= 2
    return n


def gauss_legendre(n):
    """Nodes (descending) and weights on [-1, 1] by Newton's method on the Legendre recurrence
    (src/sh.zig:32-72), vectorised over the nodes."""
    i = np.arange((n + 1) // 2, dtype=np.float64)
    z = np.cos(np.pi * (i + 0.75) / (n + 0.5))

    def p_and_d(z):
        p0 = np.ones_like(z)
        p1 = z.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
        return p1, n * (z * p1 - p0) / (z * z - 1)

    for _ in range(100):
        p, dp = p_and_d(z)
        dz = p / dp
        z = z - dz
        if np.all(np.abs(dz) < 1e-16):
            break
    _, dp = p_and_d(z)
    x = np.empty(n)
    w = np.empty(n)
    h = len(z)
    x[:h] = z
    x[n - h:] = -z[::-1]
    wh = 2 / ((1 - z * z) * dp * dp)
    w[:h] = wh
    w[n - h:] = wh[::-1]
    return x, w


def legendre_all(L, x):
    """P_l^m(x) for every pair below L: out[pair, k] for x[k] (src/sh.zig:75-102)."""
    x = np.atleast_1d(np.asarray(x, np.float64))
    s = np.sqrt(np.maximum(0.0, 1 - x * x))
    out = np.empty((pair_count(L), x.size))
    pmm = np.full(x.size, np.sqrt(1.0 / (4.0 * np.pi)))
    for m in range(L):
        if m > 0:
            pmm = pmm * (-np.sqrt((2 * m + 1) / (2 * m)) * s)
        base = pair(L, m, m)
        out[base] = pmm
        if m + 1 < L:
            out[base + 1] = np.sqrt(2 * m + 3) * x * pmm
        for l in range(m + 2, L):
            a = np.sqrt((4 * l * l - 1) / (l * l - m * m))
            b = np.sqrt(((l - 1) * (l - 1) - m * m) / (4 * (l - 1) * (l - 1) - 1))
            out[base + l - m] = a * (x * out[base + l - m - 1] - b * out[base + l - m - 2])
    return out


def wigner_half_pi(L):
    """d^l_{k,m}(pi/2) for 0 <= k, m, l < L as d[k, m, l] (zero where l < max(k, m)), from the
    Jacobi-polynomial form at 0 with log factorials (src/sh.zig:239-283)."""
    n = L
    d = np.zeros((n, n, n))
    lf = [0.0]
    for i in range(1, 2 * n + 2):
        lf.append(lf[-1] + math.log(i))
    for k in range(n):
        for m in range(k + 1):
            a = float(k - m)
            b = float(k + m)
            p2 = 0.0
            p1 = 0.0
            for nn in range(n - k):
                if nn == 0:
                    p = 1.0
                elif nn == 1:
                    p = (a - b) / 2
                else:
                    c = 2 * nn + a + b
                    p = ((c - 1) * (a * a - b * b) * p1 - 2 * (nn + a - 1) * (nn + b - 1) * c * p2) / (2 * nn * (nn + a + b) * (c - 2))
                j = nn + k
                pref = math.exp(0.5 * (lf[nn + 2 * k] + lf[nn] - lf[j + m] - lf[j - m])) * 2.0 ** (-k)
                v = pref * p
                d[m, k, j] = v
                d[k, m, j] = v if (k - m) % 2 == 0 else -v
                p2, p1 = p1, p
    return d


def d_sign(l, k, m):
    """Sign of d^l_{k,m}(pi/2) relative to d^l_{|k|,|m|}(pi/2) (src/sh.zig:286-296)."""
    if k < 0 and m < 0:
        return -1.0 if (k - m) % 2 else 1.0
    if m < 0:
        return -1.0 if (l + k) % 2 else 1.0
    if k < 0:
        return -1.0 if (l + m) % 2 else 1.0
    return 1.0


def d_half_pi(tab, l, k, m):
    ak, am = abs(k), abs(m)
    if ak > l or am > l:
        return 0.0
    return d_sign(l, k, m) * tab[ak, am, l]


def signed_coef(coef, L, l, m):
    """Coefficient of a signed order from m >= 0 storage: a_{l,-m} = (-1)^m conj(a_lm)."""
    am = abs(m)
    if am > l:
        return 0j
    c = coef[pair(L, am, l)]
    if m >= 0:
        return c
    return np.conj(c) if am % 2 == 0 else -np.conj(c)


@dataclass
class Grid:
    """A Gauss-Legendre grid with the Legendre table of bandlimit L (src/sh.zig:105-143)."""
    L: int
    n_theta: int
    n_phi: int
    x: np.ndarray
    w: np.ndarray
    plm: np.ndarray  # [pair, ring]

    @staticmethod
    def make(L, n_theta, n_phi):
        x, w = gauss_legendre(n_theta)
        return Grid(L, n_theta, n_phi, x, w, legendre_all(L, x))

    def directions(self):
        """Unit vectors of every grid point, [n_theta, n_phi, 3]."""
        ct = self.x[:, None]
        st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
        phi = 2 * np.pi * np.arange(self.n_phi) / self.n_phi
        return np.stack(np.broadcast_arrays(st * np.cos(phi), st * np.sin(phi), ct), -1)

    def analyze(self, samples):
        """samples[n_theta, n_phi] -> coefficients[pairs] (src/sh.zig:154-194): ring spectra
        F[y, m] = (2 pi / n_phi) sum_j f e^{-i m phi_j}, then the weighted Legendre sums."""
        L = self.L
        F = np.fft.fft(np.asarray(samples, np.float64).reshape(self.n_theta, self.n_phi), axis=1)[:, :L] * (2 * np.pi / self.n_phi)
        ms, _ = pairs_of(L)
        return np.einsum("py,py->p", self.plm * self.w[None, :], F[:, ms].T)

    def synthesize(self, coef):
        """coefficients -> samples[n_theta, n_phi] (src/sh.zig:197-230)."""
        L = self.L
        ms, _ = pairs_of(L)
        G = np.zeros((self.n_theta, L), np.complex128)
        for m in range(L):
            sel = ms == m
            G[:, m] = (self.plm[sel] * coef[sel][:, None]).sum(0)
        full = np.zeros((self.n_theta, self.n_phi), np.complex128)
        full[:, 0] = G[:, 0]
        full[:, 1:L] = 2 * G[:, 1:L]
        # v = G0 + 2 Re sum_m G_m e^{i m phi}
        return np.real(np.fft.ifft(full, axis=1) * self.n_phi)


def power(L, coef):
    """sum |a_lm|^2 over all signed orders."""
    ms, _ = pairs_of(L)
    k = np.where(ms == 0, 1.0, 2.0)
    return float(np.sum(k * np.abs(coef) ** 2))


@dataclass
class Rules:
    """Which coefficients a crystal's symmetry leaves non-zero (src/sh.zig:381-423)."""
    n_fold: int = 1
    mirror: bool = False
    inversion: bool = False

    def allows(self, l, m):
        l = np.asarray(l)
        m = np.asarray(m)
        ok = m % self.n_fold == 0
        if self.mirror:
            ok = ok & ((l + m) % 2 == 0)
        if self.inversion:
            ok = ok & (l % 2 == 0)
        return ok

    @staticmethod
    def detect(L, coef, tol):
        ms, ls = pairs_of(L)
        keep = ls != 0
        e = np.where(ms == 0, 1.0, 2.0) * np.abs(coef) ** 2
        e = np.where(keep, e, 0.0)
        total = e.sum()
        if total <= 0:
            return Rules()
        r = Rules(inversion=bool(e[ls % 2 != 0].sum() < tol * total), mirror=bool(e[(ls + ms) % 2 != 0].sum() < tol * total))
        for n in (6, 4, 3, 2):
            if e[ms % n != 0].sum() < tol * total:
                r.n_fold = n
                break
        return r


def fold_of(rules):
    """Largest power of two dividing the master's rotational order about z (src/master.zig:715-721)."""
    return {4: 4, 2: 2, 6: 2}.get(rules.n_fold, 1)
