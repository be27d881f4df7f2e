"""Oracle: spherical-harmonic and Wigner-d maths in float64 (src/sh.zig), by independent routes.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
import math
from fractions import Fraction

import numpy as np
import scipy.linalg
import scipy.special


def pair(L, m, l):
    return m * L - (m * (m - 1)) // 2 + (l - m)


def pairs(L):
    ms, ls = [], []
    for m in range(L):
        for l in range(m, L):
            ms.append(m)
            ls.append(l)
    return np.array(ms), np.array(ls)


def fft_size(L):
    n = 16
    while n < 2 * L - 1:
        n *= 2
    return n


def gl(n):
    """Gauss-Legendre nodes, descending (theta ascending), and weights: numpy's leggauss
    (src/sh.zig:31-72 computes the same by Newton's method)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return x[::-1].copy(), w[::-1].copy()


_PLM = {}


def plm(L, x):
    """Orthonormal P_l^m(x) with the Condon-Shortley phase for every stored pair: [pair, len(x)],
    as Y_l^m(theta, 0) from scipy (src/sh.zig:75-102 computes them by recurrence). Cached."""
    x = np.atleast_1d(np.asarray(x, np.float64))
    key = (L, x.tobytes())
    if key not in _PLM:
        _PLM[key] = _plm(L, x)
    return _PLM[key]


def _plm(L, x):
    theta = np.arccos(np.clip(x, -1, 1))
    ms, ls = pairs(L)
    out = np.empty((len(ms), len(x)))
    for p, (m, l) in enumerate(zip(ms, ls)):
        out[p] = np.real(scipy.special.sph_harm_y(l, m, theta, 0.0))
    return out


def wigner_exact(l, mp, m):
    """d^l_{m',m}(pi/2) from Wigner's sum in exact rational arithmetic: at beta = pi/2 every term
    carries cos^a sin^b = 2^{-l}, so d = 2^{-l} sqrt((l+m')!(l-m')!(l+m)!(l-m)!) S with S a sum
    of signed reciprocal products of factorials (the textbook formula of src/sh.zig:305-321)."""
    if abs(mp) > l or abs(m) > l:
        return 0.0
    f = math.factorial
    S = Fraction(0)
    for t in range(max(0, m - mp), min(l + m, l - mp) + 1):
        den = f(l + m - t) * f(t) * f(mp - m + t) * f(l - mp - t)
        sgn = -1 if (mp - m + t) % 2 else 1
        S += Fraction(sgn, den)
    A = f(l + mp) * f(l - mp) * f(l + m) * f(l - m)
    d2 = A * S * S / Fraction(4 ** l)
    if S == 0:
        return 0.0
    return math.copysign(math.sqrt(float(d2)), float(S))


def wigner_d_matrix(l, beta):
    """d^l(beta) = exp(-i beta J_y) in the |l, m> basis, rows m' and columns m from -l to l, by
    the eigen-decomposition of J_y."""
    ms = np.arange(-l, l + 1, dtype=np.float64)
    # J+ |l m> = sqrt(l(l+1) - m(m+1)) |l m+1>
    jp = np.zeros((2 * l + 1, 2 * l + 1))
    for i, m in enumerate(ms[:-1]):
        jp[i + 1, i] = math.sqrt(l * (l + 1) - m * (m + 1))
    jy = (jp - jp.T) / 2j
    w, v = np.linalg.eigh(jy)
    d = (v * np.exp(-1j * beta * w)[None, :]) @ v.conj().T
    return d.real


def wigner_table(L, beta):
    """d[l, m' + L - 1, m + L - 1] = d^l_{m',m}(beta) for every l < L (zero outside |m|, |m'| <= l)."""
    out = np.zeros((L, 2 * L - 1, 2 * L - 1))
    for l in range(L):
        d = wigner_d_matrix(l, beta)
        out[l, L - 1 - l:L + l, L - 1 - l:L + l] = d
    return out


def signed(coef, L):
    """Coefficients of every signed order: full[l, m + L - 1], a_{l,-m} = (-1)^m conj(a_lm)."""
    full = np.zeros((L, 2 * L - 1), np.complex128)
    ms, ls = pairs(L)
    for p, (m, l) in enumerate(zip(ms, ls)):
        full[l, L - 1 + m] = coef[p]
        if m > 0:
            full[l, L - 1 - m] = (-1) ** m * np.conj(coef[p])
    return full


def analyze(samples, L, x, w):
    """Analysis on a Gauss-Legendre grid: a_lm = sum_y w_y P_l^m(x_y) (2 pi / n_phi) sum_j f
    e^{-i m phi_j} (src/sh.zig:154-194)."""
    n_theta, n_phi = samples.shape
    F = np.fft.fft(samples, axis=1)[:, :L] * (2 * np.pi / n_phi)
    P = plm(L, x) * w[None, :]
    ms, _ = pairs(L)
    return np.einsum("py,yp->p", P, F[:, ms])


def synthesize(coef, L, x, n_phi):
    """Samples from coefficients: G_m(y) = sum_l a_lm P_l^m, v = G_0 + 2 Re sum_m G_m e^{i m phi}."""
    P = plm(L, x)
    ms, _ = pairs(L)
    G = np.zeros((len(x), L), np.complex128)
    np.add.at(G.T, ms, P * coef[:, None])
    phi = 2 * np.pi * np.arange(n_phi) / n_phi
    E = np.exp(1j * np.outer(np.arange(L), phi))
    k = np.where(np.arange(L) == 0, 1.0, 2.0)
    return np.real((G * k[None, :]) @ E)


def correlation_at(f, g, L, a, b, c):
    """C(a, b, c) = int g(x) f(S^-1 x) dOmega, S = Rz(a) Ry(b) Rz(c), straight from the
    coefficients: sum_{l,m,m'} f_lm conj(g_lm') D^l_{m'm}(S) with D = e^{-i m' a} d^l_{m'm}(b)
    e^{-i m c} (src/sh.zig:350-379). For tests at small L."""
    F = signed(f, L)
    G = signed(g, L)
    total = 0.0
    for l in range(L):
        d = wigner_d_matrix(l, b)
        ms = np.arange(-l, l + 1)
        fl = F[l, L - 1 - l:L + l]
        gl_ = np.conj(G[l, L - 1 - l:L + l])
        ph = np.exp(-1j * (ms[:, None] * a + ms[None, :] * c))
        total += np.real(np.sum(gl_[:, None] * fl[None, :] * d * ph))
    return total
