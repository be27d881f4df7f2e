"""Host tables of a bandwidth and of a phase, in float64, as the kernels take them
(src/spherical.zig:106-242).
"""
import numpy as np

from . import master as master_mod
from . import sh

TW_LEN = 1024


def twiddles():
    """tw[t] = exp(-2 pi i t / 1024) as float32 pairs (src/spherical.zig:155-160)."""
    ang = -2.0 * np.pi * np.arange(TW_LEN) / TW_LEN
    return np.stack([np.cos(ang), np.sin(ang)], -1).astype(np.float32)


class BandHost:
    """What depends only on L: the grid (cos theta, weight) per ring, the weighted Legendre table
    of the northern rings, d^l_{k,m}(pi/2), twiddles."""

    def __init__(self, L):
        self.L = L
        self.n = sh.fft_size(L)
        n_theta, n_phi = self.n, 2 * self.n
        self.pairs = sh.pair_count(L)
        grid = sh.Grid.make(L, n_theta, n_phi)
        self.grid = grid
        self.ring = np.stack([grid.x, grid.w], -1).astype(np.float32)
        dphi = 2 * np.pi / n_phi
        # [y][pair] for the northern rings y < n_theta / 2
        self.wplm = (grid.plm[:, : n_theta // 2].T * grid.w[: n_theta // 2, None] * dphi).astype(np.float32)
        self.wigner = sh.wigner_half_pi(L)            # [k, m, l]
        # dq[(m L + l) L + k] = d^l_{k,m}: [m][l][k]
        self.dq = np.ascontiguousarray(np.transpose(self.wigner, (1, 2, 0))).astype(np.float32)
        self.tw = twiddles()


class PhaseHost:
    """What a phase adds at a bandwidth: its harmonics, fold, and the weight tables of the master
    and of its square, [mc][L][L] complex -> float32 pairs (src/spherical.zig:192-242)."""

    def __init__(self, band, side, sheets):
        L = band.L
        self.hm = master_mod.harmonics(side, sheets, L)
        self.fold = sh.fold_of(self.hm.rules)
        self.mu_max = (L - 1) // self.fold
        self.mc = 2 * self.mu_max + 1
        self.nc = band.n // self.fold
        self.l_step = 2 if self.hm.rules.inversion else 1
        wf = master_mod.weights(L, self.hm.f, band.wigner, self.fold)
        wf2 = master_mod.weights(L, self.hm.f2, band.wigner, self.fold)
        self.w_f = np.stack([wf.real, wf.imag], -1).astype(np.float32)
        self.w_f2 = np.stack([wf2.real, wf2.imag], -1).astype(np.float32)


def bitrev(x, bits):
    x = np.asarray(x, np.int64)
    out = np.zeros_like(x)
    for b in range(bits):
        out |= ((x >> b) & 1) << (bits - 1 - b)
    return out


def volume_natural(vol, n, nc, n_items=1):
    """A volume in the slab transform's own order (zb_so3_denominators' rden: per item and
    j, position (ppos, cpos) holding a = 2 bitrev(ppos) + {0, 1} at c = bitrev(cpos)) in the
    natural order [item][j][c][a]."""
    h = n // 2
    nb = h + 1
    lh, lc = int(np.log2(h)), int(np.log2(nc))
    v = np.asarray(vol).reshape(n_items, nb, h, nc, 2)
    out = np.empty((n_items, nb, nc, n), v.dtype)
    p = bitrev(np.arange(h), lh)
    c = bitrev(np.arange(nc), lc)
    out[:, :, c[None, :, None].repeat(h, 2)[0], (2 * p)[None, :].repeat(nc, 0)] = np.transpose(v[..., 0], (0, 1, 3, 2))
    out[:, :, c[None, :, None].repeat(h, 2)[0], (2 * p + 1)[None, :].repeat(nc, 0)] = np.transpose(v[..., 1], (0, 1, 3, 2))
    return out
