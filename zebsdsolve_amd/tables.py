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


def d_beta(L, fold, wigner, betas):
    """D[j, l, m', mu] = d^l_{m', fold mu}(beta_j) for m' in [0, L), |mu| <= (L-1)/fold, from the
    table at pi/2 by d^l_{m'm}(beta) = i^{m'-m} sum_k d^l_{km'}(pi/2) d^l_{km}(pi/2) e^{-ik beta}
    (src/sh.zig:508-518), with d^l_{-k,m} = (-1)^{l+m} d^l_{k,m} folding k < 0 onto k > 0."""
    mu_max = (L - 1) // fold
    mus = np.arange(-mu_max, mu_max + 1)
    ms = fold * mus
    am = np.abs(ms)
    betas = np.asarray(betas, np.float64)
    ks = np.arange(L)
    out = np.zeros((len(betas), L, L, len(mus)))
    cosk = np.cos(np.outer(betas, ks))           # [j, k]
    sink = np.sin(np.outer(betas, ks))
    for l in range(L):
        # P[k, m'] = d^l_{k, m'} (k, m' >= 0); R[k, mu] = d^l_{k, m} for signed m
        P = wigner[:l + 1, :, l]                                  # [k, m']
        R = wigner[:l + 1, am, l].copy()                          # [k, mu], |m|
        neg = ms < 0
        # d^l_{k,-|m|} = (-1)^{l+k} d^l_{k,|m|}
        R[:, neg] *= np.where((l + np.arange(l + 1)) % 2 == 1, -1.0, 1.0)[:, None]
        wk = np.where(np.arange(l + 1) == 0, 1.0, 2.0)
        # parity of m' + m decides whether the +-k terms add to a cosine or a sine
        Sc = np.einsum("km,jk,ku->jmu", P, cosk[:, :l + 1] * wk, R, optimize=True)
        Ss = np.einsum("km,jk,ku->jmu", P, sink[:, :l + 1] * wk, R, optimize=True)
        mp = np.arange(L)[:, None]
        par = (mp + ms[None, :]) % 2
        # i^{m'-m} (sum_k ... e^{-ik beta}): even parity -> i^{m'-m} Sc (real), odd -> i^{m'-m} (-i Ss)
        ph = (mp - ms[None, :]) % 4
        even = np.where(ph == 0, 1.0, np.where(ph == 2, -1.0, 0.0))
        # i^p (-i) = i^{p-1}: p = 1 gives 1, p = 3 gives -1 (p even has even parity)
        odd = np.where(ph == 1, 1.0, np.where(ph == 3, -1.0, 0.0))
        # (P's columns m' > l and R's columns |m| > l are zero in the table: so are these)
        out[:, l] = np.where(par[None] == 0, Sc * even[None], Ss * odd[None])
    return out


def vgemm_shape(nc, slots):
    """(cols, S, nblk) of the matrix-core correlation for nc points along c and `slots` degree
    slots, by the rule its kernel launch follows (zb_vtable_shape in the C ABI): cols = 32
    ceil(2 nc / 32) table columns, S = the first of 8, 16, 32, 48, 64, 96 >= slots slots per
    wave, nblk column blocks of 32 per wave (1 when cols = 32 or S >= 64, else 2)."""
    cols = 32 * ((2 * nc + 31) // 32)
    S = next(x for x in (8, 16, 32, 48, 64, 96) if x >= slots)
    nblk = 1 if (cols == 32 or S >= 64) else 2
    return cols, S, nblk


def tile_vtable(T, S, nblk):
    """The natural table [nb][L][slots][2][cols] in the layout the kernel loads 16 bytes per lane
    from: per (j, m', column group cg of nblk x 32 columns), [S/4][nblk][64 lanes][4], lane =
    32 r + x holding row r (re, im) of column cg nblk 32 + b 32 + x for the four slots
    4 sq .. 4 sq + 3 (zero past the last slot)."""
    nb, L, slots, _, cols = T.shape
    ncg = cols // (32 * nblk)
    Tp = np.zeros((nb, L, S, 2, cols), np.float32)
    Tp[:, :, :slots] = T
    Tp = Tp.reshape(nb, L, S // 4, 4, 2, ncg, nblk, 32)
    return np.ascontiguousarray(np.transpose(Tp, (0, 1, 5, 2, 6, 4, 7, 3))).reshape(-1)


def vgemm16_shape(nc, slots):
    """(cols, groups, nblk) of the 16-bit table (zb_vtable16_shape in the C ABI): groups of 8
    slots, an even count, 2 ceil(slots / 16); nblk 1 when cols = 32, else 2."""
    cols = 32 * ((2 * nc + 31) // 32)
    return cols, 2 * ((slots + 15) // 16), 1 if cols == 32 else 2


def round16(x, kind):
    """float64 values to FP16 (kind 1) or BF16 (kind 2), nearest with ties to even: (the 16-bit
    patterns as uint16, their values as float64). BF16 rounds the float32 nearest to x."""
    if kind == 1:
        if np.abs(x).max(initial=0.0) > 65504.0:
            raise ValueError("the table exceeds the FP16 range: use BF16")
        h = x.astype(np.float16)
        return h.view(np.uint16), h.astype(np.float64)
    u = x.astype(np.float32).view(np.uint32)
    r = ((u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) >> np.uint32(16)).astype(np.uint16)
    return r, (r.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64)


def split_slot(split_degree, l_step):
    """First slot (a multiple of 16) of the 16-bit products with hi + lo: the group of 16 slots
    holding degree split_degree, so every degree >= split_degree takes them."""
    return ((split_degree + l_step - 1) // l_step) // 16 * 16


def tile_vtable16(T, groups, nblk, split, kind):
    """The natural table of one j, [L][slots][2][cols] float64, as zb_vtable.qlp's pieces for
    that j: per (m', column group cg of 32 complex columns), the groups t of 8 slots in order,
    each [2 blocks][64 lanes][8] hi values, lane = 32 r + x holding row r of column
    2 (32 cg + x) + b (block 0 the real and block 1 the imaginary part of complex column
    32 cg + x) for slots 8 t .. 8 t + 7; groups t >= split / 8 followed by the same for
    lo = round16(x - hi)."""
    L, slots, _, cols = T.shape
    if nblk != 2 or cols % 64:
        raise ValueError("the 16-bit table pairs the real and imaginary column blocks: nc >= 32")
    ncg = cols // 64
    Tp = np.zeros((L, 8 * groups, 2, cols))
    Tp[:, :slots] = T
    Tp = np.transpose(Tp.reshape(L, groups, 8, 2, ncg, 32, 2), (0, 4, 1, 6, 3, 5, 2))  # [m'][cg][t][b][r][x][i]
    hb, hv = round16(Tp, kind)
    lb, _ = round16(Tp - hv, kind)
    P = nblk * 64 * 8
    H = hb.reshape(L, ncg, groups, P)
    Lo = lb.reshape(L, ncg, groups, P)
    gs = min(split // 8, groups)
    both = np.stack((H[:, :, gs:], Lo[:, :, gs:]), axis=3).reshape(L, ncg, 2 * (groups - gs), P)
    return np.concatenate((H[:, :, :gs], both), axis=2).reshape(-1)


class VTableHost:
    """The correlation as one GEMM per (j, m'): V[j, m', c] = sum_l Q[j, m', c, l] conj(g_{l,m'}),
    with Q[j, m', c, l] = sum_mu f_{l, fold mu} d^l_{m', fold mu}(beta_j) e^{-2 pi i mu c / nc}
    (the transforms along b and c of the reference's spectrum folded into a pattern-independent
    table). Laid out as the MFMA B operand: per (j, m', slot s) two rows (k = re, im) of
    `cols` columns, c-major with (re, im) interleaved, then zero pad:
        row re: column 2c = Re Q_c, 2c + 1 = Im Q_c;   row im: 2c = Im Q_c, 2c + 1 = -Re Q_c
    so that Re g (row re) + Im g (row im) gives V = Q conj(g) as (re, im) pairs. Slot s holds
    l = s l_step."""

    def __init__(self, band, phase, coef):
        L, n = band.L, band.n
        fold, nc = phase.fold, phase.nc
        self.L, self.n, self.nc, self.fold, self.l_step = L, n, nc, fold, phase.l_step
        self.slots = (L + self.l_step - 1) // self.l_step
        self.cols = 32 * ((2 * nc + 31) // 32)
        nb = n // 2 + 1
        betas = 2 * np.pi * np.arange(nb) / n
        D = d_beta(L, fold, band.wigner, betas)                   # [j, l, m', mu]
        mu_max = (L - 1) // fold
        mus = np.arange(-mu_max, mu_max + 1)
        F = np.zeros((L, len(mus)), np.complex128)
        for i, mu in enumerate(mus):
            for l in range(abs(fold * mu), L):
                F[l, i] = sh.signed_coef(coef, L, l, fold * mu)
        E = np.exp(-2j * np.pi * np.outer(mus, np.arange(nc)) / nc)   # [mu, c]
        ls = np.arange(self.slots) * self.l_step
        Q = np.einsum("su,jsmu,uc->jmcs", F[ls], D[:, ls], E, optimize=True)      # [j, m', c, s]
        # zero where l < m'
        Q *= (ls[None, None, None, :] >= np.arange(L)[None, :, None, None])
        self.Q = Q
        T = np.zeros((nb, L, self.slots, 2, self.cols), np.float32)
        Qt = np.transpose(Q, (0, 1, 3, 2))                          # [j, m', s, c]
        T[..., 0, 0:2 * nc:2] = Qt.real
        T[..., 0, 1:2 * nc:2] = Qt.imag
        T[..., 1, 0:2 * nc:2] = Qt.imag
        T[..., 1, 1:2 * nc:2] = -Qt.real
        self.table = T
        _, self.S, self.nblk = vgemm_shape(nc, self.slots)
        self.tiled = tile_vtable(T, self.S, self.nblk)

    def table16(self, kind, split):
        """zb_vtable.qlp for FP16 (kind 1) or BF16 (kind 2) with hi + lo products from slot
        `split` on: the rows rebuilt in float64 from Q, one j at a time."""
        _, groups, nblk = vgemm16_shape(self.nc, self.slots)
        nc = self.nc
        # V = Q conj(g) with unit-norm coefficients: |V| <= the l2 norm of a row of Q, which FP16
        # V must hold (twice over)
        if kind == 1 and 2.0 * np.sqrt((np.abs(self.Q) ** 2).sum(axis=-1)).max() > 65504.0:
            raise ValueError("V could exceed the FP16 range: use BF16")
        out = []
        for Qj in self.Q:                                            # [m', c, s]
            Qt = np.transpose(Qj, (0, 2, 1))                          # [m', s, c]
            T = np.zeros((self.L, self.slots, 2, self.cols))
            T[:, :, 0, 0:2 * nc:2] = Qt.real
            T[:, :, 0, 1:2 * nc:2] = Qt.imag
            T[:, :, 1, 0:2 * nc:2] = Qt.imag
            T[:, :, 1, 1:2 * nc:2] = -Qt.real
            out.append(tile_vtable16(T, groups, nblk, split, kind))
        return np.concatenate(out)
