// Normalized cross-correlation over SO(3) from spherical-harmonic coefficients (the reference's
// so3.wgsl stage, src/spherical.zig:363-464), for gfx950.
//
// For a pattern g (sample frame) and a master f (crystal frame), C(a, b, c) = int g(x) f(S^-1 x)
// over S = Rz(a) Ry(b) Rz(c) is the Fourier series C = sum T[k, m', m] e^{-i (k b + m' a + m c)},
//     T = i^{m'-m} sum_l f_lm conj(g_lm') d^l_{k m'}(pi/2) d^l_{k m}(pi/2)
// (sh.zig:6-13, 352-379). T[-k] = (-1)^{m'+m} T[k], so b is transformed first from k >= 0 and
// only j <= n/2 kept; C is real, so only m' >= 0 is computed and a is transformed last as a real
// transform. A master with a z axis of order `fold` has only orders m = fold mu.
//
//   k_spectrum  one wave per (item, pair of orders m' = 2h, 2h+1, group of Q columns iq): lane
//               k sums over l (the tables are zero below their own l, so every lane starts at
//               the same l and the coefficients are wave-uniform loads), the two orders' columns
//               go into one complex line each (one is even in k, the other odd), an in-LDS FFT
//               along k, the halves apart into U[b][j][m'][iq], j <= n/2
//   k_slab      one block per (item, j) slab of U: FFT along c (rows m'), the real transform's
//               half-length input per column, FFT along a (columns), both in place in LDS in
//               decimation-in-frequency order (bit-reversed out: no permutation pass); then
//               either the argmax of value x rden (ties: lowest index) into one 64-bit atomic
//               per item, or the volume itself (for the denominators)
//   k_finish    per item: the peak's six neighbours summed directly from U, parabolic offsets
//               (each within half a step), and the orientation conj(Rz(a) Ry(b) Rz(c))
//
// Volumes (rden, and the window's) are stored in the slab transform's own order: for slab j,
// position (ppos, cpos) holds a = 2 bitrev(ppos) + {0, 1} (float2) at c = bitrev(cpos).
#include "common.h"

namespace zb {

struct SoArg {
    int L, n, log2n, nc, log2nc, fold, mu_max, l_step, pairs, nb, Lh, Q;
};

__device__ __forceinline__ int mu_of(int iq, int nc) { return iq >= (nc + 1) / 2 ? iq - nc : iq; }

template <int KPT, int Q>
__global__ void __launch_bounds__(64) k_spectrum(SoArg a, int n_items, const float2* __restrict__ wm, const float* __restrict__ dq,
                                                 const float2* __restrict__ coef, const float2* __restrict__ tw, float2* __restrict__ U) {
    extern __shared__ float2 lines[];  // [Q][n]
    const int ngroups = a.nc / Q;
    int blk = blockIdx.x;
    const int cg = blk % ngroups;
    blk /= ngroups;
    const int hh = blk % a.Lh;
    const int b = blk / a.Lh;
    const int L = a.L, n = a.n;
    const int ma = 2 * hh, mb = 2 * hh + 1;
    const bool has_b = mb < L;
    const int iq0 = cg * Q;
    const int lane = threadIdx.x;

    int mi[Q];
    int amu[Q];
    bool live[Q];
    int lo = L;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int mu = mu_of(iq0 + q, a.nc);
        amu[q] = mu < 0 ? -mu : mu;
        live[q] = amu[q] <= a.mu_max;
        mi[q] = min(max(mu, -a.mu_max), a.mu_max) + a.mu_max;
        if (live[q]) lo = min(lo, amu[q] * a.fold);
    }
    int l0 = max(ma, lo);
    if (a.l_step == 2) l0 += l0 & 1;

    float2 accA[KPT][Q], accB[KPT][Q];
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            accA[kk][q] = make_float2(0.f, 0.f);
            accB[kk][q] = make_float2(0.f, 0.f);
        }
    const float2* g = coef + (size_t)b * a.pairs;
    const int pa = ma * L - (ma * (ma - 1)) / 2 - ma;  // pair(ma, l) = pa + l
    const int pb = mb * L - (mb * (mb - 1)) / 2 - mb;
    for (int l = l0; l < L; l += a.l_step) {
        const float2 ga = g[pa + l];
        const float2 gb = (has_b && l >= mb) ? g[pb + l] : make_float2(0.f, 0.f);
#pragma unroll
        for (int kk = 0; kk < KPT; ++kk) {
            const int k = lane + 64 * kk;
            if (k < L) {
                const float da = dq[((size_t)ma * L + l) * L + k];
                const float db = has_b ? dq[((size_t)mb * L + l) * L + k] : 0.f;
                // conj(g) d
                const float2 xa = make_float2(ga.x * da, -ga.y * da);
                const float2 xb = make_float2(gb.x * db, -gb.y * db);
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const float2 w = wm[((size_t)mi[q] * L + l) * L + k];
                    accA[kk][q].x += w.x * xa.x - w.y * xa.y;
                    accA[kk][q].y += w.x * xa.y + w.y * xa.x;
                    accB[kk][q].x += w.x * xb.x - w.y * xb.y;
                    accB[kk][q].y += w.x * xb.y + w.y * xb.x;
                }
            }
        }
    }
    // k runs over -L < k < L in a line of n >= 2L - 1 points: zero between
    for (int e = lane; e < Q * n; e += 64) {
        const int p = e & (n - 1);
        if (p >= L && p <= n - L) lines[e] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) {
        const int k = lane + 64 * kk;
        if (k < L) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                float2 xa = live[q] ? ipow(accA[kk][q], ma) : make_float2(0.f, 0.f);
                float2 xb = (live[q] && has_b) ? ipow(accB[kk][q], mb) : make_float2(0.f, 0.f);
                // T[-k] = (-1)^{m'+m} T[k]: odd for the line m' = ma when ma + m is odd, else for mb
                const bool odd_a = ((ma + amu[q] * a.fold) & 1) != 0;
                lines[q * n + k] = cadd(xa, xb);
                if (k > 0) {
                    const float2 na = odd_a ? make_float2(-xa.x, -xa.y) : xa;
                    const float2 nb_ = odd_a ? xb : make_float2(-xb.x, -xb.y);
                    lines[q * n + n - k] = cadd(na, nb_);
                }
            }
        }
    }
    __syncthreads();
    fft_dif<64>(lines, a.log2n, Q, n, 1, tw, lane);
    // the pair apart: the even part of the transform is the even line's, the odd part the odd one's
    for (int e = lane; e < a.nb * Q; e += 64) {
        const int q = e % Q;
        const int j = e / Q;
        const float2 zj = lines[q * n + bitrev(j, a.log2n)];
        const float2 zr = lines[q * n + bitrev((n - j) & (n - 1), a.log2n)];
        const float2 ev = make_float2(0.5f * (zj.x + zr.x), 0.5f * (zj.y + zr.y));
        const float2 od = make_float2(0.5f * (zj.x - zr.x), 0.5f * (zj.y - zr.y));
        // which of the two is odd: ma even, so m alone says (m = fold mu)
        const int mu = mu_of(iq0 + q, a.nc);
        const bool a_odd = (((mu < 0 ? -mu : mu) * a.fold) & 1) != 0;
        const size_t at = (((size_t)b * a.nb + j) * L + ma) * a.nc + iq0 + q;
        U[at] = a_odd ? od : ev;
        if (has_b) U[at + a.nc] = a_odd ? ev : od;
    }
}

__device__ __forceinline__ unsigned ordered_bits(float v) {
    unsigned u = __float_as_uint(v + 0.f);  // (-0 -> +0)
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float from_ordered(unsigned o) {
    unsigned u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

// MODE 0: argmax into best[b]; MODE 1: write the volume (position order) into vol.
template <int MODE>
__global__ void __launch_bounds__(256) k_slab(SoArg a, int n_items, const float2* __restrict__ U, const float2* __restrict__ tw,
                                              const float2* __restrict__ rden, int rden_div, unsigned long long* __restrict__ best,
                                              float2* __restrict__ vol) {
    extern __shared__ float2 fm[];  // [h][nc]
    __shared__ float sv[4];
    __shared__ unsigned si[4];
    const int slab = blockIdx.x;
    const int b = slab / a.nb;
    const int j = slab - b * a.nb;
    const int L = a.L, nc = a.nc, log2nc = a.log2nc;
    const int h = a.n >> 1;
    const int log2h = a.log2n - 1;
    const int tid = threadIdx.x;
    const float2* src = U + ((size_t)b * a.nb + j) * L * nc;
    for (int e = tid; e < h * nc; e += 256) {
        const int row = e >> log2nc;
        fm[e] = row < L ? src[e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    // along c: V[m', cpos]
    fft_dif<256>(fm, log2nc, L, nc, 1, tw, tid);
    // the real transform along a from a complex one of half its length:
    //   Z_q = (X_q + X_{q+h}) + i e^{-2 pi i q / n} (X_q - X_{q+h}),  X_{q+h} = conj X_{h-q}
    // (X_m' = V[m'] for m' < L <= h, zero above; X_0 real). A thread owns the pair (q, h - q).
    const int tws = ZB_TW / a.n;
    for (int e = tid; e < nc * (h / 2 + 1); e += 256) {
        const int cpos = e & (nc - 1);
        const int q = e >> log2nc;
        const int r = h - q;
        float2 xq = fm[q * nc + cpos];
        float2 yq = make_float2(0.f, 0.f);
        float2 xr = make_float2(0.f, 0.f);
        if (q == 0) {
            xq.y = 0.f;
        } else {
            xr = fm[r * nc + cpos];
            yq = make_float2(xr.x, -xr.y);
        }
        const float2 w = tw[q * tws];
        const float2 d = csub(xq, yq);
        const float2 zq = make_float2(xq.x + yq.x - (w.x * d.y + w.y * d.x), xq.y + yq.y + (w.x * d.x - w.y * d.y));
        fm[q * nc + cpos] = zq;
        if (q != 0 && r != q) {
            const float2 ym = make_float2(xq.x, -xq.y);
            const float2 wr = tw[r * tws];
            const float2 dr = csub(xr, ym);
            fm[r * nc + cpos] = make_float2(xr.x + ym.x - (wr.x * dr.y + wr.y * dr.x), xr.y + ym.y + (wr.x * dr.x - wr.y * dr.y));
        }
    }
    __syncthreads();
    // along a, down each column
    fft_dif<256>(fm, log2h, nc, 1, nc, tw, tid);
    const size_t vbase = ((size_t)b * a.nb + j) * h * nc;
    if (MODE == 1) {
        for (int e = tid; e < h * nc; e += 256) vol[vbase + e] = fm[e];
        return;
    }
    const float2* rd = rden + ((size_t)(b / rden_div) * a.nb + j) * h * nc;
    float bv = -3.4e38f;
    unsigned bi = 0xffffffffu;
    for (int e = tid; e < h * nc; e += 256) {
        const int ppos = e >> log2nc;
        const int cpos = e & (nc - 1);
        const float2 z = fm[e];
        const float2 r = rd[e];
        const unsigned base = ((unsigned)(j * nc + bitrev(cpos, log2nc)) * a.n) + 2u * (unsigned)bitrev(ppos, log2h);
        const float ve = z.x * r.x, vo = z.y * r.y;
        if (ve > bv || (ve == bv && base < bi)) { bv = ve; bi = base; }
        if (vo > bv || (vo == bv && base + 1 < bi)) { bv = vo; bi = base + 1; }
    }
    // the block's best: value first, then the lowest index
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const unsigned oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) {
        sv[tid >> 6] = bv;
        si[tid >> 6] = bi;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < 4; ++i)
            if (sv[i] > bv || (sv[i] == bv && si[i] < bi)) { bv = sv[i]; bi = si[i]; }
        const unsigned long long key = ((unsigned long long)ordered_bits(bv) << 32) | (unsigned long long)(0xffffffffu - bi);
        atomicMax(best + b, key);
    }
}

__global__ void k_clear_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// V[m', c] of item b at slab jj taken directly from U: sum_x U[x] e^{-2 pi i x c / nc}.
__device__ __forceinline__ float2 v_at(const SoArg& a, const float2* __restrict__ U, const float2* __restrict__ tw, int b, int jj, int mp, int c) {
    const float2* row = U + (((size_t)b * a.nb + jj) * a.L + mp) * a.nc;
    const int ts = ZB_TW / a.nc;
    float2 acc = make_float2(0.f, 0.f);
    for (int x = 0; x < a.nc; ++x) acc = cadd(acc, cmul(row[x], tw[((x * c) & (a.nc - 1)) * ts]));
    return acc;
}

// Re(V e^{-i m' a}) at grid point ia along a.
__device__ __forceinline__ float a_term(const SoArg& a, const float2* __restrict__ tw, float2 v, int mp, int ia) {
    const float2 e = tw[((mp * ia) & (a.n - 1)) * (ZB_TW / a.n)];
    return v.x * e.x - v.y * e.y;
}

__device__ __forceinline__ float rden_at(const SoArg& a, const float2* __restrict__ rden, int rb, int jj, int c, int ia) {
    const int h = a.n >> 1;
    const int ppos = bitrev(ia >> 1, a.log2n - 1);
    const int cpos = bitrev(c, a.log2nc);
    const float2 r = rden[(((size_t)rb * a.nb + jj) * h + ppos) * a.nc + cpos];
    return (ia & 1) ? r.y : r.x;
}

__device__ __forceinline__ float vertex(float lo, float mid, float hi) {
    const float den = lo - 2.f * mid + hi;
    if (den >= -1e-12f) return 0.f;
    return fminf(fmaxf(0.5f * (lo - hi) / den, -0.5f), 0.5f);
}

__global__ void __launch_bounds__(64) k_finish(SoArg a, int n_items, const float2* __restrict__ U, const float2* __restrict__ tw,
                                               const float2* __restrict__ rden, int rden_div, const unsigned long long* __restrict__ best,
                                               float* __restrict__ peaks) {
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    const unsigned long long key = best[b];
    float* out = peaks + (size_t)b * 8;
    if (key == 0ull) {
        if (lane == 0) {
            out[0] = 1.f; out[1] = 0.f; out[2] = 0.f; out[3] = 0.f;
            out[4] = -3.4e38f; out[5] = 0.f; out[6] = 0.f; out[7] = 0.f;
        }
        return;
    }
    const float v0 = from_ordered((unsigned)(key >> 32));
    const unsigned idx = 0xffffffffu - (unsigned)(key & 0xffffffffull);
    const int n = a.n, nc = a.nc;
    const int ia = idx & (n - 1);
    const int iq = (idx / n) & (nc - 1);
    const int j = idx / (n * nc);
    const bool inner = j > 0 && j < n / 2;
    const int iqm = (iq + nc - 1) & (nc - 1), iqp = (iq + 1) & (nc - 1);
    const int iam = (ia + n - 1) & (n - 1), iap = (ia + 1) & (n - 1);
    float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int mp = lane; mp < a.L; mp += 64) {
        const float f = mp == 0 ? 1.f : 2.f;
        const float2 v0c = v_at(a, U, tw, b, j, mp, iq);
        const float2 vm = v_at(a, U, tw, b, j, mp, iqm);
        const float2 vp = v_at(a, U, tw, b, j, mp, iqp);
        s[0] += f * a_term(a, tw, v0c, mp, iam);
        s[1] += f * a_term(a, tw, v0c, mp, iap);
        s[2] += f * a_term(a, tw, vm, mp, ia);
        s[3] += f * a_term(a, tw, vp, mp, ia);
        if (inner) {
            s[4] += f * a_term(a, tw, v_at(a, U, tw, b, j - 1, mp, iq), mp, ia);
            s[5] += f * a_term(a, tw, v_at(a, U, tw, b, j + 1, mp, iq), mp, ia);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = warp_sum(s[k]);
    if (lane != 0) return;
    const int rb = b / rden_div;
    const float da = vertex(s[0] * rden_at(a, rden, rb, j, iq, iam), v0, s[1] * rden_at(a, rden, rb, j, iq, iap));
    const float dc = vertex(s[2] * rden_at(a, rden, rb, j, iqm, ia), v0, s[3] * rden_at(a, rden, rb, j, iqp, ia));
    float dj = 0.f;
    if (inner) dj = vertex(s[4] * rden_at(a, rden, rb, j - 1, iq, ia), v0, s[5] * rden_at(a, rden, rb, j + 1, iq, ia));
    const float step = ZB_TAU / (float)n;
    const float aa = ((float)ia + da) * step;
    const float bb = ((float)j + dj) * step;
    const float cc = ((float)iq + dc) * step;
    // S = Rz(a) Ry(b) Rz(c); the orientation is its inverse
    const float4 qa = make_float4(cosf(0.5f * aa), 0.f, 0.f, sinf(0.5f * aa));
    const float4 qb = make_float4(cosf(0.5f * bb), 0.f, sinf(0.5f * bb), 0.f);
    const float4 qc = make_float4(cosf(0.5f * cc), 0.f, 0.f, sinf(0.5f * cc));
    float4 q = qmul(qa, qmul(qb, qc));
    q = make_float4(q.x, -q.y, -q.z, -q.w);
    if (q.x < 0.f) q = make_float4(-q.x, -q.y, -q.z, -q.w);
    out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = q.w;
    out[4] = v0; out[5] = aa; out[6] = bb; out[7] = cc;
}

// 1 / sqrt(<w, f^2> - <w, f>^2 / area), 0 where that is <= 1e-6 area (so3.wgsl:614-626).
__global__ void __launch_bounds__(256) k_rden(const float* __restrict__ vf, const float* __restrict__ vf2, const float* __restrict__ stats,
                                              long per_item, long total, float* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const float ar = fmaxf(stats[(i / per_item) * 4], 1e-12f);
    const float m = vf[i];
    const float v = vf2[i] - m * m / ar;
    out[i] = v > 1e-6f * ar ? 1.f / sqrtf(v) : 0.f;
}

}  // namespace zb

using namespace zb;

static int ilog2i(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return (1 << r) == v ? r : -1;
}

static int so_arg(const zb_band* band, const zb_phase* ph, SoArg& a, const char* who) {
    if (!band || !ph || band->L < 2 || ilog2i(band->n) < 4 || band->n < 2 * band->L - 1 || band->n > 512 || 2 * band->L > band->n + 1 ||
        !(ph->fold == 1 || ph->fold == 2 || ph->fold == 4) || ph->nc * ph->fold != band->n || ph->mu_max != (band->L - 1) / ph->fold ||
        !(ph->l_step == 1 || ph->l_step == 2) || !band->dq || !band->twiddle || !ph->w_f) {
        set_error("%s: bad band or phase", who);
        return ZB_EINVAL;
    }
    a.L = band->L;
    a.n = band->n;
    a.log2n = ilog2i(band->n);
    a.nc = ph->nc;
    a.log2nc = ilog2i(ph->nc);
    a.fold = ph->fold;
    a.mu_max = ph->mu_max;
    a.l_step = ph->l_step;
    a.pairs = band->L * (band->L + 1) / 2;
    a.nb = band->n / 2 + 1;
    a.Lh = (band->L + 1) / 2;
    a.Q = ph->nc >= 8 ? 8 : ph->nc;
    if (a.nc < 2 || (a.L > 64 * 4)) {
        set_error("%s: unsupported shape (nc %d, L %d)", who, a.nc, a.L);
        return ZB_EINVAL;
    }
    // the slab must fit the LDS of one workgroup
    if ((size_t)(a.n / 2) * a.nc * 8 > 160 * 1024) {
        set_error("%s: slab of %d x %d does not fit LDS", who, a.n / 2, a.nc);
        return ZB_EINVAL;
    }
    return ZB_OK;
}

static int64_t u_bytes(const SoArg& a, int n_items) { return (int64_t)n_items * a.nb * a.L * a.nc * 8; }
static int64_t vol_bytes(const SoArg& a, int n_items) { return (int64_t)n_items * a.nb * a.n * a.nc * 4; }
static int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }

static int64_t so3_ws(const SoArg& a, int n_items) {
    int64_t corr = align256(u_bytes(a, n_items)) + align256((int64_t)n_items * 8);
    int64_t den = align256(u_bytes(a, n_items)) + 2 * align256(vol_bytes(a, n_items));
    return corr > den ? corr : den;
}

extern "C" int64_t zb_so3_workspace(const zb_band* band, const zb_phase* phase, int32_t n_items) {
    SoArg a;
    if (so_arg(band, phase, a, "zb_so3_workspace") || n_items < 0) return ZB_EINVAL;
    return so3_ws(a, n_items);
}

static int launch_spectrum(const SoArg& a, int n_items, const float* w, const float* dq, const float* alm, const float2* tw, float2* U,
                           hipStream_t s) {
    const int groups = a.nc / a.Q;
    const dim3 grid((unsigned)((long)n_items * a.Lh * groups));
    const size_t lds = (size_t)a.Q * a.n * 8;
    const int kpt = (a.L + 63) / 64;
#define ZB_SPEC(K, QQ) ZB_LAUNCH("k_spectrum", (k_spectrum<K, QQ>), grid, dim3(64), lds, s, a, n_items, (const float2*)w, dq, (const float2*)alm, tw, U)
    if (a.Q == 8) {
        if (kpt == 1) ZB_SPEC(1, 8); else if (kpt == 2) ZB_SPEC(2, 8); else if (kpt == 3) ZB_SPEC(3, 8); else ZB_SPEC(4, 8);
    } else if (a.Q == 4) {
        if (kpt == 1) ZB_SPEC(1, 4); else ZB_SPEC(2, 4);
    } else {
        if (kpt == 1) ZB_SPEC(1, 2); else ZB_SPEC(2, 2);
    }
#undef ZB_SPEC
    return check_launch("spectrum");
}

extern "C" int32_t zb_so3_correlate(const zb_band* band, const zb_phase* phase, const float* alm, int32_t n_items, const float* rden,
                                    int32_t rden_div, void* work, int64_t work_bytes, float* peaks, void* stream) {
    SoArg a;
    if (int rc = so_arg(band, phase, a, "zb_so3_correlate")) return rc;
    if (n_items < 0 || rden_div < 1) {
        set_error("zb_so3_correlate: bad n_items %d or rden_div %d", n_items, rden_div);
        return ZB_EINVAL;
    }
    if (n_items == 0) return ZB_OK;
    if (!alm || !rden || !work || !peaks) {
        set_error("zb_so3_correlate: null pointer");
        return ZB_EINVAL;
    }
    if (work_bytes < so3_ws(a, n_items)) {
        set_error("zb_so3_correlate: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)so3_ws(a, n_items));
        return ZB_ESPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    const float2* tw = (const float2*)band->twiddle;
    float2* U = (float2*)work;
    unsigned long long* best = (unsigned long long*)((char*)work + align256(u_bytes(a, n_items)));
    if (int rc = launch_spectrum(a, n_items, phase->w_f, band->dq, alm, tw, U, s)) return rc;
    ZB_LAUNCH("k_clear_u64", k_clear_u64, dim3((n_items + 255) / 256), dim3(256), 0, s, best, n_items);
    const size_t lds = (size_t)(a.n / 2) * a.nc * 8;
    if (lds > 65536 && hipFuncSetAttribute((const void*)k_slab<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        set_error("slab of %zu bytes: LDS attribute refused", lds);
        return ZB_ELAUNCH;
    }
    ZB_LAUNCH("k_slab0", k_slab<0>, dim3((unsigned)((long)n_items * a.nb)), dim3(256), lds, s, a, n_items, (const float2*)U, tw,
                       (const float2*)rden, rden_div, best, (float2*)nullptr);
    if (int rc = check_launch("zb_so3_correlate/slab")) return rc;
    ZB_LAUNCH("k_finish", k_finish, dim3(n_items), dim3(64), 0, s, a, n_items, (const float2*)U, tw, (const float2*)rden, rden_div,
                       (const unsigned long long*)best, peaks);
    return check_launch("zb_so3_correlate/finish");
}

extern "C" int32_t zb_so3_denominators(const zb_band* band, const zb_phase* phase, const float* w_alm, const float* w_stats,
                                       int32_t n_windows, void* work, int64_t work_bytes, float* rden, void* stream) {
    SoArg a;
    if (int rc = so_arg(band, phase, a, "zb_so3_denominators")) return rc;
    if (n_windows < 0 || (n_windows > 0 && (!w_alm || !w_stats || !work || !rden || !phase->w_f2))) {
        set_error("zb_so3_denominators: bad arguments");
        return ZB_EINVAL;
    }
    if (n_windows == 0) return ZB_OK;
    if (work_bytes < so3_ws(a, n_windows)) {
        set_error("zb_so3_denominators: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)so3_ws(a, n_windows));
        return ZB_ESPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    const float2* tw = (const float2*)band->twiddle;
    float2* U = (float2*)work;
    float* vf = (float*)((char*)work + align256(u_bytes(a, n_windows)));
    float* vf2 = (float*)((char*)vf + align256(vol_bytes(a, n_windows)));
    const size_t lds = (size_t)(a.n / 2) * a.nc * 8;
    if (lds > 65536 && hipFuncSetAttribute((const void*)k_slab<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        set_error("slab of %zu bytes: LDS attribute refused", lds);
        return ZB_ELAUNCH;
    }
    const float* ws[2] = {phase->w_f, phase->w_f2};
    float* vs[2] = {vf, vf2};
    for (int i = 0; i < 2; ++i) {
        if (int rc = launch_spectrum(a, n_windows, ws[i], band->dq, w_alm, tw, U, s)) return rc;
        ZB_LAUNCH("k_slab1", k_slab<1>, dim3((unsigned)((long)n_windows * a.nb)), dim3(256), lds, s, a, n_windows, (const float2*)U, tw,
                           (const float2*)nullptr, 1, (unsigned long long*)nullptr, (float2*)vs[i]);
        if (int rc = check_launch("zb_so3_denominators/volume")) return rc;
    }
    const long per_item = (long)a.nb * a.n * a.nc;
    const long total = per_item * n_windows;
    ZB_LAUNCH("k_rden", k_rden, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, vf, vf2, w_stats, per_item, total, rden);
    return check_launch("zb_so3_denominators/rden");
}
