// The SO(3) correlation as one dense GEMM per (j, m') on the matrix cores, then the last axis in
// registers (gfx950). Same correlation volume as so3.hip's spectrum + slab path, reached another
// way:
//
//   V[j, m', c] = sum_l Q[j, m', c, l] conj(g_{l m'})                        (k_vgemm, MFMA)
//   Q[j, m', c, l] = sum_mu f_{l, fold mu} d^l_{m', fold mu}(beta_j) e^{-2 pi i mu c / nc}
//
// i.e. the reference's spectrum T summed over k against e^{-ik beta_j} (its transform along b)
// and over its orders m against e^{-i m c} (its transform along c) are folded into a
// pattern-independent table Q, built once per phase and bandwidth on the host (float64). What is
// left per pattern is a complex GEMM per (j, m') with the item's coefficients (K = its degrees l
// >= m'), and the real transform along a.
//
//   k_vgemm   one wave per (m', tile of 32 MB items), looping over every j: the items'
//             coefficients for m' are the MFMA A operand (rows = items, k = (re, im) of one
//             degree), held in registers across all j; the table rows are the B operand
//             ([Q_re, Q_im] and [Q_im, -Q_re] so one real MFMA per degree gives V_re | V_im).
//             v_mfma_f32_32x32x2_f32 (float32 in, float32 accumulate). V goes to HBM as two
//             planes, [item][j][m'][c], each write a run of columns.
//   k_apeak   one lane per (item, j, c) column: the column's L orders into the half-length
//             complex input of the real transform along a, a radix-2 FFT of h = n/2 points in
//             registers (twiddles are compile-time constants), times the reciprocal denominator,
//             the best (value, lowest index) of the column, of the slab (xor shuffles over its
//             nc lanes), of the item (one 64-bit atomicMax). Or the volume itself (denominators).
//   k_vfinish the peak's six neighbours summed directly from V (already along c), parabolic
//             offsets and the orientation, as so3.hip's k_finish.
//
// Volumes here (rden, and the window's) are in natural order, transposed: [j][a][c].
#include "common.h"
#include "tw128.h"

namespace zb {

typedef float floatx16 __attribute__((ext_vector_type(16)));

struct VArg {
    int L, n, log2n, nc, log2nc, nb, slots, cols, l_step, pairs, tiles, n_items;
};

__device__ __forceinline__ int xcd_block(int id, int G) {
    // blocks are dealt to the 8 XCDs round-robin: give each XCD a contiguous run of the logical
    // order, so one XCD's L2 holds the table rows of a few m'
    return (G % 8 == 0) ? (id % 8) * (G / 8) + id / 8 : id;
}

template <int MB, int NBLK, int S>
__global__ void __launch_bounds__(64) k_vgemm(VArg a, const float* __restrict__ tab, const float2* __restrict__ coef,
                                              float* __restrict__ vre, float* __restrict__ vim) {
    const int logical = xcd_block(blockIdx.x, gridDim.x);
    const int mp = logical / a.tiles;
    const int tile = logical - mp * a.tiles;
    const int lane = threadIdx.x;
    const int half = lane >> 5;
    const int r32 = lane & 31;
    const int s0 = (mp + a.l_step - 1) / a.l_step;
    const int pbase = mp * a.L - (mp * (mp - 1)) / 2 - mp;  // pair(mp, l) = pbase + l
    float areg[MB][S];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int item = (tile * MB + mb) * 32 + r32;
        const bool ok = item < a.n_items;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            float v = 0.f;
            if (ok && s >= s0 && s < a.slots) {
                const float2 g = coef[(size_t)item * a.pairs + pbase + s * a.l_step];
                v = half ? g.y : g.x;
            }
            areg[mb][s] = v;
        }
    }
    const size_t per_item = (size_t)a.nb * a.L * a.nc;
    for (int j = 0; j < a.nb; ++j) {
        floatx16 acc[MB][NBLK];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.f;
        const float* tj = tab + ((size_t)j * a.L + mp) * a.slots * 2 * a.cols + half * a.cols + r32;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s >= s0 && s < a.slots) {
                float bv[NBLK];
#pragma unroll
                for (int nb = 0; nb < NBLK; ++nb) bv[nb] = tj[(size_t)s * 2 * a.cols + nb * 32];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NBLK; ++nb) acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(areg[mb][s], bv[nb], acc[mb][nb], 0, 0, 0);
            }
        }
        // C[item row][col]: lane holds column r32 of rows (r/4) 8 + half 4 + r % 4
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb) {
                const int col = nb * 32 + r32;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int item = (tile * MB + mb) * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                    if (item < a.n_items) {
                        const size_t at = (size_t)item * per_item + ((size_t)j * a.L + mp) * a.nc;
                        if (col < a.nc)
                            vre[at + col] = acc[mb][nb][r];
                        else if (col < 2 * a.nc)
                            vim[at + col - a.nc] = acc[mb][nb][r];
                    }
                }
            }
    }
}

__device__ __forceinline__ unsigned ord_bits(float v) {
    unsigned u = __float_as_uint(v + 0.f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float from_ord(unsigned o) {
    unsigned u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

template <int LG>
__device__ __forceinline__ int brev_c(int x) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < LG; ++b) r |= ((x >> b) & 1) << (LG - 1 - b);
    return r;
}

// exp(-2 pi i t / N) from the 128-point constants (N divides 128)
template <int N>
__device__ __forceinline__ float2 twc(int t) {
    const int k = (t * (128 / N)) & 127;
    return make_float2(kTwCos[k], -kTwSin[k]);
}

template <int H, int LH, int MODE>
__global__ void __launch_bounds__(64) k_apeak(VArg a, const float* __restrict__ vre, const float* __restrict__ vim,
                                              const float* __restrict__ rden, int rden_div, unsigned long long* __restrict__ best,
                                              float* __restrict__ vol) {
    const int lane = threadIdx.x;
    const int sl = lane >> a.log2nc;
    const int c = lane & (a.nc - 1);
    const long slab = (long)blockIdx.x * (64 >> a.log2nc) + sl;
    const bool live = slab < (long)a.n_items * a.nb;
    const long sc = live ? slab : 0;
    const int b = (int)(sc / a.nb);
    const int j = (int)(sc - (long)b * a.nb);
    const size_t base = ((size_t)b * a.nb + j) * a.L * a.nc + c;
    float2 X[H];
#pragma unroll
    for (int q = 0; q < H; ++q) X[q] = (live && q < a.L) ? make_float2(vre[base + (size_t)q * a.nc], vim[base + (size_t)q * a.nc]) : make_float2(0.f, 0.f);
    // the real transform along a (C = V_0 + 2 Re sum_{m' > 0} V_m' e^{-i m' a}) from a complex one
    // of half its length: Z_q = (X_q + X_{q+h}) + i e^{-2 pi i q / n} (X_q - X_{q+h}),
    // X_{q+h} = conj X_{h-q}, X_0 real
    float2 z[H];
    z[0] = make_float2(X[0].x, X[0].x);
#pragma unroll
    for (int q = 1; q < H; ++q) {
        const float2 x = X[q];
        const float2 y = make_float2(X[H - q].x, -X[H - q].y);
        const float2 w = twc<2 * H>(q);
        const float2 d = csub(x, y);
        z[q] = make_float2(x.x + y.x - (w.x * d.y + w.y * d.x), x.y + y.y + (w.x * d.x - w.y * d.y));
    }
    // radix-2 decimation in frequency: natural in, bit-reversed out
#pragma unroll
    for (int hh = H / 2; hh >= 1; hh >>= 1) {
#pragma unroll
        for (int blk = 0; blk < H; blk += 2 * hh) {
#pragma unroll
            for (int jj = 0; jj < hh; ++jj) {
                const float2 u = z[blk + jj], v = z[blk + jj + hh];
                z[blk + jj] = cadd(u, v);
                z[blk + jj + hh] = cmul(csub(u, v), twc<128>(jj * (64 / hh)));
            }
        }
    }
    const int n = a.n;
    if (MODE == 1) {
        if (!live) return;
        float* out = vol + ((size_t)b * a.nb + j) * n * a.nc + c;
#pragma unroll
        for (int p = 0; p < H; ++p) {
            const int f = brev_c<LH>(p);
            out[(size_t)(2 * f) * a.nc] = z[p].x;
            out[(size_t)(2 * f + 1) * a.nc] = z[p].y;
        }
        return;
    }
    const float* rd = rden + ((size_t)(b / rden_div) * a.nb + j) * n * a.nc + c;
    float bv = -3.4e38f;
    unsigned bi = 0xffffffffu;
    const unsigned ibase = (unsigned)(j * a.nc + c) * (unsigned)n;
#pragma unroll
    for (int p = 0; p < H; ++p) {
        const int f = brev_c<LH>(p);
        const float ve = z[p].x * rd[(size_t)(2 * f) * a.nc];
        const float vo = z[p].y * rd[(size_t)(2 * f + 1) * a.nc];
        const unsigned ie = ibase + 2u * f;
        if (ve > bv || (ve == bv && ie < bi)) { bv = ve; bi = ie; }
        if (vo > bv || (vo == bv && ie + 1 < bi)) { bv = vo; bi = ie + 1; }
    }
    if (!live) { bv = -3.4e38f; bi = 0xffffffffu; }
    for (int o = 1; o < a.nc; o <<= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const unsigned oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (live && c == 0) {
        const unsigned long long key = ((unsigned long long)ord_bits(bv) << 32) | (unsigned long long)(0xffffffffu - bi);
        atomicMax(best + b, key);
    }
}

__global__ void k_vclear(unsigned long long* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__device__ __forceinline__ float vertex_v(float lo, float mid, float hi) {
    const float den = lo - 2.f * mid + hi;
    if (den >= -1e-12f) return 0.f;
    return fminf(fmaxf(0.5f * (lo - hi) / den, -0.5f), 0.5f);
}

__global__ void __launch_bounds__(64) k_vfinish(VArg a, const float* __restrict__ vre, const float* __restrict__ vim, const float2* __restrict__ tw,
                                                const float* __restrict__ rden, int rden_div, const unsigned long long* __restrict__ best,
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
    const float v0 = from_ord((unsigned)(key >> 32));
    const unsigned idx = 0xffffffffu - (unsigned)(key & 0xffffffffull);
    const int n = a.n, nc = a.nc;
    const int ia = idx & (n - 1);
    const int iq = (idx >> a.log2n) & (nc - 1);
    const int j = idx >> (a.log2n + a.log2nc);
    const bool inner = j > 0 && j < n / 2;
    const int iqm = (iq + nc - 1) & (nc - 1), iqp = (iq + 1) & (nc - 1);
    const int iam = (ia + n - 1) & (n - 1), iap = (ia + 1) & (n - 1);
    const int ts = ZB_TW / n;
    float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    auto V = [&](int jj, int mp, int c) {
        const size_t at = (((size_t)b * a.nb + jj) * a.L + mp) * nc + c;
        return make_float2(vre[at], vim[at]);
    };
    auto term = [&](float2 v, int mp, int x) {
        const float2 e = tw[((mp * x) & (n - 1)) * ts];
        return v.x * e.x - v.y * e.y;
    };
    for (int mp = lane; mp < a.L; mp += 64) {
        const float f = mp == 0 ? 1.f : 2.f;
        const float2 v0c = V(j, mp, iq);
        s[0] += f * term(v0c, mp, iam);
        s[1] += f * term(v0c, mp, iap);
        s[2] += f * term(V(j, mp, iqm), mp, ia);
        s[3] += f * term(V(j, mp, iqp), mp, ia);
        if (inner) {
            s[4] += f * term(V(j - 1, mp, iq), mp, ia);
            s[5] += f * term(V(j + 1, mp, iq), mp, ia);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = warp_sum(s[k]);
    if (lane != 0) return;
    const float* rd = rden + (size_t)(b / rden_div) * a.nb * n * nc;
    auto R = [&](int jj, int c, int x) { return rd[((size_t)jj * n + x) * nc + c]; };
    const float da = vertex_v(s[0] * R(j, iq, iam), v0, s[1] * R(j, iq, iap));
    const float dc = vertex_v(s[2] * R(j, iqm, ia), v0, s[3] * R(j, iqp, ia));
    float dj = 0.f;
    if (inner) dj = vertex_v(s[4] * R(j - 1, iq, ia), v0, s[5] * R(j + 1, iq, ia));
    const float step = ZB_TAU / (float)n;
    const float aa = ((float)ia + da) * step;
    const float bb = ((float)j + dj) * step;
    const float cc = ((float)iq + dc) * step;
    const float4 qa = make_float4(cosf(0.5f * aa), 0.f, 0.f, sinf(0.5f * aa));
    const float4 qb = make_float4(cosf(0.5f * bb), 0.f, sinf(0.5f * bb), 0.f);
    const float4 qc = make_float4(cosf(0.5f * cc), 0.f, 0.f, sinf(0.5f * cc));
    float4 q = qmul(qa, qmul(qb, qc));
    q = make_float4(q.x, -q.y, -q.z, -q.w);
    if (q.x < 0.f) q = make_float4(-q.x, -q.y, -q.z, -q.w);
    out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = q.w;
    out[4] = v0; out[5] = aa; out[6] = bb; out[7] = cc;
}

__global__ void __launch_bounds__(256) k_vrden(const float* __restrict__ vf, const float* __restrict__ vf2, const float* __restrict__ stats,
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

static int ilog2v(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return (1 << r) == v ? r : -1;
}

static int v_arg(const zb_vtable* t, int n_items, VArg& a, const char* who) {
    if (!t || t->L < 2 || ilog2v(t->n) < 4 || t->n < 2 * t->L - 1 || ilog2v(t->nc) < 1 || t->nc > 64 || t->n / t->nc > 4 ||
        !(t->l_step == 1 || t->l_step == 2) || t->slots != (t->L + t->l_step - 1) / t->l_step || t->cols != 32 * ((2 * t->nc + 31) / 32) ||
        t->n / 2 > 64 || !t->q) {
        set_error("%s: bad table (L %d, n %d, nc %d, slots %d, cols %d; this path takes n <= 128, nc <= 64)", who, t ? t->L : 0, t ? t->n : 0,
                  t ? t->nc : 0, t ? t->slots : 0, t ? t->cols : 0);
        return ZB_EINVAL;
    }
    if (t->slots > 64) {
        set_error("%s: %d degree slots: at most 64", who, t->slots);
        return ZB_EINVAL;
    }
    a.L = t->L;
    a.n = t->n;
    a.log2n = ilog2v(t->n);
    a.nc = t->nc;
    a.log2nc = ilog2v(t->nc);
    a.nb = t->n / 2 + 1;
    a.slots = t->slots;
    a.cols = t->cols;
    a.l_step = t->l_step;
    a.pairs = t->L * (t->L + 1) / 2;
    a.n_items = n_items;
    a.tiles = 1;
    return ZB_OK;
}

static int64_t v_bytes(const VArg& a, int n_items) { return (int64_t)n_items * a.nb * a.L * a.nc * 4; }
static int64_t al256(int64_t v) { return (v + 255) / 256 * 256; }
// V (two planes) and the per-item best; the denominators add two volumes
static int64_t vcorr_ws(const VArg& a, int n_items, int volumes) {
    return 2 * al256(v_bytes(a, n_items)) + al256((int64_t)n_items * 8) + (volumes ? 2 * al256((int64_t)n_items * a.nb * a.n * a.nc * 4) : 0);
}

extern "C" int64_t zb_vcorr_workspace(const zb_vtable* t, int32_t n_items, int32_t denominators) {
    VArg a;
    if (v_arg(t, n_items, a, "zb_vcorr_workspace")) return ZB_EINVAL;
    if (n_items < 0) {
        set_error("zb_vcorr_workspace: n_items %d", n_items);
        return ZB_EINVAL;
    }
    return vcorr_ws(a, n_items, denominators != 0);
}

template <int MB, int NBLK, int S>
static void launch_vgemm_t(const VArg& a, const float* q, const float* alm, float* vre, float* vim, hipStream_t s) {
    const int tiles = (a.n_items + 32 * MB - 1) / (32 * MB);
    VArg b = a;
    b.tiles = tiles;
    ZB_LAUNCH("k_vgemm", (k_vgemm<MB, NBLK, S>), dim3((unsigned)(tiles * a.L)), dim3(64), 0, s, b, q, (const float2*)alm, vre, vim);
}

static int launch_vgemm(const VArg& a, const float* q, const float* alm, float* vre, float* vim, hipStream_t s) {
    const int nblk = a.cols / 32;
    const int S = a.slots <= 8 ? 8 : a.slots <= 16 ? 16 : a.slots <= 32 ? 32 : 64;
#define ZB_VG(MB, NB, SS) launch_vgemm_t<MB, NB, SS>(a, q, alm, vre, vim, s)
    if (nblk == 1) {
        if (S == 8) ZB_VG(2, 1, 8); else if (S == 16) ZB_VG(2, 1, 16); else if (S == 32) ZB_VG(2, 1, 32); else ZB_VG(1, 1, 64);
    } else if (nblk == 2) {
        if (S == 8) ZB_VG(2, 2, 8); else if (S == 16) ZB_VG(2, 2, 16); else if (S == 32) ZB_VG(2, 2, 32); else ZB_VG(1, 2, 64);
    } else {
        if (S == 8) ZB_VG(1, 4, 8); else if (S == 16) ZB_VG(1, 4, 16); else if (S == 32) ZB_VG(1, 4, 32); else ZB_VG(1, 4, 64);
    }
#undef ZB_VG
    return check_launch("k_vgemm");
}

template <int MODE>
static int launch_apeak(const VArg& a, const float* vre, const float* vim, const float* rden, int rden_div, unsigned long long* best, float* vol,
                        hipStream_t s) {
    const int per_block = 64 >> a.log2nc;
    const long slabs = (long)a.n_items * a.nb;
    const dim3 grid((unsigned)((slabs + per_block - 1) / per_block));
    const int H = a.n / 2;
#define ZB_AP(HH, LL) ZB_LAUNCH("k_apeak", (k_apeak<HH, LL, MODE>), grid, dim3(64), 0, s, a, vre, vim, rden, rden_div, best, vol)
    if (H == 8) ZB_AP(8, 3); else if (H == 16) ZB_AP(16, 4); else if (H == 32) ZB_AP(32, 5); else ZB_AP(64, 6);
#undef ZB_AP
    return check_launch("k_apeak");
}

extern "C" int32_t zb_vcorr_correlate(const zb_vtable* t, const float* twiddle, const float* alm, int32_t n_items, const float* rden,
                                      int32_t rden_div, void* work, int64_t work_bytes, float* peaks, void* stream) {
    VArg a;
    if (int rc = v_arg(t, n_items, a, "zb_vcorr_correlate")) return rc;
    if (n_items < 0 || rden_div < 1) {
        set_error("zb_vcorr_correlate: bad n_items %d or rden_div %d", n_items, rden_div);
        return ZB_EINVAL;
    }
    if (n_items == 0) return ZB_OK;
    if (!twiddle || !alm || !rden || !work || !peaks) {
        set_error("zb_vcorr_correlate: null pointer");
        return ZB_EINVAL;
    }
    if (work_bytes < vcorr_ws(a, n_items, 0)) {
        set_error("zb_vcorr_correlate: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)vcorr_ws(a, n_items, 0));
        return ZB_ESPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    float* vre = (float*)work;
    float* vim = (float*)((char*)work + al256(v_bytes(a, n_items)));
    unsigned long long* best = (unsigned long long*)((char*)vim + al256(v_bytes(a, n_items)));
    if (int rc = launch_vgemm(a, t->q, alm, vre, vim, s)) return rc;
    ZB_LAUNCH("k_vclear", k_vclear, dim3((n_items + 255) / 256), dim3(256), 0, s, best, n_items);
    if (int rc = launch_apeak<0>(a, vre, vim, rden, rden_div, best, nullptr, s)) return rc;
    ZB_LAUNCH("k_vfinish", k_vfinish, dim3(n_items), dim3(64), 0, s, a, vre, vim, (const float2*)twiddle, rden, rden_div,
              (const unsigned long long*)best, peaks);
    return check_launch("zb_vcorr_correlate/finish");
}

extern "C" int32_t zb_vcorr_denominators(const zb_vtable* t_f, const zb_vtable* t_f2, const float* w_alm, const float* w_stats, int32_t n_windows,
                                         void* work, int64_t work_bytes, float* rden, void* stream) {
    VArg a, a2;
    if (int rc = v_arg(t_f, n_windows, a, "zb_vcorr_denominators")) return rc;
    if (int rc = v_arg(t_f2, n_windows, a2, "zb_vcorr_denominators")) return rc;
    if (a.L != a2.L || a.n != a2.n || a.nc != a2.nc || a.l_step != a2.l_step || n_windows < 0 ||
        (n_windows > 0 && (!w_alm || !w_stats || !work || !rden))) {
        set_error("zb_vcorr_denominators: bad arguments");
        return ZB_EINVAL;
    }
    if (n_windows == 0) return ZB_OK;
    if (work_bytes < vcorr_ws(a, n_windows, 1)) {
        set_error("zb_vcorr_denominators: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)vcorr_ws(a, n_windows, 1));
        return ZB_ESPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    float* vre = (float*)work;
    float* vim = (float*)((char*)work + al256(v_bytes(a, n_windows)));
    char* after = (char*)vim + al256(v_bytes(a, n_windows)) + al256((int64_t)n_windows * 8);
    const int64_t vb = al256((int64_t)n_windows * a.nb * a.n * a.nc * 4);
    float* vf = (float*)after;
    float* vf2 = (float*)(after + vb);
    if (int rc = launch_vgemm(a, t_f->q, w_alm, vre, vim, s)) return rc;
    if (int rc = launch_apeak<1>(a, vre, vim, nullptr, 1, nullptr, vf, s)) return rc;
    if (int rc = launch_vgemm(a2, t_f2->q, w_alm, vre, vim, s)) return rc;
    if (int rc = launch_apeak<1>(a, vre, vim, nullptr, 1, nullptr, vf2, s)) return rc;
    const long per_item = (long)a.nb * a.n * a.nc;
    const long total = per_item * n_windows;
    ZB_LAUNCH("k_vrden", k_vrden, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, vf, vf2, w_stats, per_item, total, rden);
    return check_launch("zb_vcorr_denominators/rden");
}
