// Patterns onto the sphere and into spherical harmonics (the reference's sphere.wgsl stage,
// src/spherical.zig:319-361), for gfx950.
//
// An item is one (detector position t, pattern) pair. The sphere is sampled on a Gauss-Legendre
// grid of n rings by 2n points in the sample frame: each grid direction's ray meets the detector
// plane, and the bilinear pattern value there counts where it lands on the detector and inside
// the mask (the window). Four kernels:
//   k_sums     per item, the window's quadrature sums (for its mean)
//   k_rings    ring y with its mirror n-1-y as one complex line (real: the ring, imaginary: the
//              mirror), mean-subtracted and windowed, FFT along phi in LDS, split apart; of the
//              orders m < L the sum and difference of the two rings' spectra are kept
//   k_legendre a_lm = sum_y wplm[y, pair] (sum or difference by the parity of l + m): per order
//              m and tile of 16 items, two GEMMs on the matrix cores (one per parity)
//   k_norm     unit harmonic norm and the item's stats
// Workspace: per item the window sums (float2), and the ring spectra [L][items][n/2] float4
// (order-major, so that a k_legendre block's items lie side by side, not a power-of-two stride
// apart).
#include <mutex>

#include "common.h"

namespace zb {

struct SphArg {
    float c0[3], c1[3], c2[3];
    float mask[3];
    int h, w;
    int L, n, log2nphi, pairs, flags;
};

// The sample of an item along grid direction d (sphere.wgsl:67-101): the ray meets the detector
// plane at h = (c2.t / c2.d) d - t, so row = c0.h and col = c1.h. Given nd = c2.d, p0 = c0.d and
// p1 = c1.d, and the item's t0 = c0.t, t1 = c1.t, t2 = c2.t: row = s p0 - t0, col = s p1 - t1 with
// s = t2 / nd. Returns (pattern value, window): the bilinear value where the ray lands on the
// detector and inside the mask, else (0, 0).
struct ItemGeom {
    float t0, t1, t2;
    const float* pat;
};

__device__ __forceinline__ ItemGeom item_geom(const SphArg& a, const float* __restrict__ pats, float4 it) {
    ItemGeom g;
    g.t0 = a.c0[0] * it.x + a.c0[1] * it.y + a.c0[2] * it.z;
    g.t1 = a.c1[0] * it.x + a.c1[1] * it.y + a.c1[2] * it.z;
    g.t2 = a.c2[0] * it.x + a.c2[1] * it.y + a.c2[2] * it.z;
    g.pat = pats + (size_t)(int)it.w * a.h * a.w;
    return g;
}

// A grid column (angle phi): the frame's rows dotted with the horizontal direction (cos phi,
// sin phi, 0), so that for a ring (cos theta = ct, sin theta = st) c_k.d = st q_k + c_k.z ct.
struct ColGeom {
    float q0, q1, q2;
};

__device__ __forceinline__ ColGeom col_geom(const SphArg& a, float2 e) {  // e = exp(-i phi)
    const float cx = e.x, sy = -e.y;
    return {a.c0[0] * cx + a.c0[1] * sy, a.c1[0] * cx + a.c1[1] * sy, a.c2[0] * cx + a.c2[1] * sy};
}

__device__ __forceinline__ float2 sample_dir(const SphArg& a, const ItemGeom& g, float nd, float p0, float p1) {
    float val = 0.f, w = 0.f;
    if (nd > 1e-4f) {
        // (the hardware reciprocal, 1 ulp: the ray parameter of a sample, not a decision)
        const float s = g.t2 * __builtin_amdgcn_rcpf(nd);
        const float row = s * p0 - g.t0;
        const float col = s * p1 - g.t1;
        const float dr = row - a.mask[0], dc = col - a.mask[1];
        if (row >= 0.f && row <= (float)(a.h - 1) && col >= 0.f && col <= (float)(a.w - 1) &&
            (a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2])) {
            w = 1.f;
            val = 1.f;
            if ((a.flags & ZB_FLAG_WINDOW) == 0) {
                const int r0 = min((int)row, a.h - 2), q0 = min((int)col, a.w - 2);
                const float fr = row - (float)r0, fc = col - (float)q0;
                const float* p = g.pat + r0 * a.w + q0;
                const float top = lerpf(p[0], p[1], fc);
                const float bot = lerpf(p[a.w], p[a.w + 1], fc);
                val = lerpf(top, bot, fr);
            }
        }
    }
    return make_float2(val, w);
}

// A ring pair's values as a block's threads read them from LDS: cos theta of the northern ring,
// the sine both rings share, and their quadrature weight. Computed once per ring by the block
// (the sine's square root takes some twenty instructions) instead of once per sample pair.
__device__ __forceinline__ float4 ring_entry(const float* __restrict__ ring, int y) {
    const float ct = ring[2 * y];
    return make_float4(ct, sqrtf(fmaxf(1.f - ct * ct, 0.f)), ring[2 * y + 1], 0.f);
}

// The samples of an item at grid column cg on ring y (cos theta = ct, sin theta = st) and on its
// mirror ring n - 1 - y (cos theta = -ct, the same sine).
__device__ __forceinline__ void sample_pair(const SphArg& a, const ItemGeom& g, const ColGeom& cg, float ct, float st, float2& north,
                                            float2& south) {
    const float h0 = st * cg.q0, h1 = st * cg.q1, h2 = st * cg.q2;
    const float z0 = a.c0[2] * ct, z1 = a.c1[2] * ct, z2 = a.c2[2] * ct;
    north = sample_dir(a, g, h2 + z2, h0 + z0, h1 + z1);
    south = sample_dir(a, g, h2 - z2, h0 - z0, h1 - z1);
}

// Per item: sum of w wq value and of w wq over its whole grid (sphere.wgsl:114-151). Each thread
// keeps grid columns (phi) and walks the ring pairs (y, n - 1 - y), whose weights are equal.
__global__ void __launch_bounds__(256) k_sums(SphArg a, const float* __restrict__ ring, const float2* __restrict__ tw,
                                              const float* __restrict__ pats, const float4* __restrict__ items, float2* __restrict__ sums) {
    __shared__ float2 red[4];
    __shared__ float4 rt[ZB_TW / 4];  // the ring pairs' entries (n / 2 <= ZB_TW / 4)
    const int b = blockIdx.x;
    const ItemGeom g = item_geom(a, pats, items[b]);
    const int nphi = 1 << a.log2nphi;
    const int nh = a.n >> 1;
    const int cols = nphi < 256 ? nphi : 256;   // columns a pass covers
    const int rows = 256 / cols;                // ring pairs a pass covers
    const int tid = threadIdx.x;
    for (int y = tid; y < nh; y += 256) rt[y] = ring_entry(ring, y);
    __syncthreads();
    float sv = 0.f, sw = 0.f;
    for (int j = tid % cols; j < nphi; j += cols) {
        const ColGeom cg = col_geom(a, tw[j << (10 - a.log2nphi)]);
        for (int y = tid / cols; y < nh; y += rows) {
            float2 gn, gs;
            const float4 r = rt[y];
            sample_pair(a, g, cg, r.x, r.y, gn, gs);
            const float wq = r.z;
            sv += (gn.x * gn.y + gs.x * gs.y) * wq;
            sw += (gn.y + gs.y) * wq;
        }
    }
    sv = warp_sum(sv);
    sw = warp_sum(sw);
    if ((tid & 63) == 0) red[tid >> 6] = make_float2(sv, sw);
    __syncthreads();
    if (tid == 0) {
        float2 t = red[0];
        for (int i = 1; i < 4; ++i) t = make_float2(t.x + red[i].x, t.y + red[i].y);
        sums[b] = t;
    }
}

__device__ __forceinline__ float window_mean(const SphArg& a, float2 tot) {
    return (a.flags & ZB_FLAG_CENTER) ? tot.x / fmaxf(tot.y, 1e-20f) : 0.f;
}

// Ring pairs y0 .. y0 + G - 1 of one item: sample, FFT along phi, keep m < L (sphere.wgsl:157-207).
// LOG2NPHI: the points a ring has, 2^LOG2NPHI, known at compile time (the FFT's index arithmetic
// folded into constants); 0: any, from the arguments.
template <int LOG2NPHI>
__global__ void __launch_bounds__(256) k_rings(SphArg a, int G, const float* __restrict__ ring, const float2* __restrict__ tw,
                                               const float* __restrict__ pats, const float4* __restrict__ items,
                                               const float2* __restrict__ sums, float4* __restrict__ spectra) {
    extern __shared__ float2 buf[];  // [G][nphi] lines, then the group's G ring entries
    if constexpr (LOG2NPHI > 0) a.log2nphi = LOG2NPHI;
    const int nh = a.n >> 1;
    const int groups = nh / G;
    const int b = blockIdx.x / groups;
    const int y0 = (blockIdx.x - b * groups) * G;
    const int nphi = 1 << a.log2nphi;
    float4* rt = reinterpret_cast<float4*>(buf + G * nphi);
    for (int i = threadIdx.x; i < G; i += 256) rt[i] = ring_entry(ring, y0 + i);
    __syncthreads();
    const ItemGeom g = item_geom(a, pats, items[b]);
    const float mean = window_mean(a, sums[b]);
    // each thread keeps grid columns and walks the group's rings
    const int cols = nphi < 256 ? nphi : 256;
    const int rows = 256 / cols;
    for (int j = threadIdx.x % cols; j < nphi; j += cols) {
        const ColGeom cg = col_geom(a, tw[j << (10 - a.log2nphi)]);
        for (int i = threadIdx.x / cols; i < G; i += rows) {
            float2 gn, gs;
            const float4 r = rt[i];
            sample_pair(a, g, cg, r.x, r.y, gn, gs);
            buf[i * nphi + j] = make_float2((gn.x - mean) * gn.y, (gs.x - mean) * gs.y);
        }
    }
    __syncthreads();
    if constexpr (LOG2NPHI > 0)
        fft_dif_fixed<256, LOG2NPHI>(buf, G, tw, threadIdx.x);
    else
        fft_dif<256>(buf, a.log2nphi, G, nphi, 1, tw, threadIdx.x);
    // Z = N + i S for real N, S: N_m = (Z_m + conj Z_{-m}) / 2, S_m = (Z_m - conj Z_{-m}) / 2i
    const int L = a.L;
    const int lg = __builtin_ctz((unsigned)G);  // G is a power of two
    const int n_items = gridDim.x / groups;
    for (int e = threadIdx.x; e < G * L; e += 256) {
        const int m = e >> lg;
        const int i = e & (G - 1);
        const float2 z = buf[i * nphi + bitrev(m, a.log2nphi)];
        const float2 zr = buf[i * nphi + bitrev((nphi - m) & (nphi - 1), a.log2nphi)];
        const float2 rn = make_float2(0.5f * (z.x + zr.x), 0.5f * (z.y - zr.y));
        const float2 rs = make_float2(0.5f * (z.y + zr.y), -0.5f * (z.x - zr.x));
        spectra[((size_t)m * n_items + b) * nh + y0 + i] = make_float4(rn.x + rs.x, rn.y + rs.y, rn.x - rs.x, rn.y - rs.y);
    }
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// a_lm of 16 items at one order m on the matrix cores (sphere.wgsl:209-226): per parity of l + m
// a GEMM with rows = (item, re / im) of the ring-pair spectra (N + S for even l + m, N - S for
// odd), k = the ring y, columns = the degrees of that parity, against the Legendre table.
// v_mfma_f32_32x32x2_f32: A lane -> (row lane % 32, k lane / 32), B lane -> (k lane / 32,
// column lane % 32), C register r -> (row (r / 4) 8 + (lane / 32) 4 + r % 4, column lane % 32).
__global__ void __launch_bounds__(64) k_legendre(SphArg a, int n_items, const float* __restrict__ wplm, const float4* __restrict__ spectra,
                                                 float2* __restrict__ alm) {
    const int lane = threadIdx.x;
    const int half = lane >> 5;
    const int r32 = lane & 31;
    const int m = blockIdx.y;
    const int b0 = blockIdx.x * 16;
    const int nh = a.n >> 1;
    const int L = a.L;
    const int base = m * L - (m * (m - 1)) / 2 - m;  // pair(m, l) = base + l
    // this lane's A row: item r32 / 2, part r32 % 2
    const int item = b0 + (r32 >> 1);
    const bool im = (r32 & 1) != 0;
    const bool iok = item < n_items;
    const float4* sp = spectra + ((size_t)m * n_items + (iok ? item : 0)) * nh;
    const int ncol = (L - m + 1) / 2;  // degrees l = m, m + 2, ...
    for (int ct = 0; ct * 32 < ncol; ++ct) {
        const int le = m + 2 * (ct * 32 + r32);  // this lane's column: even l + m
        const int lo = le + 1;                   // and odd
        floatx16 ce, co;
#pragma unroll
        for (int r = 0; r < 16; ++r) ce[r] = co[r] = 0.f;
        for (int ks = 0; ks < nh / 2; ++ks) {
            const int y = 2 * ks + half;
            const float4 v = iok ? sp[y] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float* wy = wplm + (size_t)y * a.pairs + base;
            const float be = le < L ? wy[le] : 0.f;
            const float bo = lo < L ? wy[lo] : 0.f;
            ce = __builtin_amdgcn_mfma_f32_32x32x2f32(im ? v.y : v.x, be, ce, 0, 0, 0);
            co = __builtin_amdgcn_mfma_f32_32x32x2f32(im ? v.w : v.z, bo, co, 0, 0, 0);
        }
        // rows (r / 4) 8 + half 4 + r % 4: registers r, r + 1 (r even) are (re, im) of one item
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int row = (r >> 2) * 8 + half * 4 + (r & 3);
            const int it = b0 + (row >> 1);
            if (it < n_items) {
                float2* out = alm + (size_t)it * a.pairs + base;
                if (le < L) out[le] = make_float2(ce[r], ce[r + 1]);
                if (lo < L) out[lo] = make_float2(co[r], co[r + 1]);
            }
        }
    }
}

// The same sums with 16-bit products (FP16, or BF16 with BF) and FP32 accumulation, on
// v_mfma_f32_32x32x16_{f16,bf16}: 16 rings a product (A lane -> (row lane % 32, rings
// (lane / 32) 8 .. + 8), B lane -> (those rings, column lane % 32); the output layout is the
// 32 x 32 one above). Each value x is carried as hi = round16(x) and lo = round16(x - hi); the
// column tiles holding a degree >= split take hi hi + hi lo + lo hi (an error near FP32's), the
// others hi hi. FP16's exponent range is narrow: the Legendre table's values (below ~1e-3 at
// L = 96) and their lo parts would fall into its subnormals, so with FP16 the table is scaled by
// 2^12 on the way in and the sums by 2^-12 on the way out (exact: powers of two).
template <bool BF>
struct Leg16;
template <>
struct Leg16<false> {
    typedef _Float16 T;
    typedef _Float16 V8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ floatx16 mfma(V8 a, V8 b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
};
template <>
struct Leg16<true> {
    typedef __bf16 T;
    typedef __bf16 V8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ floatx16 mfma(V8 a, V8 b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};

template <bool BF>
__global__ void __launch_bounds__(64) k_legendre16(SphArg a, int n_items, int split, const float* __restrict__ wplm,
                                                   const float4* __restrict__ spectra, float2* __restrict__ alm) {
    typedef typename Leg16<BF>::T T;
    typedef typename Leg16<BF>::V8 V8;
    const int lane = threadIdx.x;
    const int half = lane >> 5;
    const int r32 = lane & 31;
    const int m = blockIdx.y;
    const int b0 = blockIdx.x * 16;
    const int nh = a.n >> 1;
    const int L = a.L;
    const int base = m * L - (m * (m - 1)) / 2 - m;
    const int item = b0 + (r32 >> 1);
    const bool im = (r32 & 1) != 0;
    const bool iok = item < n_items;
    const float4* sp = spectra + ((size_t)m * n_items + (iok ? item : 0)) * nh;
    const int ncol = (L - m + 1) / 2;
    constexpr float kScale = BF ? 1.f : 4096.f;
    for (int ct = 0; ct * 32 < ncol; ++ct) {
        const int le = m + 2 * (ct * 32 + r32);
        const int lo = le + 1;
        // the tile's highest degree decides the split (uniform over the wave)
        const bool two = m + 2 * (ct * 32 + 31) + 1 >= split;
        floatx16 ce, co;
#pragma unroll
        for (int r = 0; r < 16; ++r) ce[r] = co[r] = 0.f;
        for (int ks = 0; ks < nh / 16; ++ks) {
            V8 aeh, aol, ael, aoh, beh, bel, boh, bol;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int y = 16 * ks + 8 * half + i;
                const float4 v = iok ? sp[y] : make_float4(0.f, 0.f, 0.f, 0.f);
                const float xe = im ? v.y : v.x, xo = im ? v.w : v.z;
                const float* wy = wplm + (size_t)y * a.pairs + base;
                const float we = le < L ? kScale * wy[le] : 0.f, wo = lo < L ? kScale * wy[lo] : 0.f;
                const T h0 = (T)xe, h1 = (T)xo, h2 = (T)we, h3 = (T)wo;
                aeh[i] = h0;
                aoh[i] = h1;
                beh[i] = h2;
                boh[i] = h3;
                ael[i] = (T)(xe - (float)h0);
                aol[i] = (T)(xo - (float)h1);
                bel[i] = (T)(we - (float)h2);
                bol[i] = (T)(wo - (float)h3);
            }
            ce = Leg16<BF>::mfma(aeh, beh, ce);
            co = Leg16<BF>::mfma(aoh, boh, co);
            if (two) {
                ce = Leg16<BF>::mfma(aeh, bel, ce);
                ce = Leg16<BF>::mfma(ael, beh, ce);
                co = Leg16<BF>::mfma(aoh, bol, co);
                co = Leg16<BF>::mfma(aol, boh, co);
            }
        }
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int row = (r >> 2) * 8 + half * 4 + (r & 3);
            const int it = b0 + (row >> 1);
            if (it < n_items) {
                float2* out = alm + (size_t)it * a.pairs + base;
                constexpr float inv = 1.f / kScale;
                if (le < L) out[le] = make_float2(ce[r] * inv, ce[r + 1] * inv);
                if (lo < L) out[lo] = make_float2(co[r] * inv, co[r + 1] * inv);
            }
        }
    }
}

// Unit norm sqrt(sum_m k |a|^2), k = 1 for m = 0 else 2, and the stats (sphere.wgsl:228-262).
// PER > 0: a thread holds its PER coefficients (PER * 256 >= pairs) in registers between the norm
// and the scaling; 0: any size, the coefficients read again.
template <int PER>
__global__ void __launch_bounds__(256) k_norm(SphArg a, const float2* __restrict__ sums, float2* __restrict__ alm, float* __restrict__ stats) {
    __shared__ float red[4];
    const int b = blockIdx.x;
    float2* c = alm + (size_t)b * a.pairs;
    float2 v[PER > 0 ? PER : 1];
    float e = 0.f;
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int p = threadIdx.x + 256 * k;
            v[k] = p < a.pairs ? c[p] : make_float2(0.f, 0.f);
            e += (p < a.L ? 1.f : 2.f) * (v[k].x * v[k].x + v[k].y * v[k].y);
        }
    } else {
        for (int p = threadIdx.x; p < a.pairs; p += 256) {
            const float2 x = c[p];
            e += (p < a.L ? 1.f : 2.f) * (x.x * x.x + x.y * x.y);
        }
    }
    e = warp_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    const float tot = red[0] + red[1] + red[2] + red[3];
    const float nrm = sqrtf(fmaxf(tot, 1e-30f));
    const bool scale = (a.flags & ZB_FLAG_NORM) != 0;
    const float inv = 1.f / nrm;
    float c0 = 0.f;  // the (0, 0) coefficient as stored
    if constexpr (PER > 0) {
        if (scale) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int p = threadIdx.x + 256 * k;
                v[k] = make_float2(v[k].x * inv, v[k].y * inv);
                if (p < a.pairs) c[p] = v[k];
            }
        }
        c0 = v[0].x;
    } else {
        if (scale)
            for (int p = threadIdx.x; p < a.pairs; p += 256) c[p] = make_float2(c[p].x * inv, c[p].y * inv);
        __syncthreads();
        c0 = c[0].x;
    }
    if (threadIdx.x == 0) {
        const float2 t = sums[b];
        stats[4 * b] = t.y * ZB_TAU / (float)(1 << a.log2nphi);
        stats[4 * b + 1] = window_mean(a, t);
        stats[4 * b + 2] = nrm;
        stats[4 * b + 3] = c0;
    }
}

}  // namespace zb

using namespace zb;

static int64_t sh_ws(int L, int n, int n_items) {
    const int64_t nh = n / 2;
    int64_t sums = ((int64_t)n_items * 8 + 255) / 256 * 256;
    return sums + (int64_t)n_items * L * nh * 16;
}

extern "C" int64_t zb_sh_workspace(const zb_band* band, int32_t n_items) {
    if (!band || band->L < 2 || n_items < 0) return ZB_EINVAL;
    return sh_ws(band->L, band->n, n_items);
}

static int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return (1 << r) == v ? r : -1;
}

// The arguments zb_sh_coefficients and zb_sh_coefficients_noalloc share, checked; 1 when there
// is nothing to do.
static int sh_check(const char* who, const zb_band* band, const zb_frame* frame, const float* pats, const float* items, int32_t n_items,
                    int32_t flags, float* alm, float* stats) {
    if (!band || !frame || n_items < 0 || band->L < 2 || ilog2(band->n) < 4 || band->n < 2 * band->L - 1 || 2 * band->n > ZB_TW ||
        frame->pat_h < 2 || frame->pat_w < 2) {
        set_error("%s: bad band or frame", who);
        return ZB_EINVAL;
    }
    if ((flags & ZB_FLAG_FP16) && (flags & ZB_FLAG_BF16)) {
        set_error("%s: FP16 and BF16 products both asked for", who);
        return ZB_EINVAL;
    }
    if (n_items == 0) return 1;
    if (!items || !alm || !stats || !band->ring || !band->wplm || !band->twiddle || (!(flags & ZB_FLAG_WINDOW) && !pats)) {
        set_error("%s: null pointer", who);
        return ZB_EINVAL;
    }
    return ZB_OK;
}

// The four launches over a workspace of sh_ws bytes at `work` (a device address).
static int32_t sh_run(const zb_band* band, const zb_frame* frame, const float* pats, const float* items, int32_t n_items, int32_t flags,
                      char* work, float* alm, float* stats, hipStream_t s);

extern "C" int32_t zb_sh_coefficients(const zb_band* band, const zb_frame* frame, const float* pats, const float* items,
                                      int32_t n_items, int32_t flags, void* work, int64_t work_bytes, float* alm, float* stats,
                                      void* stream) {
    const int rc = sh_check("zb_sh_coefficients", band, frame, pats, items, n_items, flags, alm, stats);
    if (rc) return rc < 0 ? rc : ZB_OK;
    if (!work) {
        set_error("zb_sh_coefficients: null pointer");
        return ZB_EINVAL;
    }
    if (work_bytes < sh_ws(band->L, band->n, n_items)) {
        set_error("zb_sh_coefficients: workspace %lld bytes, need %lld", (long long)work_bytes,
                  (long long)sh_ws(band->L, band->n, n_items));
        return ZB_ESPACE;
    }
    return sh_run(band, frame, pats, items, n_items, flags, (char*)work, alm, stats, (hipStream_t)stream);
}

// zb_sh_coefficients_noalloc's workspace: one array per device for the life of the process.
// Host code never passes `scratch` itself to a kernel (on the host it names a stand-in, not the
// device copy): it uses the device address hipGetSymbolAddress returns, kept per device.
static __device__ __align__(256) char scratch[ZB_SH_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[ZB_MAX_DEVICES];

extern "C" int32_t zb_sh_coefficients_noalloc(const zb_band* band, const zb_frame* frame, const float* pats, const float* items,
                                              int32_t n_items, int32_t flags, float* alm, float* stats, void* stream) {
    const int rc = sh_check("zb_sh_coefficients_noalloc", band, frame, pats, items, n_items, flags, alm, stats);
    if (rc) return rc < 0 ? rc : ZB_OK;
    const int64_t need = sh_ws(band->L, band->n, n_items);
    if (need > (int64_t)ZB_SH_SCRATCH_BYTES) {
        int64_t most = ZB_SH_SCRATCH_BYTES / ((int64_t)band->L * (band->n / 2) * 16 + 8);
        while (most > 0 && sh_ws(band->L, band->n, (int)most) > (int64_t)ZB_SH_SCRATCH_BYTES) --most;
        set_error("zb_sh_coefficients_noalloc: %d items need %lld bytes of scratch, the library holds %lld (at most %lld items at L = %d)",
                  n_items, (long long)need, (long long)ZB_SH_SCRATCH_BYTES, (long long)most, band->L);
        return ZB_ESPACE;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= ZB_MAX_DEVICES) {
        set_error("zb_sh_coefficients_noalloc: no current device (or device %d past %d)", dev, ZB_MAX_DEVICES);
        return ZB_EINVAL;
    }
    void* ptr = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                set_error("zb_sh_coefficients_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
                return ZB_ELAUNCH;
            }
            g_scratch_ptr[dev] = p;
        }
        ptr = g_scratch_ptr[dev];
    }
    return sh_run(band, frame, pats, items, n_items, flags, (char*)ptr, alm, stats, (hipStream_t)stream);
}

static int32_t sh_run(const zb_band* band, const zb_frame* frame, const float* pats, const float* items, int32_t n_items, int32_t flags,
                      char* work, float* alm, float* stats, hipStream_t s) {
    const int L = band->L, n = band->n;
    SphArg a;
    for (int i = 0; i < 3; ++i) {
        a.c0[i] = frame->c0[i];
        a.c1[i] = frame->c1[i];
        a.c2[i] = frame->c2[i];
        a.mask[i] = frame->mask[i];
    }
    a.h = frame->pat_h;
    a.w = frame->pat_w;
    a.L = L;
    a.n = n;
    a.log2nphi = ilog2(2 * n);
    a.pairs = L * (L + 1) / 2;
    a.flags = flags;
    char* ws = work;
    float2* sums = (float2*)ws;
    float4* spectra = (float4*)(ws + ((int64_t)n_items * 8 + 255) / 256 * 256);
    const float2* tw = (const float2*)band->twiddle;
    const float4* its = (const float4*)items;
    ZB_LAUNCH("k_sums", k_sums, dim3(n_items), dim3(256), 0, s, a, band->ring, tw, pats, its, sums);
    if (int rc = check_launch("zb_sh_coefficients/sums")) return rc;
    const int nh = n / 2;
    const int nphi = 2 * n;
    // ring pairs a block transforms: the knob's (a power of two), else 16 KB of LDS worth; at
    // most n / 2 and 64 KB of LDS
    int G = knob(KNOB_RINGS_GROUP) > 0 ? knob(KNOB_RINGS_GROUP) : 2048 / nphi;
    while (G & (G - 1)) G &= G - 1;
    if (G > 8192 / nphi) G = 8192 / nphi;
    if (G < 1) G = 1;
    if (G > nh) G = nh;
    const dim3 rgrid(n_items * (nh / G));
    const size_t rlds = (size_t)G * nphi * 8 + (size_t)G * 16;
    if (a.log2nphi == 7)
        ZB_LAUNCH("k_rings", k_rings<7>, rgrid, dim3(256), rlds, s, a, G, band->ring, tw, pats, its, (const float2*)sums, spectra);
    else if (a.log2nphi == 8)
        ZB_LAUNCH("k_rings", k_rings<8>, rgrid, dim3(256), rlds, s, a, G, band->ring, tw, pats, its, (const float2*)sums, spectra);
    else if (a.log2nphi == 9)
        ZB_LAUNCH("k_rings", k_rings<9>, rgrid, dim3(256), rlds, s, a, G, band->ring, tw, pats, its, (const float2*)sums, spectra);
    else
        ZB_LAUNCH("k_rings", k_rings<0>, rgrid, dim3(256), rlds, s, a, G, band->ring, tw, pats, its, (const float2*)sums, spectra);
    if (int rc = check_launch("zb_sh_coefficients/rings")) return rc;
    if ((flags & (ZB_FLAG_FP16 | ZB_FLAG_BF16)) && nh >= 16) {
        // 16-bit products, 16 rings a product (below n = 32 the FP32 sums stand in)
        const int split = (flags >> ZB_FLAG_SPLIT_SHIFT) & 0xff;
        if (flags & ZB_FLAG_BF16)
            ZB_LAUNCH("k_legendre", k_legendre16<true>, dim3((n_items + 15) / 16, L), dim3(64), 0, s, a, n_items, split, band->wplm,
                      (const float4*)spectra, (float2*)alm);
        else
            ZB_LAUNCH("k_legendre", k_legendre16<false>, dim3((n_items + 15) / 16, L), dim3(64), 0, s, a, n_items, split, band->wplm,
                      (const float4*)spectra, (float2*)alm);
    } else {
        ZB_LAUNCH("k_legendre", k_legendre, dim3((n_items + 15) / 16, L), dim3(64), 0, s, a, n_items, band->wplm, (const float4*)spectra,
                  (float2*)alm);
    }
    if (int rc = check_launch("zb_sh_coefficients/legendre")) return rc;
    if (a.pairs <= 256 * 10)
        ZB_LAUNCH("k_norm", k_norm<10>, dim3(n_items), dim3(256), 0, s, a, (const float2*)sums, (float2*)alm, stats);
    else if (a.pairs <= 256 * 20)
        ZB_LAUNCH("k_norm", k_norm<20>, dim3(n_items), dim3(256), 0, s, a, (const float2*)sums, (float2*)alm, stats);
    else
        ZB_LAUNCH("k_norm", k_norm<0>, dim3(n_items), dim3(256), 0, s, a, (const float2*)sums, (float2*)alm, stats);
    return check_launch("zb_sh_coefficients/norm");
}
