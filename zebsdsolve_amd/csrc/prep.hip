// Patterns in, masters baked: per-pattern normalization under the mask, simulated patterns,
// and the master's smoothing and gradient on the sphere (setup, once per master and geometry).
#include <algorithm>

#include "common.h"

namespace zb {

template <int NT>
__device__ float block_sum(float v, float* red) {
    v = warp_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    float t = 0.f;
    for (int i = 0; i < NT / 64; ++i) t += red[i];
    __syncthreads();
    return t;
}

__device__ __forceinline__ bool in_mask(float row, float col, const float* m) {
    float dr = row - m[0], dc = col - m[1];
    return m[2] <= 0.f || dr * dr + dc * dc <= m[2];
}

struct MaskArg { float m[3]; };

// VEC consecutive pixels as one load (VEC 4: 16 bytes of float32, 8 of u16, 4 of u8).
template <typename T, int VEC>
struct PixVec;
template <typename T>
struct PixVec<T, 1> {
    static __device__ __forceinline__ void load(const T* p, float* x) { x[0] = (float)p[0]; }
};
template <>
struct PixVec<float, 4> {
    static __device__ __forceinline__ void load(const float* p, float* x) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        x[0] = t.x, x[1] = t.y, x[2] = t.z, x[3] = t.w;
    }
};
template <>
struct PixVec<uint16_t, 4> {
    static __device__ __forceinline__ void load(const uint16_t* p, float* x) {
        const ushort4 t = *reinterpret_cast<const ushort4*>(p);
        x[0] = (float)t.x, x[1] = (float)t.y, x[2] = (float)t.z, x[3] = (float)t.w;
    }
};
template <>
struct PixVec<uint8_t, 4> {
    static __device__ __forceinline__ void load(const uint8_t* p, float* x) {
        const uchar4 t = *reinterpret_cast<const uchar4*>(p);
        x[0] = (float)t.x, x[1] = (float)t.y, x[2] = (float)t.z, x[3] = (float)t.w;
    }
};

// One block of 256 per pattern: mean and norm under the mask (prep.wgsl:405-443). The pattern
// is read once into registers (PER loads of VEC consecutive pixels a thread, PER * VEC * 256 >=
// h w; with VEC 4, h w a multiple of 4 and the rows 4-pixel aligned), written once.
template <typename T, int PER, int VEC>
__global__ void __launch_bounds__(256) k_normalize(const T* __restrict__ raw, int h, int w, MaskArg mask, float* __restrict__ out) {
    static_assert(PER * VEC <= 64, "one mask bit a pixel in 64 bits");
    __shared__ float red[4];
    const int b = blockIdx.x;
    const int hw = h * w;
    const T* src = raw + (size_t)b * hw;
    float* dst = out + (size_t)b * hw;
    const float inv_w = 1.f / (float)w;
    float v[PER * VEC];
    unsigned long long in = 0ull;  // bit k VEC + u: pixel u of load k is under the mask
    float s = 0.f, c = 0.f;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i0 = VEC * (threadIdx.x + 256 * k);
        float x[VEC];
#pragma unroll
        for (int u = 0; u < VEC; ++u) x[u] = 0.f;
        if (i0 < hw) PixVec<T, VEC>::load(src + i0, x);
#pragma unroll
        for (int u = 0; u < VEC; ++u) {
            const int i = i0 + u;
            v[k * VEC + u] = 0.f;
            // row and column without an integer division (exact for h w <= 65536)
            const int row = (int)(((float)i + 0.5f) * inv_w);
            if (i < hw && in_mask((float)row, (float)(i - row * w), mask.m)) {
                v[k * VEC + u] = x[u];
                in |= 1ull << (k * VEC + u);
                s += x[u];
                c += 1.f;
            }
        }
    }
    s = block_sum<256>(s, red);
    c = block_sum<256>(c, red);
    const float mean = s / fmaxf(c, 1.f);
    float ss = 0.f;
#pragma unroll
    for (int k = 0; k < PER * VEC; ++k) {
        v[k] = ((in >> k) & 1ull) ? v[k] - mean : 0.f;
        ss += v[k] * v[k];
    }
    ss = block_sum<256>(ss, red);
    const float inv = 1.f / sqrtf(fmaxf(ss, 1e-30f));
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i0 = VEC * (threadIdx.x + 256 * k);
        if (i0 < hw) {
            if constexpr (VEC == 4)
                *reinterpret_cast<float4*>(dst + i0) = make_float4(v[4 * k] * inv, v[4 * k + 1] * inv, v[4 * k + 2] * inv, v[4 * k + 3] * inv);
            else
                dst[i0] = v[k] * inv;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_normalize_big(const T* __restrict__ raw, int h, int w, MaskArg mask, float* __restrict__ out) {
    __shared__ float red[4];
    const int b = blockIdx.x;
    const int hw = h * w;
    const T* src = raw + (size_t)b * hw;
    float* dst = out + (size_t)b * hw;
    float s = 0.f, c = 0.f;
    for (int i = threadIdx.x; i < hw; i += 256) {
        if (in_mask((float)(i / w), (float)(i % w), mask.m)) {
            s += (float)src[i];
            c += 1.f;
        }
    }
    s = block_sum<256>(s, red);
    c = block_sum<256>(c, red);
    const float mean = s / fmaxf(c, 1.f);
    float ss = 0.f;
    for (int i = threadIdx.x; i < hw; i += 256) {
        float v = 0.f;
        if (in_mask((float)(i / w), (float)(i % w), mask.m)) v = (float)src[i] - mean;
        dst[i] = v;
        ss += v * v;
    }
    ss = block_sum<256>(ss, red);
    const float inv = 1.f / sqrtf(fmaxf(ss, 1e-30f));
    for (int i = threadIdx.x; i < hw; i += 256) dst[i] *= inv;
}

// ── the pattern filters, fused with the normalization (prep.wgsl:57-287, 405-443; src/prep.zig:
// 105-204) ──
//
// One block per pattern with the image in LDS: the static background subtracted as it is
// loaded, the dynamic background (a Gaussian blur of the pattern, subtracted or divided by),
// then zero mean and unit norm under the mask; HBM sees the raw pattern once and the result
// once, where the reference's chain passes every image through memory at each step. The
// Gaussian of deviation sigma (pixels) is computed as the reference computes it: below 12 pixels
// directly and separably, taps exp(-d^2 / 2 sigma^2) out to ceil(3 sigma) renormalized to the
// taps inside the image; from 12 pixels from sums over boxes of s = floor(sigma / 6) pixels,
// blurred on the coarse grid (deviation sqrt(sigma^2 - s^2 - (s^2 - 1) / 12), out to 4 of them,
// with a margin of 3 whole cells) and spread back by a Gaussian of deviation s; sums of pixels
// and of counts travel together and are divided at the end.
struct BlurArg {
    int box;       // 0: direct taps, 1: the box path
    int radius;    // direct: taps each way
    float k;       // direct: 1 / (2 sigma^2)
    int s, pad;    // box: cell size in pixels, margin in cells
    float reach;   // box: ceil(4 sigma_coarse), pixels
    float k0, k2;  // box: 1 / (2 sigma_coarse^2), 1 / (2 s^2)
    int hs, ws, hp, wp;
};

struct PrepArg {
    int h, w;
    float mask[3];
    int dyn;  // 0 none, 1 subtract, 2 divide
    BlurArg dynb;
    // CLAHE: tiles x tiles, bins levels, clip; the image padded by pad_h, pad_w to tiles of
    // tile_h x tile_w
    int clahe, tiles, bins;
    float clip;
    int pad_h, pad_w, tile_h, tile_w;
    // band pass: the fine Gaussian (or the image itself) minus the coarse one
    int bp, bp_fine;
    BlurArg fine, coarse;
};

// centre (pixels) and width of coarse cell j along an axis of n pixels (cells outside are whole)
__device__ __forceinline__ float2 prep_cell(int j, int n, int s) {
    const int lo = j * s;
    int hi = lo + s;
    if (j >= 0 && lo < n) hi = min(hi, n);
    return make_float2(0.5f * (float)(lo + hi - 1), (float)(hi - lo));
}

// The Gaussian of the h x w image `img` (LDS, read only), through the scratch `aux` (LDS); the
// last pass hands each pixel's blurred value to emit(i, value), pixel i on thread i mod 256.
template <typename Emit>
__device__ void gauss(const float* img, float* aux, int h, int w, const BlurArg& g, Emit emit) {
    const int hw = h * w;
    const int tid = threadIdx.x;
    if (!g.box) {
        float* wt = aux + ((hw + 3) & ~3);  // exp(-d^2 k), d <= radius
        for (int d = tid; d <= g.radius; d += 256) wt[d] = expf(-(float)(d * d) * g.k);
        __syncthreads();
        for (int i = tid; i < hw; i += 256) {
            const int r = i / w, c = i - (i / w) * w;
            float acc = 0.f, ws = 0.f;
            const int x1 = min(c + g.radius, w - 1);
            for (int x = max(c - g.radius, 0); x <= x1; ++x) {
                const float q = wt[abs(x - c)];
                acc += q * img[r * w + x];
                ws += q;
            }
            aux[i] = acc / ws;
        }
        __syncthreads();
        for (int i = tid; i < hw; i += 256) {
            const int r = i / w, c = i - (i / w) * w;
            float acc = 0.f, ws = 0.f;
            const int y1 = min(r + g.radius, h - 1);
            for (int y = max(r - g.radius, 0); y <= y1; ++y) {
                const float q = wt[abs(y - r)];
                acc += q * aux[y * w + c];
                ws += q;
            }
            emit(i, acc / ws);
        }
        __syncthreads();
        return;
    }
    const int s = g.s, pad = g.pad;
    const int big = g.hp * (g.wp > w ? g.wp : w);
    float2* c0 = reinterpret_cast<float2*>(aux);
    float2* c1 = c0 + big;
    // pixel sums and counts over the s x s boxes: [hs][ws]
    for (int t = tid; t < g.hs * g.ws; t += 256) {
        const int i = t / g.ws, j = t - (t / g.ws) * g.ws;
        const int r1 = min((i + 1) * s, h), q1 = min((j + 1) * s, w);
        float acc = 0.f;
        for (int r = i * s; r < r1; ++r)
            for (int q = j * s; q < q1; ++q) acc += img[r * w + q];
        c0[t] = make_float2(acc, (float)((r1 - i * s) * (q1 - j * s)));
    }
    __syncthreads();
    // along the rows onto the coarse grid with its margin: [hs][wp]
    for (int t = tid; t < g.hs * g.wp; t += 256) {
        const int line = t / g.wp, tc = t - (t / g.wp) * g.wp;
        const float cc = prep_cell(tc - pad, w, s).x;
        const int j0 = max((int)floorf((cc - g.reach) / (float)s), 0);
        const int j1 = min((int)floorf((cc + g.reach) / (float)s), g.ws - 1);
        float2 acc = make_float2(0.f, 0.f);
        for (int j = j0; j <= j1; ++j) {
            const float d = prep_cell(j, w, s).x - cc;
            const float q = expf(-d * d * g.k0);
            const float2 v = c0[line * g.ws + j];
            acc = make_float2(acc.x + q * v.x, acc.y + q * v.y);
        }
        c1[t] = acc;
    }
    __syncthreads();
    // along the columns: [hp][wp]
    for (int t = tid; t < g.hp * g.wp; t += 256) {
        const int tr = t / g.wp, col = t - (t / g.wp) * g.wp;
        const float cc = prep_cell(tr - pad, h, s).x;
        const int i0 = max((int)floorf((cc - g.reach) / (float)s), 0);
        const int i1 = min((int)floorf((cc + g.reach) / (float)s), g.hs - 1);
        float2 acc = make_float2(0.f, 0.f);
        for (int i = i0; i <= i1; ++i) {
            const float d = prep_cell(i, h, s).x - cc;
            const float q = expf(-d * d * g.k0);
            const float2 v = c1[i * g.wp + col];
            acc = make_float2(acc.x + q * v.x, acc.y + q * v.y);
        }
        c0[t] = acc;
    }
    __syncthreads();
    // spread over the pixels of each row from the 2 pad + 1 cells around: [hp][w]
    for (int t = tid; t < g.hp * w; t += 256) {
        const int tr = t / w, x = t - (t / w) * w;
        const int own = x / s;
        float2 acc = make_float2(0.f, 0.f);
        for (int q = 0; q <= 2 * pad; ++q) {
            const float2 cw = prep_cell(own + q - pad, w, s);
            const float d = (float)x - cw.x;
            const float e = cw.y * expf(-d * d * g.k2);
            const float2 v = c0[tr * g.wp + own + q];
            acc = make_float2(acc.x + e * v.x, acc.y + e * v.y);
        }
        c1[t] = acc;
    }
    __syncthreads();
    // and of each column, into the pixels
    for (int i = tid; i < hw; i += 256) {
        const int y = i / w, x = i - (i / w) * w;
        const int own = y / s;
        float2 acc = make_float2(0.f, 0.f);
        for (int q = 0; q <= 2 * pad; ++q) {
            const float2 cw = prep_cell(own + q - pad, h, s);
            const float d = (float)y - cw.x;
            const float e = cw.y * expf(-d * d * g.k2);
            const float2 v = c1[(own + q) * w + x];
            acc = make_float2(acc.x + e * v.x, acc.y + e * v.y);
        }
        emit(i, acc.x / acc.y);
    }
    __syncthreads();
}

__device__ __forceinline__ int clahe_reflect(int i, int n, int pad) {
    if (n <= 1) return 0;
    int x = abs(i - pad);
    if (x >= n) x = 2 * n - 2 - x;
    return min(max(x, 0), n - 1);
}

// One tile's table from its histogram (prep.wgsl:343-364), in place (each level read before it
// is written): clipped at clip x the mean count, what was cut spread evenly (the remainder one
// count each to the first levels), accumulated and scaled by the tile's pixel count.
__device__ void clahe_table(unsigned* hist, int B, float voxels, float clip) {
    float* cdf = reinterpret_cast<float*>(hist);
    const float nb = (float)B;
    const float limit = fmaxf(clip * voxels / nb, 1.f);
    float sumc = 0.f;
    for (int i = 0; i < B; ++i) sumc += fminf((float)hist[i], limit);
    const float clipped = voxels - sumc;
    const int residual = (int)(clipped - nb * floorf(clipped / nb));
    const float redist = (clipped - (float)residual) / nb;
    float acc = 0.f;
    for (int i = 0; i < B; ++i) {
        float c = fminf((float)hist[i], limit) + redist;
        if (i < residual) c += 1.f;
        acc += c;
        cdf[i] = fminf(fmaxf(acc / voxels, 0.f), 1.f);
    }
}

// The equalized value of pixel (row, col) holding v (prep.wgsl:366-397): the four nearest
// tiles' tables [T][T][B], each read between its two nearest levels.
__device__ float clahe_value(float v, int row, int col, int h, int w, int T, int B, const float* cdf) {
    auto mix = [](float x, float y, float t) { return x * (1.f - t) + y * t; };
    v = fminf(fmaxf(v, 0.f), 1.f);
    const float gy = ((float)row + 0.5f) / (float)h * (float)(T - 1);
    const float gx = ((float)col + 0.5f) / (float)w * (float)(T - 1);
    const int y0 = (int)floorf(gy), x0 = (int)floorf(gx);
    const int y1 = min(y0 + 1, T - 1), x1 = min(x0 + 1, T - 1);
    const float fy = gy - (float)y0, fx = gx - (float)x0;
    const float bf = v * (float)(B - 1);
    const int b0 = (int)fminf(fmaxf(floorf(bf), 0.f), (float)(B - 2));
    const float fb = bf - (float)b0;
    auto at = [&](int ty, int tx) {
        const float* c = cdf + (ty * T + tx) * B;
        return mix(c[b0], c[b0 + 1], fb);
    };
    return mix(mix(at(y0, x0), at(y0, x1), fx), mix(at(y1, x0), at(y1, x1), fx), fy);
}

__device__ __forceinline__ int clahe_bin(float v, int B) { return (int)fminf(fmaxf(fminf(fmaxf(v, 0.f), 1.f) * (float)(B - 1), 0.f), (float)(B - 1)); }

// CLAHE of the h x w image in LDS, in place (prep.wgsl:289-397): stretched to [0, 1] by its
// extremes; padded by reflection at its top and left to whole tiles; each tile's histogram
// (LDS atomics, `hist` tiles^2 bins words) turned into its table; each pixel equalized.
__device__ void clahe_block(float* img, unsigned* hist, const PrepArg& a, float* red) {
    const int h = a.h, w = a.w, hw = h * w;
    const int tid = threadIdx.x;
    const int T = a.tiles, B = a.bins;
    float lo = 3.4e38f, hi = -3.4e38f;
    for (int i = tid; i < hw; i += 256) {
        lo = fminf(lo, img[i]);
        hi = fmaxf(hi, img[i]);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    if ((tid & 63) == 0) {
        red[tid >> 6] = lo;
        red[4 + (tid >> 6)] = hi;
    }
    for (int i = tid; i < T * T * B; i += 256) hist[i] = 0u;
    __syncthreads();
    lo = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    hi = fmaxf(fmaxf(red[4], red[5]), fmaxf(red[6], red[7]));
    const float k = 1.f / fmaxf(hi - lo, 1e-20f);
    for (int i = tid; i < hw; i += 256) img[i] = (img[i] - lo) * k;
    __syncthreads();
    const int hp = h + a.pad_h, wp = w + a.pad_w;
    for (int t = tid; t < hp * wp; t += 256) {
        const int py = t / wp, px = t - (t / wp) * wp;
        const int bin = clahe_bin(img[clahe_reflect(py, h, a.pad_h) * w + clahe_reflect(px, w, a.pad_w)], B);
        const int tile = (py / a.tile_h) * T + px / a.tile_w;
        atomicAdd(&hist[tile * B + bin], 1u);
    }
    __syncthreads();
    for (int t = tid; t < T * T; t += 256) clahe_table(hist + t * B, B, (float)max(a.tile_h * a.tile_w, 1), a.clip);
    __syncthreads();
    const float* cdf = reinterpret_cast<const float*>(hist);
    for (int i = tid; i < hw; i += 256) {
        const int row = i / w, col = i - (i / w) * w;
        img[i] = clahe_value(img[i], row, col, h, w, T, B, cdf);
    }
    __syncthreads();
}

// CLAHE of images too large for one block's LDS (a master's sheets), in three launches over
// global memory: the extremes and the stretch, one block an image; the tables, one block a
// (image, tile) with its histogram in LDS; the equalized pixels.
struct ClaheArg {
    int n, h, w, tiles, bins;
    float clip;
    int pad_h, pad_w, tile_h, tile_w;
};

__global__ void __launch_bounds__(256) k_clahe_stretch(float* __restrict__ img, ClaheArg a) {
    __shared__ float red[8];
    float* x = img + (size_t)blockIdx.x * a.h * a.w;
    const int hw = a.h * a.w, tid = threadIdx.x;
    float lo = 3.4e38f, hi = -3.4e38f;
    for (int i = tid; i < hw; i += 256) {
        lo = fminf(lo, x[i]);
        hi = fmaxf(hi, x[i]);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    if ((tid & 63) == 0) {
        red[tid >> 6] = lo;
        red[4 + (tid >> 6)] = hi;
    }
    __syncthreads();
    lo = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    hi = fmaxf(fmaxf(red[4], red[5]), fmaxf(red[6], red[7]));
    const float k = 1.f / fmaxf(hi - lo, 1e-20f);
    for (int i = tid; i < hw; i += 256) x[i] = (x[i] - lo) * k;
}

__global__ void __launch_bounds__(256) k_clahe_tables(const float* __restrict__ img, ClaheArg a, float* __restrict__ cdf) {
    __shared__ unsigned hist[256];
    const int T = a.tiles, B = a.bins;
    const int b = blockIdx.x / (T * T), tile = blockIdx.x - b * (T * T);
    const int ty = tile / T, tx = tile - ty * T;
    const float* x = img + (size_t)b * a.h * a.w;
    for (int i = threadIdx.x; i < B; i += 256) hist[i] = 0u;
    __syncthreads();
    for (int t = threadIdx.x; t < a.tile_h * a.tile_w; t += 256) {
        const int py = ty * a.tile_h + t / a.tile_w, px = tx * a.tile_w + t % a.tile_w;
        atomicAdd(&hist[clahe_bin(x[clahe_reflect(py, a.h, a.pad_h) * a.w + clahe_reflect(px, a.w, a.pad_w)], B)], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) clahe_table(hist, B, (float)max(a.tile_h * a.tile_w, 1), a.clip);
    __syncthreads();
    for (int i = threadIdx.x; i < B; i += 256) cdf[(size_t)blockIdx.x * B + i] = reinterpret_cast<const float*>(hist)[i];
}

__global__ void __launch_bounds__(256) k_clahe_apply(float* __restrict__ img, ClaheArg a, const float* __restrict__ cdf) {
    const long tid = (long)blockIdx.x * 256 + threadIdx.x;
    const long hw = (long)a.h * a.w;
    if (tid >= (long)a.n * hw) return;
    const int b = (int)(tid / hw);
    const int pix = (int)(tid - (long)b * hw);
    img[tid] = clahe_value(img[tid], pix / a.w, pix % a.w, a.h, a.w, a.tiles, a.bins, cdf + (size_t)b * a.tiles * a.tiles * a.bins);
}

// The filter chain and the normalization of one pattern, the block's 256 threads over the
// image `img` and the scratch `aux` (LDS, or a block's slot of global memory for images past
// what LDS holds): the static background subtracted on the way in, the dynamic background,
// CLAHE, the band pass (the coarse Gaussian parked in the pattern's output row), then zero mean
// and unit norm under the mask. The same arithmetic in the same order wherever img and aux live.
template <typename T>
__device__ void prep_one(const T* __restrict__ src, const float* __restrict__ bg, const PrepArg& a, float* __restrict__ dst, float* img,
                         float* aux, float* red) {
    const int hw = a.h * a.w;
    const int tid = threadIdx.x;
    for (int i = tid; i < hw; i += 256) img[i] = (float)src[i] - (bg ? bg[i] : 0.f);
    __syncthreads();
    if (a.dyn) {
        const int dyn = a.dyn;
        gauss(img, aux, a.h, a.w, a.dynb, [&](int i, float b) { img[i] = dyn == 1 ? img[i] - b : img[i] / fmaxf(b, 1e-6f); });
    }
    if (a.clahe) clahe_block(img, reinterpret_cast<unsigned*>(aux), a, red);
    if (a.bp) {
        // the coarse Gaussian into the output row (each pixel back on the thread that wrote it)
        gauss(img, aux, a.h, a.w, a.coarse, [&](int i, float b) { dst[i] = b; });
        if (a.bp_fine) {
            gauss(img, aux, a.h, a.w, a.fine, [&](int i, float b) { img[i] = b - dst[i]; });
        } else {
            for (int i = tid; i < hw; i += 256) img[i] -= dst[i];
            __syncthreads();
        }
    }
    // zero mean and unit norm under the mask (as k_normalize)
    float sm = 0.f, cn = 0.f;
    for (int i = tid; i < hw; i += 256) {
        if (in_mask((float)(i / a.w), (float)(i % a.w), a.mask)) {
            sm += img[i];
            cn += 1.f;
        }
    }
    sm = block_sum<256>(sm, red);
    cn = block_sum<256>(cn, red);
    const float mean = sm / fmaxf(cn, 1.f);
    float ss = 0.f;
    for (int i = tid; i < hw; i += 256) {
        float v = 0.f;
        if (in_mask((float)(i / a.w), (float)(i % a.w), a.mask)) v = img[i] - mean;
        img[i] = v;
        ss += v * v;
    }
    ss = block_sum<256>(ss, red);
    const float inv = 1.f / sqrtf(fmaxf(ss, 1e-30f));
    for (int i = tid; i < hw; i += 256) dst[i] = img[i] * inv;
}

// One block per pattern with the image in LDS: HBM sees the raw pattern and the result once each.
template <typename T>
__global__ void __launch_bounds__(256) k_prep(const T* __restrict__ raw, const float* __restrict__ bg, PrepArg a, float* __restrict__ out) {
    extern __shared__ float lds[];
    __shared__ float red[8];
    const int hw = a.h * a.w;
    prep_one(raw + (size_t)blockIdx.x * hw, bg, a, out + (size_t)blockIdx.x * hw, lds, lds + ((hw + 3) & ~3), red);
}

// Images past one block's LDS: each block works in its own slot of `work` (slot floats, the
// image then the scratch) and takes patterns blockIdx.x, + gridDim.x, ... in turn.
template <typename T>
__global__ void __launch_bounds__(256) k_prep_g(const T* __restrict__ raw, const float* __restrict__ bg, PrepArg a, int n, float* __restrict__ work,
                                                long slot, float* __restrict__ out) {
    __shared__ float red[8];
    const int hw = a.h * a.w;
    float* img = work + (size_t)blockIdx.x * slot;
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        prep_one(raw + (size_t)b * hw, bg, a, out + (size_t)b * hw, img, img + ((hw + 3) & ~3), red);
        __syncthreads();  // every thread done with the slot before the next pattern's loads
    }
}

struct FrameArg {
    float c0[3], c1[3], c2[3];
    int h, w;
    float mask[3];
};

static FrameArg frame_arg(const zb_frame* f) {
    FrameArg a;
    for (int i = 0; i < 3; ++i) {
        a.c0[i] = f->c0[i];
        a.c1[i] = f->c1[i];
        a.c2[i] = f->c2[i];
        a.mask[i] = f->mask[i];
    }
    a.h = f->pat_h;
    a.w = f->pat_w;
    return a;
}

// Simulated pattern pixels: the master in direction q . normalize(ray) (detector.wgsl:254-263).
__global__ void __launch_bounds__(256) k_project(const float* __restrict__ master, int side, FrameArg fr, const float* __restrict__ poses, int n, float* __restrict__ out) {
    const int hw = fr.h * fr.w;
    const long tid = (long)blockIdx.x * 256 + threadIdx.x;
    if (tid >= (long)n * hw) return;
    const int b = (int)(tid / hw);
    const int pix = (int)(tid - (long)b * hw);
    const float row = (float)(pix / fr.w), col = (float)(pix % fr.w);
    const float* p = poses + (size_t)b * 8;
    float3 u = make_float3(fr.c0[0] * row + fr.c1[0] * col + p[0], fr.c0[1] * row + fr.c1[1] * col + p[1], fr.c0[2] * row + fr.c1[2] * col + p[2]);
    float4 q = make_float4(p[4], p[5], p[6], p[7]);
    out[tid] = sheet_bilinear<float>(master, side, qrot(q, normalize3(u)));
}

__device__ __forceinline__ float3 texel_dir(int tid, int side) {
    const int plane = side * side;
    const int pix = tid % plane;
    const float last = (float)(side - 1);
    float2 p = make_float2(2.f * (float)(pix % side) / last - 1.f, 2.f * (float)(pix / side) / last - 1.f);
    return square_to_sphere(p, tid >= plane);
}

__device__ __forceinline__ void tangent_frame(float3 d, float3& e1, float3& e2) {
    float3 ref = fabsf(d.x) > 0.8f ? make_float3(0.f, 1.f, 0.f) : make_float3(1.f, 0.f, 0.f);
    e1 = normalize3(cross3(d, ref));
    e2 = cross3(d, e1);
}

// Gaussian on the sphere, one pass: a 9 x 9 lattice in each texel's tangent plane
// (mastersph.wgsl:87-109).
__global__ void __launch_bounds__(256) k_blur(const float* __restrict__ src, float* __restrict__ dst, int side, float step, float k2) {
    const int tid = blockIdx.x * 256 + threadIdx.x;
    if (tid >= 2 * side * side) return;
    float3 d = texel_dir(tid, side);
    float3 e1, e2;
    tangent_frame(d, e1, e2);
    float acc = 0.f, wsum = 0.f;
    for (int i = -4; i <= 4; ++i) {
        for (int j = -4; j <= 4; ++j) {
            float a = (float)i * step, b = (float)j * step;
            float w = expf(-(a * a + b * b) * k2);
            float3 v = normalize3(make_float3(d.x + a * e1.x + b * e2.x, d.y + a * e1.y + b * e2.y, d.z + a * e1.z + b * e2.z));
            acc += w * sheet_bilinear<float>(src, side, v);
            wsum += w;
        }
    }
    dst[tid] = acc / wsum;
}

// (grad M, M) at every texel by central differences one texel apart (mastersph.wgsl:124-138).
__global__ void __launch_bounds__(256) k_gradient(const float* __restrict__ src, int side, float h, float4* __restrict__ grad) {
    const int tid = blockIdx.x * 256 + threadIdx.x;
    if (tid >= 2 * side * side) return;
    float3 c = texel_dir(tid, side);
    float3 e1, e2;
    tangent_frame(c, e1, e2);
    float g1 = (sheet_bilinear<float>(src, side, normalize3(make_float3(c.x + h * e1.x, c.y + h * e1.y, c.z + h * e1.z))) -
                sheet_bilinear<float>(src, side, normalize3(make_float3(c.x - h * e1.x, c.y - h * e1.y, c.z - h * e1.z)))) / (2.f * h);
    float g2 = (sheet_bilinear<float>(src, side, normalize3(make_float3(c.x + h * e2.x, c.y + h * e2.y, c.z + h * e2.z))) -
                sheet_bilinear<float>(src, side, normalize3(make_float3(c.x - h * e2.x, c.y - h * e2.y, c.z - h * e2.z)))) / (2.f * h);
    grad[tid] = make_float4(g1 * e1.x + g2 * e2.x, g1 * e1.y + g2 * e2.y, g1 * e1.z + g2 * e2.z, src[tid]);
}

// The mean over the master's point group (ops [n_ops][3][3], row-major): each texel the
// average of the sheets at op . d (mastersph.wgsl:111-122).
__global__ void __launch_bounds__(256) k_symmetrize(const float* __restrict__ src, float* __restrict__ dst, int side, const float* __restrict__ ops,
                                                    int n_ops) {
    const int tid = blockIdx.x * 256 + threadIdx.x;
    if (tid >= 2 * side * side) return;
    const float3 d = texel_dir(tid, side);
    float acc = 0.f;
    for (int g = 0; g < n_ops; ++g) {
        const float* m = ops + g * 9;
        const float3 v = make_float3(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[3] * d.x + m[4] * d.y + m[5] * d.z, m[6] * d.x + m[7] * d.y + m[8] * d.z);
        acc += sheet_bilinear<float>(src, side, normalize3(v));
    }
    dst[tid] = acc / (float)max(n_ops, 1);
}

}  // namespace zb

using namespace zb;

extern "C" int32_t zb_normalize(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], float* out,
                                void* stream) {
    if (n < 0 || h < 2 || w < 2 || (n > 0 && (!raw || !out)) || !mask || pixel < 0 || pixel > 2) {
        set_error("zb_normalize: bad arguments (n %d, h %d, w %d, pixel %d)", n, h, w, pixel);
        return ZB_EINVAL;
    }
    if (n == 0) return ZB_OK;
    MaskArg m;
    m.m[0] = mask[0];
    m.m[1] = mask[1];
    m.m[2] = mask[2];
    hipStream_t s = (hipStream_t)stream;
    const long hw = (long)h * w;
    // four pixels a load where every pattern's row of h w starts 4-pixel aligned in both buffers
    const bool vec = hw % 4 == 0 && ((uintptr_t)raw % (4 * ((pixel == 0) ? 1 : (pixel == 1) ? 2 : 4))) == 0 && ((uintptr_t)out % 16) == 0;
#define ZB_NORM(T)                                                                                                            \
    do {                                                                                                                      \
        if (vec && hw <= 4096)                                                                                                \
            ZB_LAUNCH("k_normalize", (k_normalize<T, 4, 4>), dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);        \
        else if (vec && hw <= 16384)                                                                                          \
            ZB_LAUNCH("k_normalize", (k_normalize<T, 16, 4>), dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);       \
        else if (hw <= 4096)                                                                                                  \
            ZB_LAUNCH("k_normalize", (k_normalize<T, 16, 1>), dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);       \
        else if (hw <= 16384)                                                                                                 \
            ZB_LAUNCH("k_normalize", (k_normalize<T, 64, 1>), dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);       \
        else                                                                                                                  \
            ZB_LAUNCH("k_normalize", k_normalize_big<T>, dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);            \
    } while (0)
    if (pixel == 0)
        ZB_NORM(uint8_t);
    else if (pixel == 1)
        ZB_NORM(uint16_t);
    else
        ZB_NORM(float);
#undef ZB_NORM
    return check_launch("zb_normalize");
}

// The Gaussian's parameters for deviation sigma pixels on an h x w image, and the LDS words its
// scratch takes (beside the image).
static size_t blur_arg(double sigma, int h, int w, BlurArg& g) {
    const long hw = (long)h * w;
    const int s = (int)floor(sigma / 6);
    g = BlurArg{};
    if (s < 2) {
        g.box = 0;
        g.radius = (int)fmax(1.0, ceil(3 * sigma));
        g.k = (float)(1.0 / (2 * sigma * sigma));
        return (size_t)((hw + 3) & ~3L) + (size_t)(g.radius + 1);
    }
    const double cs = sqrt(sigma * sigma - (double)s * s - ((double)s * s - 1) / 12);
    g.box = 1;
    g.s = s;
    g.pad = 3;
    g.reach = (float)ceil(4 * cs);
    g.k0 = (float)(1.0 / (2 * cs * cs));
    g.k2 = (float)(1.0 / (2.0 * s * s));
    g.hs = (h + s - 1) / s;
    g.ws = (w + s - 1) / s;
    g.hp = g.hs + 2 * g.pad;
    g.wp = g.ws + 2 * g.pad;
    return (size_t)2 * g.hp * (g.wp > w ? g.wp : w) * 2;
}

// The filter chain's parameters, and the LDS words its scratch takes beside the image; -1 (with
// the reason kept) for bad arguments.
static long prep_arg(const char* who, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* o, PrepArg& a) {
    if (n < 0 || h < 2 || w < 2 || !mask || !o || pixel < 0 || pixel > 2 || o->dynamic < 0 || o->dynamic > 2 ||
        (o->dynamic && !(o->dynamic_sigma > 0.f)) ||
        (o->clahe && (o->clahe_grid < 1 || o->clahe_grid > 16 || o->clahe_bins < 2 || o->clahe_bins > 256 || !(o->clahe_clip > 0.f))) ||
        (o->bandpass && !(o->bp_fine >= 0.f && o->bp_coarse > o->bp_fine)) || (long)h * w > (1L << 28)) {
        set_error("%s: bad arguments (n %d, h %d, w %d, pixel %d, dynamic %d, clahe %d, bandpass %d)", who, n, h, w, pixel, o ? o->dynamic : 0,
                  o ? o->clahe : 0, o ? o->bandpass : 0);
        return -1;
    }
    a = PrepArg{};
    a.h = h;
    a.w = w;
    for (int i = 0; i < 3; ++i) a.mask[i] = mask[i];
    a.dyn = o->dynamic;
    size_t scratch = 0;  // words beside the image
    if (a.dyn) scratch = std::max(scratch, blur_arg(o->dynamic_sigma, h, w, a.dynb));
    if (o->clahe) {
        a.clahe = 1;
        a.tiles = o->clahe_grid;
        a.bins = o->clahe_bins;
        a.clip = o->clahe_clip;
        a.pad_h = (a.tiles - h % a.tiles) % a.tiles;
        a.pad_w = (a.tiles - w % a.tiles) % a.tiles;
        a.tile_h = std::max(1, (h + a.pad_h) / a.tiles);
        a.tile_w = std::max(1, (w + a.pad_w) / a.tiles);
        scratch = std::max(scratch, (size_t)a.tiles * a.tiles * a.bins);
    }
    if (o->bandpass) {
        a.bp = 1;
        scratch = std::max(scratch, blur_arg(o->bp_coarse, h, w, a.coarse));
        if (o->bp_fine > 0.05f) {
            a.bp_fine = 1;
            scratch = std::max(scratch, blur_arg(o->bp_fine, h, w, a.fine));
        }
    }
    return (long)scratch;
}

// Bytes of LDS a block of k_prep takes, and what the device gives one.
static int prep_lds(const PrepArg& a, long scratch, size_t& lds, int& limit) {
    const long hw = (long)a.h * a.w;
    lds = ((size_t)((hw + 3) & ~3L) + (size_t)scratch) * 4;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return -1;
    return 0;
}

// Blocks of k_prep_g, each with its slot of the workspace.
constexpr int kPrepBlocks = 1024;

static long prep_slot(const PrepArg& a, long scratch) {
    const long hw = (long)a.h * a.w;
    return ((((hw + 3) & ~3L) + scratch) + 63) & ~63L;  // floats, slots 256-byte aligned
}

template <typename T>
static void prep_launch(const void* raw, const float* bg, const PrepArg& a, int n, bool global, size_t lds, float* work, long slot, int grid,
                        float* out, hipStream_t s) {
    if (global)
        ZB_LAUNCH("k_prep", k_prep_g<T>, dim3(grid), dim3(256), 0, s, (const T*)raw, bg, a, n, work, slot, out);
    else
        ZB_LAUNCH("k_prep", k_prep<T>, dim3(n), dim3(256), lds, s, (const T*)raw, bg, a, out);
}

static int32_t prep_run(const char* who, const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* o,
                        void* work, int64_t work_bytes, bool may_global, float* out, void* stream) {
    PrepArg a;
    const long scratch = prep_arg(who, pixel, n, h, w, mask, o, a);
    if (scratch < 0) return ZB_EINVAL;
    if (n == 0) return ZB_OK;
    if (!raw || !out) {
        set_error("%s: null pointer", who);
        return ZB_EINVAL;
    }
    if (((uintptr_t)work & 7) != 0) {  // the box blur's scratch is read as float pairs
        set_error("%s: the workspace must be 8-byte aligned", who);
        return ZB_EINVAL;
    }
    size_t lds = 0;
    int limit = 0;
    if (prep_lds(a, scratch, lds, limit)) {
        set_error("%s: no device", who);
        return ZB_ELAUNCH;
    }
    const bool global = lds > (size_t)limit || (may_global && knob(KNOB_PREP_GLOBAL) != 0);
    if (global && !may_global) {
        set_error("%s: a %d x %d pattern with these filters needs %zu bytes of LDS, the device gives a block %d (zb_prep_work takes it "
                  "through global memory)",
                  who, h, w, lds, limit);
        return ZB_EINVAL;
    }
    const long slot = prep_slot(a, scratch);
    int grid = 0;
    if (global) {
        // a block a slot: as many as the workspace holds, up to kPrepBlocks (fewer run slower)
        grid = (int)std::min<long>(std::min<long>(n, kPrepBlocks), work ? work_bytes / (slot * 4) : 0);
        if (grid < 1) {
            set_error("%s: workspace %lld bytes, need at least %lld (one slot; %lld for the full grid)", who, (long long)work_bytes,
                      (long long)slot * 4, (long long)std::min(n, kPrepBlocks) * slot * 4);
            return ZB_ESPACE;
        }
    }
    hipStream_t s = (hipStream_t)stream;
    const float* bg = o->background;
    if (pixel == 0)
        prep_launch<uint8_t>(raw, bg, a, n, global, lds, (float*)work, slot, grid, out, s);
    else if (pixel == 1)
        prep_launch<uint16_t>(raw, bg, a, n, global, lds, (float*)work, slot, grid, out, s);
    else
        prep_launch<float>(raw, bg, a, n, global, lds, (float*)work, slot, grid, out, s);
    return check_launch(who);
}

extern "C" int32_t zb_prep(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* o,
                           float* out, void* stream) {
    return prep_run("zb_prep", raw, pixel, n, h, w, mask, o, nullptr, 0, false, out, stream);
}

extern "C" int64_t zb_prep_workspace(int32_t n, int32_t h, int32_t w, const zb_prep_opts* o) {
    PrepArg a;
    const float mask[3] = {0.f, 0.f, 0.f};
    const long scratch = prep_arg("zb_prep_workspace", 2, n, h, w, mask, o, a);
    if (scratch < 0) return ZB_EINVAL;
    size_t lds = 0;
    int limit = 0;
    if (prep_lds(a, scratch, lds, limit)) {
        set_error("zb_prep_workspace: no device");
        return ZB_ELAUNCH;
    }
    if (lds <= (size_t)limit && knob(KNOB_PREP_GLOBAL) == 0) return 0;
    return (int64_t)std::min(n, kPrepBlocks) * prep_slot(a, scratch) * 4;
}

extern "C" int32_t zb_prep_work(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* o,
                                void* work, int64_t work_bytes, float* out, void* stream) {
    return prep_run("zb_prep_work", raw, pixel, n, h, w, mask, o, work, work_bytes, true, out, stream);
}

extern "C" int32_t zb_project(const float* master, int32_t side, const zb_frame* frame, const float* poses, int32_t n, float* out,
                              void* stream) {
    if (!frame || side < 2 || n < 0 || frame->pat_h < 2 || frame->pat_w < 2 || (n > 0 && (!master || !poses || !out))) {
        set_error("zb_project: bad arguments");
        return ZB_EINVAL;
    }
    if (n == 0) return ZB_OK;
    FrameArg fr = frame_arg(frame);
    long total = (long)n * fr.h * fr.w;
    ZB_LAUNCH("k_project", k_project, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, master, side, fr, poses, n, out);
    return check_launch("zb_project");
}

extern "C" int32_t zb_master_blur(const float* src, float* dst, int32_t side, float sigma_rad, void* stream) {
    if (!src || !dst || side < 2 || !(sigma_rad > 0.f)) {
        set_error("zb_master_blur: bad arguments");
        return ZB_EINVAL;
    }
    int total = 2 * side * side;
    ZB_LAUNCH("k_blur", k_blur, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, src, dst, side, 0.6f * sigma_rad,
                       1.f / (2.f * sigma_rad * sigma_rad));
    return check_launch("zb_master_blur");
}

extern "C" int32_t zb_master_gradient(const float* sheets, int32_t side, float* grad, void* stream) {
    if (!sheets || !grad || side < 2) {
        set_error("zb_master_gradient: bad arguments");
        return ZB_EINVAL;
    }
    int total = 2 * side * side;
    float h = (float)(2.5066282746310002 / (double)side);  // sqrt(2 pi) / side
    ZB_LAUNCH("k_gradient", k_gradient, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, sheets, side, h, (float4*)grad);
    return check_launch("zb_master_gradient");
}

extern "C" int64_t zb_clahe_workspace(int32_t n, int32_t grid, int32_t bins) {
    if (n < 0 || grid < 1 || grid > 64 || bins < 2 || bins > 256) {
        set_error("zb_clahe_workspace: bad arguments (n %d, grid %d, bins %d)", n, grid, bins);
        return ZB_EINVAL;
    }
    return (int64_t)n * grid * grid * bins * 4;
}

extern "C" int32_t zb_clahe(float* img, int32_t n, int32_t h, int32_t w, int32_t grid, int32_t bins, float clip, void* work, int64_t work_bytes,
                            void* stream) {
    if (n < 0 || h < 2 || w < 2 || grid < 1 || grid > 64 || bins < 2 || bins > 256 || !(clip > 0.f) || (n > 0 && (!img || !work))) {
        set_error("zb_clahe: bad arguments (n %d, h %d, w %d, grid %d, bins %d)", n, h, w, grid, bins);
        return ZB_EINVAL;
    }
    if (n == 0) return ZB_OK;
    if (work_bytes < zb_clahe_workspace(n, grid, bins)) {
        set_error("zb_clahe: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)zb_clahe_workspace(n, grid, bins));
        return ZB_ESPACE;
    }
    ClaheArg a;
    a.n = n;
    a.h = h;
    a.w = w;
    a.tiles = grid;
    a.bins = bins;
    a.clip = clip;
    a.pad_h = (grid - h % grid) % grid;
    a.pad_w = (grid - w % grid) % grid;
    a.tile_h = std::max(1, (h + a.pad_h) / grid);
    a.tile_w = std::max(1, (w + a.pad_w) / grid);
    hipStream_t s = (hipStream_t)stream;
    float* cdf = (float*)work;
    ZB_LAUNCH("k_clahe_stretch", k_clahe_stretch, dim3(n), dim3(256), 0, s, img, a);
    ZB_LAUNCH("k_clahe_tables", k_clahe_tables, dim3((unsigned)(n * grid * grid)), dim3(256), 0, s, (const float*)img, a, cdf);
    const long total = (long)n * h * w;
    ZB_LAUNCH("k_clahe_apply", k_clahe_apply, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, img, a, (const float*)cdf);
    return check_launch("zb_clahe");
}

extern "C" int32_t zb_master_symmetrize(const float* src, float* dst, int32_t side, const float* ops, int32_t n_ops, void* stream) {
    if (!src || !dst || src == dst || side < 2 || n_ops < 1 || !ops) {
        set_error("zb_master_symmetrize: bad arguments");
        return ZB_EINVAL;
    }
    const int total = 2 * side * side;
    ZB_LAUNCH("k_symmetrize", k_symmetrize, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, src, dst, side, ops, n_ops);
    return check_launch("zb_master_symmetrize");
}

// The gradient field in half precision (each float rounded to nearest even), 8 bytes a texel,
// each texel stored with the texel below it (16 bytes; the last row of a sheet with itself): a
// lookup's four texels are two adjacent 16-byte loads, (o, o + side) and (o + 1, o + 1 + side),
// which mostly share one cache line, where a row-major field puts the two rows in two.
__global__ void __launch_bounds__(256) k_pack_half(const float4* __restrict__ grad, int side, int total, uint4* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int row = (i % (side * side)) / side;
    const float4 g = grad[i];
    const float4 u = grad[row < side - 1 ? i + side : i];
    out[i] = make_uint4(pack_half2(g.x, g.y), pack_half2(g.z, g.w), pack_half2(u.x, u.y), pack_half2(u.z, u.w));
}

extern "C" int32_t zb_master_pack(const float* grad, int32_t side, uint16_t* grad_h, void* stream) {
    if (!grad || !grad_h || side < 2) {
        set_error("zb_master_pack: bad arguments");
        return ZB_EINVAL;
    }
    const int total = 2 * side * side;
    ZB_LAUNCH("k_pack_half", k_pack_half, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, (const float4*)grad, side, total,
              (uint4*)grad_h);
    return check_launch("zb_master_pack");
}
