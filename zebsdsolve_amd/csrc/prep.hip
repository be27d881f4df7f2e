// Patterns in, masters baked: per-pattern normalization under the mask, simulated patterns,
// and the master's smoothing and gradient on the sphere (setup, once per master and geometry).
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

// One block of 256 per pattern: mean and norm under the mask (prep.wgsl:405-443). The pattern
// is read once into registers (PER pixels a thread, PER * 256 >= h w), written once.
template <typename T, int PER>
__global__ void __launch_bounds__(256) k_normalize(const T* __restrict__ raw, int h, int w, MaskArg mask, float* __restrict__ out) {
    __shared__ float red[4];
    const int b = blockIdx.x;
    const int hw = h * w;
    const T* src = raw + (size_t)b * hw;
    float* dst = out + (size_t)b * hw;
    const float inv_w = 1.f / (float)w;
    float v[PER];
    float s = 0.f, c = 0.f;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 256 * k;
        v[k] = 0.f;
        if (i < hw) {
            // row and column without an integer division (exact for h w <= 65536)
            const int row = (int)(((float)i + 0.5f) * inv_w);
            if (in_mask((float)row, (float)(i - row * w), mask.m)) {
                v[k] = (float)src[i];
                s += v[k];
                c += 1.f;
            }
        }
    }
    s = block_sum<256>(s, red);
    c = block_sum<256>(c, red);
    const float mean = s / fmaxf(c, 1.f);
    float ss = 0.f;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 256 * k;
        const int row = (int)(((float)i + 0.5f) * inv_w);
        const bool in = i < hw && in_mask((float)row, (float)(i - row * w), mask.m);
        v[k] = in ? v[k] - mean : 0.f;
        ss += v[k] * v[k];
    }
    ss = block_sum<256>(ss, red);
    const float inv = 1.f / sqrtf(fmaxf(ss, 1e-30f));
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + 256 * k;
        if (i < hw) dst[i] = v[k] * inv;
    }
}

// The same for frames of more than 16384 pixels, in passes over memory.
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
struct PrepArg {
    int h, w;
    float mask[3];
    int dyn;       // 0 none, 1 subtract, 2 divide
    int box;       // 0: direct taps, 1: the box path
    int radius;    // direct: taps each way
    float k;       // direct: 1 / (2 sigma^2)
    int s, pad;    // box: cell size in pixels, margin in cells
    float reach;   // box: ceil(4 sigma_coarse), pixels
    float k0, k2;  // box: 1 / (2 sigma_coarse^2), 1 / (2 s^2)
    int hs, ws, hp, wp;
};

// centre (pixels) and width of coarse cell j along an axis of n pixels (cells outside are whole)
__device__ __forceinline__ float2 prep_cell(int j, int n, int s) {
    const int lo = j * s;
    int hi = lo + s;
    if (j >= 0 && lo < n) hi = min(hi, n);
    return make_float2(0.5f * (float)(lo + hi - 1), (float)(hi - lo));
}

template <typename T>
__global__ void __launch_bounds__(256) k_prep(const T* __restrict__ raw, const float* __restrict__ bg, PrepArg a, float* __restrict__ out) {
    extern __shared__ float lds[];
    __shared__ float red[4];
    const int hw = a.h * a.w;
    const int tid = threadIdx.x;
    float* img = lds;
    float* aux = lds + ((hw + 3) & ~3);
    const T* src = raw + (size_t)blockIdx.x * hw;
    for (int i = tid; i < hw; i += 256) img[i] = (float)src[i] - (bg ? bg[i] : 0.f);
    __syncthreads();
    auto combine = [&](float v, float b) { return a.dyn == 1 ? v - b : v / fmaxf(b, 1e-6f); };
    if (a.dyn && !a.box) {
        float* wt = aux + ((hw + 3) & ~3);  // exp(-d^2 k), d <= radius
        for (int d = tid; d <= a.radius; d += 256) wt[d] = expf(-(float)(d * d) * a.k);
        __syncthreads();
        for (int i = tid; i < hw; i += 256) {
            const int r = i / a.w, c = i - (i / a.w) * a.w;
            float acc = 0.f, ws = 0.f;
            const int x1 = min(c + a.radius, a.w - 1);
            for (int x = max(c - a.radius, 0); x <= x1; ++x) {
                const float g = wt[abs(x - c)];
                acc += g * img[r * a.w + x];
                ws += g;
            }
            aux[i] = acc / ws;
        }
        __syncthreads();
        for (int i = tid; i < hw; i += 256) {
            const int r = i / a.w, c = i - (i / a.w) * a.w;
            float acc = 0.f, ws = 0.f;
            const int y1 = min(r + a.radius, a.h - 1);
            for (int y = max(r - a.radius, 0); y <= y1; ++y) {
                const float g = wt[abs(y - r)];
                acc += g * aux[y * a.w + c];
                ws += g;
            }
            img[i] = combine(img[i], acc / ws);
        }
        __syncthreads();
    } else if (a.dyn) {
        const int s = a.s, pad = a.pad;
        const int big = a.hp * (a.wp > a.w ? a.wp : a.w);
        float2* c0 = reinterpret_cast<float2*>(aux);
        float2* c1 = c0 + big;
        // pixel sums and counts over the s x s boxes: [hs][ws]
        for (int t = tid; t < a.hs * a.ws; t += 256) {
            const int i = t / a.ws, j = t - (t / a.ws) * a.ws;
            const int r1 = min((i + 1) * s, a.h), q1 = min((j + 1) * s, a.w);
            float acc = 0.f;
            for (int r = i * s; r < r1; ++r)
                for (int q = j * s; q < q1; ++q) acc += img[r * a.w + q];
            c0[t] = make_float2(acc, (float)((r1 - i * s) * (q1 - j * s)));
        }
        __syncthreads();
        // along the rows onto the coarse grid with its margin: [hs][wp]
        for (int t = tid; t < a.hs * a.wp; t += 256) {
            const int line = t / a.wp, tc = t - (t / a.wp) * a.wp;
            const float cc = prep_cell(tc - pad, a.w, s).x;
            const int j0 = max((int)floorf((cc - a.reach) / (float)s), 0);
            const int j1 = min((int)floorf((cc + a.reach) / (float)s), a.ws - 1);
            float2 acc = make_float2(0.f, 0.f);
            for (int j = j0; j <= j1; ++j) {
                const float d = prep_cell(j, a.w, s).x - cc;
                const float g = expf(-d * d * a.k0);
                const float2 v = c0[line * a.ws + j];
                acc = make_float2(acc.x + g * v.x, acc.y + g * v.y);
            }
            c1[t] = acc;
        }
        __syncthreads();
        // along the columns: [hp][wp]
        for (int t = tid; t < a.hp * a.wp; t += 256) {
            const int tr = t / a.wp, col = t - (t / a.wp) * a.wp;
            const float cc = prep_cell(tr - pad, a.h, s).x;
            const int i0 = max((int)floorf((cc - a.reach) / (float)s), 0);
            const int i1 = min((int)floorf((cc + a.reach) / (float)s), a.hs - 1);
            float2 acc = make_float2(0.f, 0.f);
            for (int i = i0; i <= i1; ++i) {
                const float d = prep_cell(i, a.h, s).x - cc;
                const float g = expf(-d * d * a.k0);
                const float2 v = c1[i * a.wp + col];
                acc = make_float2(acc.x + g * v.x, acc.y + g * v.y);
            }
            c0[t] = acc;
        }
        __syncthreads();
        // spread over the pixels of each row from the 2 pad + 1 cells around: [hp][w]
        for (int t = tid; t < a.hp * a.w; t += 256) {
            const int tr = t / a.w, x = t - (t / a.w) * a.w;
            const int own = x / s;
            float2 acc = make_float2(0.f, 0.f);
            for (int q = 0; q <= 2 * pad; ++q) {
                const float2 cw = prep_cell(own + q - pad, a.w, s);
                const float d = (float)x - cw.x;
                const float g = cw.y * expf(-d * d * a.k2);
                const float2 v = c0[tr * a.wp + own + q];
                acc = make_float2(acc.x + g * v.x, acc.y + g * v.y);
            }
            c1[t] = acc;
        }
        __syncthreads();
        // and of each column, into the pixels: the blurred value, combined
        for (int i = tid; i < hw; i += 256) {
            const int y = i / a.w, x = i - (i / a.w) * a.w;
            const int own = y / s;
            float2 acc = make_float2(0.f, 0.f);
            for (int q = 0; q <= 2 * pad; ++q) {
                const float2 cw = prep_cell(own + q - pad, a.h, s);
                const float d = (float)y - cw.x;
                const float g = cw.y * expf(-d * d * a.k2);
                const float2 v = c1[(own + q) * a.w + x];
                acc = make_float2(acc.x + g * v.x, acc.y + g * v.y);
            }
            img[i] = combine(img[i], acc.x / acc.y);
        }
        __syncthreads();
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
    float* dst = out + (size_t)blockIdx.x * hw;
    for (int i = tid; i < hw; i += 256) dst[i] = img[i] * inv;
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
#define ZB_NORM(T)                                                                                                            \
    do {                                                                                                                      \
        if (hw <= 4096)                                                                                                       \
            ZB_LAUNCH("k_normalize", (k_normalize<T, 16>), dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);          \
        else if (hw <= 16384)                                                                                                 \
            ZB_LAUNCH("k_normalize", (k_normalize<T, 64>), dim3(n), dim3(256), 0, s, (const T*)raw, h, w, m, out);          \
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

extern "C" int32_t zb_prep(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* o,
                           float* out, void* stream) {
    if (n < 0 || h < 2 || w < 2 || (n > 0 && (!raw || !out)) || !mask || !o || pixel < 0 || pixel > 2 || o->dynamic < 0 ||
        o->dynamic > 2 || (o->dynamic && !(o->dynamic_sigma > 0.f))) {
        set_error("zb_prep: bad arguments (n %d, h %d, w %d, pixel %d)", n, h, w, pixel);
        return ZB_EINVAL;
    }
    if (n == 0) return ZB_OK;
    PrepArg a = {};
    a.h = h;
    a.w = w;
    for (int i = 0; i < 3; ++i) a.mask[i] = mask[i];
    a.dyn = o->dynamic;
    const long hw = (long)h * w;
    size_t lds = (size_t)((hw + 3) & ~3L) * 4;  // the image
    if (a.dyn) {
        const double sigma = o->dynamic_sigma;
        const int s = (int)floor(sigma / 6);
        if (s < 2) {
            a.box = 0;
            a.radius = (int)fmax(1.0, ceil(3 * sigma));
            a.k = (float)(1.0 / (2 * sigma * sigma));
            lds += (size_t)((hw + 3) & ~3L) * 4 + (size_t)(a.radius + 1) * 4;
        } else {
            const double cs = sqrt(sigma * sigma - (double)s * s - ((double)s * s - 1) / 12);
            a.box = 1;
            a.s = s;
            a.pad = 3;
            a.reach = (float)ceil(4 * cs);
            a.k0 = (float)(1.0 / (2 * cs * cs));
            a.k2 = (float)(1.0 / (2.0 * s * s));
            a.hs = (h + s - 1) / s;
            a.ws = (w + s - 1) / s;
            a.hp = a.hs + 2 * a.pad;
            a.wp = a.ws + 2 * a.pad;
            lds += (size_t)2 * a.hp * (a.wp > w ? a.wp : w) * 8;
        }
    }
    int dev = 0, limit = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
        set_error("zb_prep: no device");
        return ZB_ELAUNCH;
    }
    if (lds > (size_t)limit) {
        set_error("zb_prep: a %d x %d pattern with these filters needs %zu bytes of LDS, the device gives a block %d", h, w, lds, limit);
        return ZB_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const float* bg = o->background;
    if (pixel == 0)
        ZB_LAUNCH("k_prep", k_prep<uint8_t>, dim3(n), dim3(256), lds, s, (const uint8_t*)raw, bg, a, out);
    else if (pixel == 1)
        ZB_LAUNCH("k_prep", k_prep<uint16_t>, dim3(n), dim3(256), lds, s, (const uint16_t*)raw, bg, a, out);
    else
        ZB_LAUNCH("k_prep", k_prep<float>, dim3(n), dim3(256), lds, s, (const float*)raw, bg, a, out);
    return check_launch("zb_prep");
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

// The gradient field in half precision (each float rounded to nearest even): 8 bytes a texel, so
// that a texel and its right-hand neighbour come in one 16-byte load.
__global__ void __launch_bounds__(256) k_pack_half(const float4* __restrict__ grad, int total, uint2* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const float4 g = grad[i];
    out[i] = make_uint2(pack_half2(g.x, g.y), pack_half2(g.z, g.w));
}

extern "C" int32_t zb_master_pack(const float* grad, int32_t side, uint16_t* grad_h, void* stream) {
    if (!grad || !grad_h || side < 2) {
        set_error("zb_master_pack: bad arguments");
        return ZB_EINVAL;
    }
    const int total = 2 * side * side;
    ZB_LAUNCH("k_pack_half", k_pack_half, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, (const float4*)grad, total,
              (uint2*)grad_h);
    return check_launch("zb_master_pack");
}
