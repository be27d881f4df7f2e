// Shared device code for the indexing kernels (gfx950, wave64): float32 rotations, the
// equal-area map, and an in-place LDS line FFT.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zebsd.h"

#define ZB_PI 3.14159265358979f
#define ZB_TAU 6.28318530717959f
// twiddle table length: tw[t] = exp(-2 pi i t / ZB_TW)
#define ZB_TW 1024
// devices a per-device table of the library (the no-alloc scratch addresses) has room for
#define ZB_MAX_DEVICES 64

// Launch a kernel, bracketed by timing events when profiling is on.
#define ZB_LAUNCH(name, kern, grid, block, lds, stream, ...)                      \
    do {                                                                          \
        const bool zb_p_ = zb::prof_on() && zb::prof_wants(name);                 \
        if (zb_p_) zb::prof_begin(name, stream);                                  \
        hipLaunchKernelGGL(kern, grid, block, lds, stream, __VA_ARGS__);          \
        if (zb_p_) zb::prof_end(stream);                                          \
    } while (0)

namespace zb {

// ── error reporting and kernel timing (api.hip) ──
void set_error(const char* fmt, ...);
int check_launch(const char* what);
// With profiling on, prof_begin records an event on `s` before a launch and prof_end one after
// it, under the kernel's name; zb_profile_report sums them per name.
// The launch-shape knobs (zb_knob_set): their current values, by id.
enum { KNOB_REFINE_QUEUE = 0, KNOB_RINGS_GROUP = 1, KNOB_PREP_GLOBAL = 2, KNOB_COUNT = 3 };
int knob(int id);

bool prof_on();
// whether kernel `name` is timed (zb_profile_select's list, all when empty)
bool prof_wants(const char* name);
void prof_begin(const char* name, hipStream_t s);
void prof_end(hipStream_t s);

// Blocks are dealt to the 8 XCDs round-robin (block id mod 8); this maps block `id` of `G` to a
// logical index so that XCD x takes a contiguous run of the logical order (G / 8 blocks, one
// more when x < G % 8): work adjacent in that order shares the XCD's L2.
__device__ __forceinline__ int xcd_block(int id, int G) {
    const int q = G >> 3, r = G & 7, x = id & 7;
    return x * q + min(x, r) + (id >> 3);
}

// ── quaternions, (w, x, y, z) in (.x, .y, .z, .w), acting as q v q* ──
__device__ __forceinline__ float4 qmul(float4 a, float4 b) {
    return make_float4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w,
                       a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
                       a.x * b.z - a.y * b.w + a.z * b.x + a.w * b.y,
                       a.x * b.w + a.y * b.z - a.z * b.y + a.w * b.x);
}

__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ float3 qrot(float4 q, float3 v) {
    float3 u = make_float3(q.y, q.z, q.w);
    float3 t = cross3(u, v);
    t = make_float3(2.f * t.x, 2.f * t.y, 2.f * t.z);
    float3 c = cross3(u, t);
    return make_float3(v.x + q.x * t.x + c.x, v.y + q.x * t.y + c.y, v.z + q.x * t.z + c.z);
}

__device__ __forceinline__ float4 qnormalize(float4 q) {
    float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n > 0.f) return make_float4(q.x / n, q.y / n, q.z / n, q.w / n);
    return make_float4(1.f, 0.f, 0.f, 0.f);
}

// exp of a rotation vector (common.wgsl qexp)
__device__ __forceinline__ float4 qexp(float3 p) {
    float th2 = dot3(p, p);
    float th = sqrtf(th2);
    float k = 0.5f - th2 / 48.f;
    if (th > 1e-3f) k = sinf(0.5f * th) / th;
    return make_float4(cosf(0.5f * th), k * p.x, k * p.y, k * p.z);
}

// (the hardware reciprocal square root, 1 ulp)
__device__ __forceinline__ float3 normalize3(float3 v) {
    float r = __builtin_amdgcn_rsqf(dot3(v, v));
    return make_float3(v.x * r, v.y * r, v.z * r);
}

// arctan(r) for r in [0, 1] as r P(r^2), P of degree 7: coefficients fitted here to arctan on
// [0, 1] by weighted least squares reweighted toward equal ripple; in float32 Horner evaluation
// the error is at most 1.2e-7 (2 ulp of pi / 4).
__device__ __forceinline__ float atan01(float r) {
    const float x = r * r;
    float p = -0.0040545594f;
    p = fmaf(p, x, 0.021862937f);
    p = fmaf(p, x, -0.055912312f);
    p = fmaf(p, x, 0.096421979f);
    p = fmaf(p, x, -0.13908631f);
    p = fmaf(p, x, 0.19946566f);
    p = fmaf(p, x, -0.33329862f);
    p = fmaf(p, x, 0.99999934f);
    return p * r;
}

// Unit vector -> equal-area square coordinates in [-1, 1]^2 of its own hemisphere. 1 - |z| is
// taken as (x^2 + y^2) / (1 + |z|), which keeps its digits near the pole (common.wgsl:44-57).
// The square root and reciprocals are the hardware's (1 ulp).
__device__ __forceinline__ float2 sphere_to_square(float3 v) {
    float ax = fabsf(v.x), ay = fabsf(v.y);
    float xn = __builtin_amdgcn_sqrtf(fmaxf(ZB_PI * 0.5f * (ax * ax + ay * ay) * __builtin_amdgcn_rcpf(1.f + fabsf(v.z)), 0.f));
    // one arctangent of the smaller over the larger component (the same value either branch of
    // the reference computes, without a divergent pair of calls)
    const bool xbig = ax >= ay;
    const float t = (1.2732395447351627f * xn) * atan01(  // 4 / pi
        (xbig ? ay : ax) * __builtin_amdgcn_rcpf(fmaxf(xbig ? ax : ay, 1e-30f)));
    const float u = xbig ? xn : t, w = xbig ? t : xn;
    const float sc = 0.79788456080286536f;  // 1 / sqrt(pi / 2)
    return make_float2(u * sc * (v.x < 0.f ? -1.f : 1.f), w * sc * (v.y < 0.f ? -1.f : 1.f));
}

// Square coordinates -> unit vector on the northern or southern hemisphere (f32).
__device__ __forceinline__ float3 square_to_sphere(float2 p, bool south) {
    const float sc = 1.2533141373155003f;  // sqrt(pi / 2)
    float xa = fabsf(p.x) * sc, ya = fabsf(p.y) * sc;
    bool swap = xa >= ya;
    float xn = fmaxf(xa, ya), yn = fminf(xa, ya);
    float3 v = make_float3(0.f, 0.f, 1.f);
    if (xn > 0.f) {
        float r = (2.f * xn / ZB_PI) * sqrtf(fmaxf(ZB_PI - xn * xn, 0.f));
        float ang = ZB_PI * yn / (4.f * xn);
        float hx = r * cosf(ang), hy = r * sinf(ang);
        if (!swap) { float t = hx; hx = hy; hy = t; }
        v = normalize3(make_float3(hx * (p.x < 0.f ? -1.f : 1.f), hy * (p.y < 0.f ? -1.f : 1.f), 1.f - 2.f * xn * xn / ZB_PI));
    }
    if (south) v.z = -v.z;
    return v;
}

// Bilinear lookup of a [north | south] Lambert sheet pair of side x side texels of T (float or
// float4) in direction v: the texel row/column clamps as the reference's master_at does.
template <typename T>
__device__ __forceinline__ T sheet_bilinear(const T* __restrict__ s, int side, float3 v);

template <>
__device__ __forceinline__ float sheet_bilinear<float>(const float* __restrict__ s, int side, float3 v) {
    float2 uv = sphere_to_square(v);
    float last = (float)(side - 1);
    float fj = fminf(fmaxf((uv.x + 1.f) * 0.5f * last, 0.f), last);
    float fi = fminf(fmaxf((uv.y + 1.f) * 0.5f * last, 0.f), last);
    int c0 = min((int)fj, side - 2), r0 = min((int)fi, side - 2);
    int o = r0 * side + c0 + (v.z < 0.f ? side * side : 0);
    float tj = fj - (float)c0, ti = fi - (float)r0;
    float a = s[o], b = s[o + 1], c = s[o + side], d = s[o + side + 1];
    float top = a + (b - a) * tj, bot = c + (d - c) * tj;
    return top + (bot - top) * ti;
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + (b - a) * t; }

// The same lookup in two halves, so that a lane can have the loads of several points in flight
// before it blends any: sheet_taps computes the texel and issues the four loads, taps_blend
// interpolates (the arithmetic of sheet_bilinear, operation for operation).
template <typename T>
struct Taps {
    T a, b, c, d;
    float tj, ti;
};

// The texel (offset of its top-left corner over both sheets) and the fractions for direction v.
__device__ __forceinline__ int sheet_texel(int side, float3 v, float& tj, float& ti) {
    float2 uv = sphere_to_square(v);
    float last = (float)(side - 1);
    float fj = fminf(fmaxf((uv.x + 1.f) * 0.5f * last, 0.f), last);
    float fi = fminf(fmaxf((uv.y + 1.f) * 0.5f * last, 0.f), last);
    int c0 = min((int)fj, side - 2), r0 = min((int)fi, side - 2);
    tj = fj - (float)c0;
    ti = fi - (float)r0;
    return r0 * side + c0 + (v.z < 0.f ? side * side : 0);
}

template <typename T>
__device__ __forceinline__ void sheet_taps(const T* __restrict__ s, int side, float3 v, Taps<T>& t) {
    const int o = sheet_texel(side, v, t.tj, t.ti);
    t.a = s[o];
    t.b = s[o + 1];
    t.c = s[o + side];
    t.d = s[o + side + 1];
}

// Only the value channel (w) of a (grad M, M) sheet: a quarter of the bytes.
__device__ __forceinline__ void sheet_taps_w(const float4* __restrict__ s, int side, float3 v, Taps<float>& t) {
    const int o = sheet_texel(side, v, t.tj, t.ti);
    t.a = s[o].w;
    t.b = s[o + 1].w;
    t.c = s[o + side].w;
    t.d = s[o + side + 1].w;
}

// ── half-precision texels: (grad M, M) as four halves, 8 bytes ──
typedef _Float16 zb_half2 __attribute__((ext_vector_type(2)));
typedef unsigned zb_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_half2(float a, float b) {
    const zb_half2 h = {(_Float16)a, (_Float16)b};
    return __builtin_bit_cast(unsigned, h);
}

__device__ __forceinline__ float2 unpack_half2(unsigned u) {
    const zb_half2 h = __builtin_bit_cast(zb_half2, u);
    return make_float2((float)h.x, (float)h.y);
}

// The four texels of a lookup in the half-precision field, where each texel is stored with the
// one below it (16 bytes, zb_master_pack): texels o and o + side in one 16-byte load (left),
// o + 1 and o + 1 + side in the next (right), through a buffer resource over the field.
struct TapsH {
    zb_u32x4 left, right;
    float tj, ti;
};

__device__ __forceinline__ void sheet_taps_h(__amdgpu_buffer_rsrc_t field, int side, float3 v, TapsH& t) {
    const int o = sheet_texel(side, v, t.tj, t.ti);
    t.left = __builtin_amdgcn_raw_buffer_load_b128(field, o * 16, 0, 0);
    t.right = __builtin_amdgcn_raw_buffer_load_b128(field, o * 16 + 16, 0, 0);
}

// The four texels of a lookup in float32 sheets: each row's pair in one 8-byte load.
struct TapsM {
    float2 top, bot;
    float tj, ti;
};

__device__ __forceinline__ void sheet_taps_m(__amdgpu_buffer_rsrc_t sheets, int side, float3 v, TapsM& t) {
    const int o = sheet_texel(side, v, t.tj, t.ti);
    t.top = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(sheets, o * 4, 0, 0));
    t.bot = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(sheets, (o + side) * 4, 0, 0));
}

__device__ __forceinline__ float taps_blend(const TapsM& t) {
    float top = lerpf(t.top.x, t.top.y, t.tj), bot = lerpf(t.bot.x, t.bot.y, t.tj);
    return lerpf(top, bot, t.ti);
}

__device__ __forceinline__ float taps_blend(const Taps<float>& t) {
    float top = lerpf(t.a, t.b, t.tj), bot = lerpf(t.c, t.d, t.tj);
    return lerpf(top, bot, t.ti);
}

__device__ __forceinline__ float4 taps_blend(const Taps<float4>& t) {
    const float4 a = t.a, b = t.b, c = t.c, d = t.d;
    const float tj = t.tj, ti = t.ti;
    float4 top = make_float4(lerpf(a.x, b.x, tj), lerpf(a.y, b.y, tj), lerpf(a.z, b.z, tj), lerpf(a.w, b.w, tj));
    float4 bot = make_float4(lerpf(c.x, d.x, tj), lerpf(c.y, d.y, tj), lerpf(c.z, d.z, tj), lerpf(c.w, d.w, tj));
    return make_float4(lerpf(top.x, bot.x, ti), lerpf(top.y, bot.y, ti), lerpf(top.z, bot.z, ti), lerpf(top.w, bot.w, ti));
}

// a + t (b - a) as a - t a + t b: two FMAs on values converted exactly from half precision (the
// compiler may read the halves in place with mixed-precision FMAs).
__device__ __forceinline__ float lerp_m(float a, float b, float t) { return __builtin_fmaf(t, b, __builtin_fmaf(-t, a, a)); }

__device__ __forceinline__ float4 taps_blend(const TapsH& t) {
    // left: texel o then the one below; right: texel o + 1 then the one below
    const float2 a0 = unpack_half2(t.left.x), a1 = unpack_half2(t.left.y);
    const float2 c0 = unpack_half2(t.left.z), c1 = unpack_half2(t.left.w);
    const float2 b0 = unpack_half2(t.right.x), b1 = unpack_half2(t.right.y);
    const float2 d0 = unpack_half2(t.right.z), d1 = unpack_half2(t.right.w);
    const float tj = t.tj, ti = t.ti;
    const float4 top = make_float4(lerp_m(a0.x, b0.x, tj), lerp_m(a0.y, b0.y, tj), lerp_m(a1.x, b1.x, tj), lerp_m(a1.y, b1.y, tj));
    const float4 bot = make_float4(lerp_m(c0.x, d0.x, tj), lerp_m(c0.y, d0.y, tj), lerp_m(c1.x, d1.x, tj), lerp_m(c1.y, d1.y, tj));
    return make_float4(lerpf(top.x, bot.x, ti), lerpf(top.y, bot.y, ti), lerpf(top.z, bot.z, ti), lerpf(top.w, bot.w, ti));
}

template <>
__device__ __forceinline__ float4 sheet_bilinear<float4>(const float4* __restrict__ s, int side, float3 v) {
    Taps<float4> t;
    sheet_taps<float4>(s, side, v, t);
    return taps_blend(t);
}

// ── complex helpers ──
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (-i)
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }
// i^p x
__device__ __forceinline__ float2 ipow(float2 x, int p) {
    switch (p & 3) {
        case 1: return make_float2(-x.y, x.x);
        case 2: return make_float2(-x.x, -x.y);
        case 3: return make_float2(x.y, -x.x);
        default: return x;
    }
}

__device__ __forceinline__ int bitrev(int x, int bits) { return (int)(__brev((unsigned)x) >> (32 - bits)); }

// In-place forward (e^{-i...}) radix-2 decimation-in-frequency FFT of `nlines` lines of 2^log2n
// points held in LDS: point p of line l at buf[l * lstride + p * pstride]. Natural order in,
// bit-reversed order out. Stages run in pairs as radix-4 butterflies (a thread takes four points
// through both stages in registers); an odd last stage runs alone. All NT threads of the block
// call it; it ends with a barrier.
template <int NT>
__device__ void fft_dif(float2* buf, int log2n, int nlines, int lstride, int pstride, const float2* __restrict__ tw, int tid) {
    int s = log2n;
    const int nq = 1 << (log2n - 2 < 0 ? 0 : log2n - 2);
    for (; s >= 2; s -= 2) {
        const int lq = s - 2;
        const int q = 1 << lq;
        const int t4 = ZB_TW >> s;        // W_{4q}^j  = tw[j * t4]
        const int t2 = ZB_TW >> (s - 1);  // W_{2q}^j  = tw[j * t2]
        const int total = nlines * nq;
        for (int e = tid; e < total; e += NT) {
            int line = e >> (log2n - 2);
            int r = e & (nq - 1);
            int j = r & (q - 1);
            int i0 = ((r >> lq) << s) + j;
            float2* b = buf + line * lstride;
            float2 x0 = b[i0 * pstride];
            float2 x1 = b[(i0 + q) * pstride];
            float2 x2 = b[(i0 + 2 * q) * pstride];
            float2 x3 = b[(i0 + 3 * q) * pstride];
            float2 w4 = tw[j * t4];
            float2 w2 = tw[j * t2];
            float2 a0 = cadd(x0, x2);
            float2 a2 = cmul(csub(x0, x2), w4);
            float2 a1 = cadd(x1, x3);
            float2 a3 = cmul(cmul_mi(csub(x1, x3)), w4);
            b[i0 * pstride] = cadd(a0, a1);
            b[(i0 + q) * pstride] = cmul(csub(a0, a1), w2);
            b[(i0 + 2 * q) * pstride] = cadd(a2, a3);
            b[(i0 + 3 * q) * pstride] = cmul(csub(a2, a3), w2);
        }
        __syncthreads();
    }
    if (s == 1) {
        const int half = 1 << (log2n - 1);
        const int total = nlines * half;
        for (int e = tid; e < total; e += NT) {
            int line = e >> (log2n - 1);
            int r = e & (half - 1);
            float2* b = buf + line * lstride;
            float2 x0 = b[(2 * r) * pstride];
            float2 x1 = b[(2 * r + 1) * pstride];
            b[(2 * r) * pstride] = cadd(x0, x1);
            b[(2 * r + 1) * pstride] = csub(x0, x1);
        }
        __syncthreads();
    }
}

// fft_dif for contiguous lines (pstride 1, lstride 2^LOG2N) of a size known at compile time:
// the same butterflies in the same order (bit-identical results), with the stage sizes, strides
// and twiddle steps constants, so that the index arithmetic folds away.
template <int NT, int LOG2N, int S>
__device__ __forceinline__ void fft_dif_stages(float2* buf, int nlines, const float2* __restrict__ tw, int tid) {
    if constexpr (S >= 2) {
        constexpr int lq = S - 2, q = 1 << lq, nq = 1 << (LOG2N - 2);
        constexpr int t4 = ZB_TW >> S, t2 = ZB_TW >> (S - 1);
        const int total = nlines * nq;
        for (int e = tid; e < total; e += NT) {
            const int line = e >> (LOG2N - 2);
            const int r = e & (nq - 1);
            const int j = r & (q - 1);
            float2* b = buf + (line << LOG2N) + ((r >> lq) << S) + j;
            float2 x0 = b[0];
            float2 x1 = b[q];
            float2 x2 = b[2 * q];
            float2 x3 = b[3 * q];
            float2 w4 = tw[j * t4];
            float2 w2 = tw[j * t2];
            float2 a0 = cadd(x0, x2);
            float2 a2 = cmul(csub(x0, x2), w4);
            float2 a1 = cadd(x1, x3);
            float2 a3 = cmul(cmul_mi(csub(x1, x3)), w4);
            b[0] = cadd(a0, a1);
            b[q] = cmul(csub(a0, a1), w2);
            b[2 * q] = cadd(a2, a3);
            b[3 * q] = cmul(csub(a2, a3), w2);
        }
        __syncthreads();
        fft_dif_stages<NT, LOG2N, S - 2>(buf, nlines, tw, tid);
    } else if constexpr (S == 1) {
        constexpr int half = 1 << (LOG2N - 1);
        const int total = nlines * half;
        for (int e = tid; e < total; e += NT) {
            const int line = e >> (LOG2N - 1);
            const int r = e & (half - 1);
            float2* b = buf + (line << LOG2N) + 2 * r;
            float2 x0 = b[0];
            float2 x1 = b[1];
            b[0] = cadd(x0, x1);
            b[1] = csub(x0, x1);
        }
        __syncthreads();
    }
}

template <int NT, int LOG2N>
__device__ void fft_dif_fixed(float2* buf, int nlines, const float2* __restrict__ tw, int tid) {
    fft_dif_stages<NT, LOG2N, LOG2N>(buf, nlines, tw, tid);
}

__device__ __forceinline__ float warp_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace zb
