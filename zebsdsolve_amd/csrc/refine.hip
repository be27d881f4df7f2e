// Orientation refinement on the detector, fused: one wave (64 lanes) per item runs the seed, every
// Gauss-Newton iteration, the normalized mutual information and the record, in one launch
// (the reference's detector.wgsl seed, refine_init, sums_rot + refine_rot_step per iteration,
// nmi and collect; src/detector.zig:178-212).
//
// Per iteration each lane adds up its pixels' share of the 19 sums the step needs (1, M, M^2,
// M e, d, M d, e d and the upper triangle of d d', with d = c x grad M(c) for c = q . ray); an
// xor-butterfly over the wave leaves every lane with the same totals (a + b == b + a bit for
// bit), so every lane takes the same step with no broadcast and no barrier. The master's
// gradient field holds (grad M, M) per texel: one bilinear lookup reads both.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "common.h"

namespace zb {

struct RefArg {
    float c0[3], c1[3], c2[3];
    float mask[3];
    int h, w, side;
    int iters, nmi_bins, seed_stride, out0;
    float max_step, scale0, scale_min, scale_max;
    float stamp;
    // 1: the masked pixels as a list of 16-bit indices in LDS (frames up to 65 536 pixels whose
    // list fits); 0: the whole frame walked, each pixel tested against the mask (larger frames)
    int use_list;
};

struct Sums {
    float n, m, mm, me;
    float3 sd, sv, se, hd, ho;
};

__device__ __forceinline__ float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 f3s(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }

__device__ __forceinline__ float wsum(float v) {
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float3 wsum3(float3 v) { return make_float3(wsum(v.x), wsum(v.y), wsum(v.z)); }

// A value every lane holds alike, read from the first lane into a scalar register.
__device__ __forceinline__ float ufloat(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ float4 uniform4(float4 v) { return make_float4(ufloat(v.x), ufloat(v.y), ufloat(v.z), ufloat(v.w)); }

// The masked pixels of the frame as pixel indices, built by each block into LDS (the same list
// for every item): the passes then walk only the pixels under the mask, 64 at a time, instead of
// the whole frame with a quarter of the lanes idle. The order is by 8 x 8 tiles of the frame (in
// row-major tile order, each tile's pixels row-major), so that the 64 lanes of a pass look up a
// compact patch of the detector, whose rays land on nearby texels of the master, rather than a
// row segment, whose rays trace a long curve across it: fewer cache lines per gather (the
// refinement 6.8 % faster than with row-major order).
__device__ int build_list(const RefArg& a, unsigned short* list, int lane) {
    const int tw = (a.w + 7) >> 3, th = (a.h + 7) >> 3, nt = tw * th;
    auto inside = [&](int r, int c) {
        const float dr = (float)r - a.mask[0], dc = (float)c - a.mask[1];
        return a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2];
    };
    int total = 0;
    for (int t0 = 0; t0 < nt; t0 += 64) {
        // lane t counts its tile's masked pixels; an exclusive prefix over the 64 tiles places them
        const int t = t0 + lane;
        const int r0 = (t / tw) * 8, c0 = (t % tw) * 8;
        int cnt = 0;
        if (t < nt)
            for (int r = r0; r < min(r0 + 8, a.h); ++r)
                for (int c = c0; c < min(c0 + 8, a.w); ++c) cnt += inside(r, c) ? 1 : 0;
        int inc = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        int at = total + inc - cnt;
        if (t < nt)
            for (int r = r0; r < min(r0 + 8, a.h); ++r)
                for (int c = c0; c < min(c0 + 8, a.w); ++c)
                    if (inside(r, c)) list[at++] = (unsigned short)(r * a.w + c);
        total += __shfl(inc, 63, 64);
    }
    __syncthreads();
    return total;
}

// Pixel k of a pass: the list's k-th (the masked pixels, in LDS) or, without a list, the
// frame's k-th; `ok` says whether it counts (k inside the pass, the pixel under the mask). Past
// the end the last pixel is looked up again and counts for nothing.
template <bool LIST>
__device__ __forceinline__ int pass_pixel(const RefArg& a, const unsigned short* list, int n_px, int k, bool& ok) {
    const int kk = min(k, n_px - 1);
    if constexpr (LIST) {
        ok = k < n_px;
        return list[kk];
    }
    const int row = kk / a.w;
    const float dr = (float)row - a.mask[0], dc = (float)(kk - row * a.w) - a.mask[1];
    ok = k < n_px && (a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2]);
    return kk;
}

// The ray of pixel `pix` (row-major index) from the item's origin t, in the sample frame; the
// row by a float product below 65 536 pixels (exact there), by division above.
template <bool LIST>
__device__ __forceinline__ float3 pixel_ray(const RefArg& a, int pix, float inv_w, float3 t) {
    const int row = LIST ? (int)(((float)pix + 0.5f) * inv_w) : pix / a.w;
    const float fr = (float)row, fc = (float)(pix - row * a.w);
    return make_float3(a.c0[0] * fr + a.c1[0] * fc + t.x, a.c0[1] * fr + a.c1[1] * fc + t.y, a.c0[2] * fr + a.c1[2] * fc + t.z);
}

// Pixels per lane whose gathers are in flight together: a lane issues the loads of kBatch pixels
// (3 each) before it blends any.
constexpr int kBatch = 4;

template <bool LIST>
__device__ Sums rot_sums(const RefArg& a, __amdgpu_buffer_rsrc_t mgh, const float* __restrict__ pat, const unsigned short* list, int n_px,
                         float inv_w, float4 q, float3 t, int lane) {
    Sums s;
    s.n = s.m = s.mm = s.me = 0.f;
    s.sd = s.sv = s.se = s.hd = s.ho = make_float3(0.f, 0.f, 0.f);
    for (int k0 = lane; k0 < n_px; k0 += 64 * kBatch) {
        TapsH tp[kBatch];
        float3 c[kBatch];
        float e[kBatch];
        bool ok[kBatch];
#pragma unroll
        for (int p = 0; p < kBatch; ++p) {
            // lanes past the end (or off the mask) look up a pixel and add nothing
            const int pix = pass_pixel<LIST>(a, list, n_px, k0 + 64 * p, ok[p]);
            c[p] = qrot(q, normalize3(pixel_ray<LIST>(a, pix, inv_w, t)));
            sheet_taps_h(mgh, a.side, c[p], tp[p]);
            e[p] = pat[pix];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int p = 0; p < kBatch; ++p) {
            if (ok[p]) {
                const float4 m = taps_blend(tp[p]);
                const float3 d = cross3(c[p], make_float3(m.x, m.y, m.z));
                s.n += 1.f;
                s.m += m.w;
                s.mm += m.w * m.w;
                s.me += m.w * e[p];
                s.sd = f3add(s.sd, d);
                s.sv = f3add(s.sv, f3s(d, m.w));
                s.se = f3add(s.se, f3s(d, e[p]));
                s.hd = f3add(s.hd, make_float3(d.x * d.x, d.y * d.y, d.z * d.z));
                s.ho = f3add(s.ho, make_float3(d.x * d.y, d.x * d.z, d.y * d.z));
            }
        }
    }
    s.n = wsum(s.n);
    s.m = wsum(s.m);
    s.mm = wsum(s.mm);
    s.me = wsum(s.me);
    s.sd = wsum3(s.sd);
    s.sv = wsum3(s.sv);
    s.se = wsum3(s.se);
    s.hd = wsum3(s.hd);
    s.ho = wsum3(s.ho);
    return s;
}

// The master's values at orientation q for the list pixels k0 + 64 p, p < kBatch (lanes past the
// end get the last pixel's), with the pattern's values there: all gathers issued before any
// blend. msh: a buffer resource over the float32 sheets.
template <bool LIST>
__device__ __forceinline__ void sims_batch(const RefArg& a, __amdgpu_buffer_rsrc_t msh, const float* __restrict__ pat,
                                           const unsigned short* list, int n_px, float inv_w, float4 q, float3 t, int k0,
                                           float* sim, float* e, bool* ok) {
    TapsM tp[kBatch];
#pragma unroll
    for (int p = 0; p < kBatch; ++p) {
        const int pix = pass_pixel<LIST>(a, list, n_px, k0 + 64 * p, ok[p]);
        sheet_taps_m(msh, a.side, qrot(q, normalize3(pixel_ray<LIST>(a, pix, inv_w, t))), tp[p]);
        e[p] = pat[pix];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int p = 0; p < kBatch; ++p) sim[p] = taps_blend(tp[p]);
}

__device__ __forceinline__ float det3(float3 c0, float3 c1, float3 c2) {
    return c0.x * (c1.y * c2.z - c2.y * c1.z) - c1.x * (c0.y * c2.z - c2.y * c0.z) + c2.x * (c0.y * c1.z - c1.y * c0.z);
}

// The symmetric 3 x 3 system (h: xx, xy, xz, yy, yz, zz) by Cramer's rule (detector.wgsl:519-527).
__device__ __forceinline__ float3 solve3(const float* h, float3 g) {
    const float3 m0 = make_float3(h[0], h[1], h[2]);
    const float3 m1 = make_float3(h[1], h[3], h[4]);
    const float3 m2 = make_float3(h[2], h[4], h[5]);
    const float det = det3(m0, m1, m2);
    if (fabsf(det) < 1e-30f) return make_float3(0.f, 0.f, 0.f);
    return make_float3(det3(g, m1, m2) / det, det3(m0, g, m2) / det, det3(m0, m1, g) / det);
}

// One item's refinement, NMI and record (at records[out0 + b]), with the masked-pixel list (LIST)
// already built and the histogram cleared.
// The Gauss-Newton passes read the half-precision gradient field (mgh: two adjacent 16-byte loads
// a lookup, each a texel and the one below it; half the vector-memory instructions of float32
// texels); the recorded NCC and
// the NMI are those of the float32 master (ms, the values the field holds in .w) at the best
// orientation, from the NMI's first pass (one 8-byte load per texel row pair).
// MAXP > 0: each lane keeps the master values of its (at most MAXP) pixels at the best
// orientation from the first NMI pass for the second; MAXP = 0: the second pass looks them up
// again.
template <int MAXP, bool LIST>
__device__ __forceinline__ void refine_item(const RefArg& a, const float* __restrict__ ms, const uint2* __restrict__ mgh, const float* __restrict__ pats,
                                            const float4* __restrict__ items, const float* __restrict__ seeds, int seed_stride,
                                            float* __restrict__ records, int out0, unsigned* hist, const unsigned short* list, int n_px, int b,
                                            int lane) {
    static_assert(MAXP % kBatch == 0, "MAXP: whole batches");
    const float inv_w = 1.f / (float)a.w;
    // the item's values are the same in every lane: held in scalar registers
    const float4 it = uniform4(items[b]);
    const float3 t = make_float3(it.x, it.y, it.z);
    const float* pat = pats + (size_t)(int)it.w * a.h * a.w;
    const float* sd = seeds + (size_t)b * seed_stride;
    float4 cur = uniform4(make_float4(sd[0], sd[1], sd[2], sd[3]));

    // refinement state (detector.wgsl refine_init)
    float4 qbest = cur;
    float best = -2.f;
    float mu = a.scale0;
    float3 gbest = make_float3(0.f, 0.f, 0.f);
    float hbest[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float last_step = 0.f, promised = 0.f, step_size = 0.f, accepted = 0.f;
    const __amdgpu_buffer_rsrc_t fh = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint2*>(mgh), (short)0, 2 * a.side * a.side * 16, 0x00020000);
    const __amdgpu_buffer_rsrc_t fm = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(ms), (short)0, 2 * a.side * a.side * 4, 0x00020000);

    for (int itn = 0; itn <= a.iters; ++itn) {
        const Sums s = rot_sums<LIST>(a, fh, pat, list, n_px, inv_w, cur, t, lane);
        // NCC, its gradient and Gauss-Newton matrix over the rotation (detector.wgsl:561-583)
        const float n = fmaxf(s.n, 1.f);
        const float mean = s.m / n;
        const float var_n = s.mm - n * mean * mean;
        float ncc = -2.f;
        float3 grad = make_float3(0.f, 0.f, 0.f);
        float hess[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (var_n > 1e-12f) {
            const float inv = 1.f / sqrtf(var_n);
            ncc = s.me * inv;
            const float3 au = f3s(make_float3(s.sv.x - mean * s.sd.x, s.sv.y - mean * s.sd.y, s.sv.z - mean * s.sd.z), inv);
            grad = f3s(make_float3(s.se.x - ncc * au.x, s.se.y - ncc * au.y, s.se.z - ncc * au.z), inv);
            const float iv = inv * inv;
            hess[0] = (s.hd.x - s.sd.x * s.sd.x / n - au.x * au.x) * iv;
            hess[1] = (s.ho.x - s.sd.x * s.sd.y / n - au.x * au.y) * iv;
            hess[2] = (s.ho.y - s.sd.x * s.sd.z / n - au.x * au.z) * iv;
            hess[3] = (s.hd.y - s.sd.y * s.sd.y / n - au.y * au.y) * iv;
            hess[4] = (s.ho.z - s.sd.y * s.sd.z / n - au.y * au.z) * iv;
            hess[5] = (s.hd.z - s.sd.z * s.sd.z / n - au.z * au.z) * iv;
        }
        // accept against the best so far, rescaling the curvature by gain / promise
        if (ncc > best) {
            if (accepted > 0.5f && promised > 1e-6f) {
                const float rho = (ncc - best) / promised;
                mu = fminf(fmaxf(mu * (2.f - rho), mu / 3.f), mu * 3.f);
            }
            qbest = cur;
            best = ncc;
            gbest = grad;
#pragma unroll
            for (int i = 0; i < 6; ++i) hbest[i] = hess[i];
            last_step = step_size;
            accepted += 1.f;
        } else {
            mu = mu * 3.f;
        }
        mu = fminf(fmaxf(mu, a.scale_min), a.scale_max);
        if (itn == a.iters) break;
        float hh[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) hh[i] = hbest[i] * mu;
        const float tr = fmaxf(hh[0] + hh[3] + hh[5], 1e-20f);
        hh[0] += 1e-6f * tr;
        hh[3] += 1e-6f * tr;
        hh[5] += 1e-6f * tr;
        float3 st = solve3(hh, gbest);
        const float len = sqrtf(dot3(st, st));
        if (len > a.max_step) st = f3s(st, a.max_step / len);
        const float quad = hh[0] * st.x * st.x + hh[3] * st.y * st.y + hh[5] * st.z * st.z +
                           2.f * (hh[1] * st.x * st.y + hh[2] * st.x * st.z + hh[4] * st.y * st.z);
        promised = dot3(gbest, st) - 0.5f * quad;
        step_size = sqrtf(dot3(st, st));
        cur = qnormalize(qmul(qexp(st), qbest));
    }

    // normalized mutual information of the pattern and the master at the best orientation,
    // each binned between its own extremes under the mask (detector.wgsl:739-832)
    const int nb = a.nmi_bins;
    float emin = 3.4e38f, emax = -3.4e38f, smin = 3.4e38f, smax = -3.4e38f;
    float f_n = 0.f, f_m = 0.f, f_mm = 0.f, f_me = 0.f;  // the NCC's sums there (float32 master)
    auto extremes = [&](float e, float sim) {
        emin = fminf(emin, e);
        emax = fmaxf(emax, e);
        smin = fminf(smin, sim);
        smax = fmaxf(smax, sim);
        f_n += 1.f;
        f_m += sim;
        f_mm += sim * sim;
        f_me += sim * e;
    };
    float simv[MAXP > 0 ? MAXP : 1];
    if constexpr (MAXP > 0) {
#pragma unroll
        for (int i0 = 0; i0 < MAXP; i0 += kBatch) {
            float sim[kBatch] = {}, e[kBatch] = {};
            bool ok[kBatch] = {};
            if (64 * i0 < n_px) sims_batch<LIST>(a, fm, pat, list, n_px, inv_w, qbest, t, lane + 64 * i0, sim, e, ok);
#pragma unroll
            for (int p = 0; p < kBatch; ++p) {
                simv[i0 + p] = sim[p];
                if (ok[p]) extremes(e[p], sim[p]);
            }
        }
    } else {
        for (int k0 = lane; k0 < n_px; k0 += 64 * kBatch) {
            float sim[kBatch], e[kBatch];
            bool ok[kBatch];
            sims_batch<LIST>(a, fm, pat, list, n_px, inv_w, qbest, t, k0, sim, e, ok);
#pragma unroll
            for (int p = 0; p < kBatch; ++p)
                if (ok[p]) extremes(e[p], sim[p]);
        }
    }
    for (int o = 1; o < 64; o <<= 1) {
        emin = fminf(emin, __shfl_xor(emin, o, 64));
        emax = fmaxf(emax, __shfl_xor(emax, o, 64));
        smin = fminf(smin, __shfl_xor(smin, o, 64));
        smax = fmaxf(smax, __shfl_xor(smax, o, 64));
    }
    // the recorded NCC: that of the float32 master at the best orientation (the rule of the
    // Gauss-Newton passes)
    float ncc_best = -2.f;
    {
        const float n = fmaxf(wsum(f_n), 1.f);
        const float mean = wsum(f_m) / n;
        const float var_n = wsum(f_mm) - n * mean * mean;
        const float me = wsum(f_me);
        if (var_n > 1e-12f) ncc_best = me * (1.f / sqrtf(var_n));
    }
    const float ke = (float)nb / fmaxf(emax - emin, 1e-20f);
    const float ks = (float)nb / fmaxf(smax - smin, 1e-20f);
    auto bin = [&](float e, float sim) {
        const int be = min((int)((e - emin) * ke), nb - 1);
        const int bs = min((int)((sim - smin) * ks), nb - 1);
        atomicAdd(&hist[be * nb + bs], 1u);
    };
    if constexpr (MAXP > 0) {
#pragma unroll
        for (int i0 = 0; i0 < MAXP; i0 += kBatch) {
            if (64 * i0 < n_px) {
                float e[kBatch];
                bool ok[kBatch];
#pragma unroll
                for (int p = 0; p < kBatch; ++p) e[p] = pat[pass_pixel<LIST>(a, list, n_px, lane + 64 * (i0 + p), ok[p])];
#pragma unroll
                for (int p = 0; p < kBatch; ++p)
                    if (ok[p]) bin(e[p], simv[i0 + p]);
            }
        }
    } else {
        for (int k0 = lane; k0 < n_px; k0 += 64 * kBatch) {
            float sim[kBatch], e[kBatch];
            bool ok[kBatch];
            sims_batch<LIST>(a, fm, pat, list, n_px, inv_w, qbest, t, k0, sim, e, ok);
#pragma unroll
            for (int p = 0; p < kBatch; ++p)
                if (ok[p]) bin(e[p], sim[p]);
        }
    }
    __syncthreads();
    float hj = 0.f, tot = 0.f, hm = 0.f;
    for (int i = lane; i < nb * nb; i += 64) {
        const float c = (float)hist[i];
        tot += c;
        if (c > 0.f) hj -= c * logf(c);
    }
    if (lane < nb) {
        float rs = 0.f, cs = 0.f;
        for (int k = 0; k < nb; ++k) {
            rs += (float)hist[lane * nb + k];
            cs += (float)hist[k * nb + lane];
        }
        if (rs > 0.f) hm -= rs * logf(rs);
        if (cs > 0.f) hm -= cs * logf(cs);
    }
    hj = wsum(hj);
    tot = wsum(tot);
    hm = wsum(hm);
    if (lane == 0) {
        const float nn = fmaxf(tot, 1.f);
        const float h_joint = logf(nn) + hj / nn;
        const float h_marg = 2.f * logf(nn) + hm / nn;
        const float nmi = h_marg > 1e-12f ? 2.f * (h_marg - h_joint) / h_marg : 0.f;
        float* r = records + (size_t)(out0 + b) * 8;
        r[0] = qbest.x;
        r[1] = qbest.y;
        r[2] = qbest.z;
        r[3] = qbest.w;
        r[4] = ncc_best;
        r[5] = nmi;
        r[6] = last_step;
        r[7] = a.stamp;
    }
}

// One block per item. With `order` (zb_refine_order's permutation), block i refines item
// order[xcd_block(i)]: the blocks of one XCD take a contiguous run of items sorted by the master
// region they look up, so that their gathers share that XCD's L2. Without it, block i refines
// item i.
template <int MAXP, bool LIST>
__global__ void __launch_bounds__(64) k_refine(RefArg a, const float* __restrict__ ms, const uint2* __restrict__ mgh, const float* __restrict__ pats,
                                               const float4* __restrict__ items, const float* __restrict__ seeds, int n_items,
                                               const int* __restrict__ order, float* __restrict__ records) {
    extern __shared__ unsigned lds[];
    unsigned* hist = lds;                                               // [nmi_bins][nmi_bins]
    unsigned short* list = LIST ? (unsigned short*)(lds + a.nmi_bins * a.nmi_bins) : nullptr;  // the masked pixels
    const int b = order ? order[xcd_block(blockIdx.x, gridDim.x)] : (int)blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < a.nmi_bins * a.nmi_bins; i += 64) hist[i] = 0u;
    int n_px = a.h * a.w;
    if constexpr (LIST) n_px = build_list(a, list, lane);
    else __syncthreads();
    refine_item<MAXP, LIST>(a, ms, mgh, pats, items, seeds, a.seed_stride, records, a.out0, hist, list, n_px, b, lane);
}

// Several phases' refinements of one batch in one launch. The n_phases x n_items work items are
// taken phase-major (item k of phase p is p * n_items + k, in that phase's order).
// k_refine_grid: one block per work item, block i taking work item xcd_block(i), so that each XCD
// takes a contiguous run of the sequence; the hardware deals blocks to CUs as others finish.
// k_refine_queue: a persistent grid (as many blocks as stay resident) pulling work items from
// device-side queues: the sequence cut into kQueues contiguous runs, one per XCD, each with its
// own counter; a block first drains its own XCD's run (block id mod 8, the round-robin dealing
// xcd_block assumes), then the other runs in turn, and leaves when all are empty, so that every
// block reaches that exit. Each block builds the masked-pixel list once for all its items.
// The phases' array pointers are separate restrict kernel arguments (so that the compiler keeps
// k_refine's register budget: without the no-alias guarantee it holds more values live); their
// sides, seed strides and orders come as one more argument to the grid form, and for the queue
// form sit in a table in the work buffer after the queue counters, written with them by
// k_queue_setup on the same stream before the launch.
constexpr int kMaxRefinePhases = 4;
constexpr int kQueues = 8;
constexpr int kQueueStride = 16;
struct RefPhase {
    const int* order;
    int side, seed_stride;
};
struct RefPhases {
    RefPhase p[kMaxRefinePhases];
};
constexpr int kTableOffset = kQueues * kQueueStride * 4;
static_assert(kTableOffset + sizeof(RefPhases) <= ZB_REFINE_QUEUE_BYTES, "the counters and the phase table fit the work buffer");

// the phases' master, half-precision field, seeds and records, as restrict arguments
#define ZB_PHASE_ARGS(i) \
    const float *__restrict__ ms##i, const uint2 *__restrict__ mgh##i, const float *__restrict__ sd##i, float *__restrict__ rec##i
template <typename T>
__device__ __forceinline__ T pick(int p, T x0, T x1, T x2, T x3) {
    return p == 0 ? x0 : p == 1 ? x1 : p == 2 ? x2 : x3;
}

// Work item w of the sequence: phase p, item b (in that phase's order); then the item refined.
template <int MAXP, bool LIST>
__device__ __forceinline__ void refine_work(const RefArg& a, RefPhase e, int p, int i, ZB_PHASE_ARGS(0), ZB_PHASE_ARGS(1), ZB_PHASE_ARGS(2),
                                            ZB_PHASE_ARGS(3), const float* __restrict__ pats, const float4* __restrict__ items, unsigned* hist,
                                            const unsigned short* list, int n_px, int lane) {
    const int* order = e.order;
    const int b = order ? __builtin_amdgcn_readfirstlane(order[i]) : i;
    RefArg ap = a;
    ap.side = __builtin_amdgcn_readfirstlane(e.side);
    refine_item<MAXP, LIST>(ap, pick(p, ms0, ms1, ms2, ms3), pick(p, mgh0, mgh1, mgh2, mgh3), pats, items, pick(p, sd0, sd1, sd2, sd3),
                            __builtin_amdgcn_readfirstlane(e.seed_stride), pick(p, rec0, rec1, rec2, rec3), a.out0, hist, list, n_px, b, lane);
}

// Work item w of the sequence: phase p and item i (in that phase's order).
__device__ __forceinline__ void work_item(unsigned w, int n_items, int& p, int& i) {
    p = __builtin_amdgcn_readfirstlane((int)(w / (unsigned)n_items));
    i = __builtin_amdgcn_readfirstlane((int)(w - (unsigned)p * (unsigned)n_items));
}

template <int MAXP, bool LIST>
__global__ void __launch_bounds__(64) k_refine_grid(RefArg a, ZB_PHASE_ARGS(0), ZB_PHASE_ARGS(1), ZB_PHASE_ARGS(2), ZB_PHASE_ARGS(3),
                                                    const float* __restrict__ pats, const float4* __restrict__ items, int n_items,
                                                    RefPhases ph) {
    extern __shared__ unsigned lds[];
    unsigned* hist = lds;
    unsigned short* list = LIST ? (unsigned short*)(lds + a.nmi_bins * a.nmi_bins) : nullptr;
    const int lane = threadIdx.x;
    for (int j = lane; j < a.nmi_bins * a.nmi_bins; j += 64) hist[j] = 0u;
    int n_px = a.h * a.w;
    if constexpr (LIST) n_px = build_list(a, list, lane);
    else __syncthreads();
    int p, i;
    work_item((unsigned)xcd_block(blockIdx.x, gridDim.x), n_items, p, i);
    // the phase's entry picked from the argument (an index into it would place it in memory)
    const RefPhase e = pick(p, ph.p[0], ph.p[1], ph.p[2], ph.p[3]);
    refine_work<MAXP, LIST>(a, e, p, i, ms0, mgh0, sd0, rec0, ms1, mgh1, sd1, rec1, ms2, mgh2, sd2, rec2, ms3, mgh3, sd3, rec3, pats, items, hist,
                            list, n_px, lane);
}

template <int MAXP, bool LIST>
__global__ void __launch_bounds__(64) k_refine_queue(RefArg a, ZB_PHASE_ARGS(0), ZB_PHASE_ARGS(1), ZB_PHASE_ARGS(2), ZB_PHASE_ARGS(3),
                                                     int n_phases, const float* __restrict__ pats, const float4* __restrict__ items, int n_items,
                                                     unsigned* __restrict__ work) {
    extern __shared__ unsigned lds[];
    unsigned* hist = lds;
    unsigned short* list = LIST ? (unsigned short*)(lds + a.nmi_bins * a.nmi_bins) : nullptr;
    const RefPhase* table = (const RefPhase*)((const char*)work + kTableOffset);
    const int lane = threadIdx.x;
    int n_px = a.h * a.w;
    if constexpr (LIST) n_px = build_list(a, list, lane);
    const unsigned total = (unsigned)n_items * (unsigned)n_phases;  // below 2^31 (the entry point's check)
    for (int r = 0; r < kQueues; ++r) {
        const unsigned x = (blockIdx.x + r) & (kQueues - 1);
        const unsigned lo = (unsigned)((unsigned long long)total * x / kQueues);
        const unsigned len = (unsigned)((unsigned long long)total * (x + 1) / kQueues) - lo;
        for (;;) {
            unsigned k = 0;
            if (lane == 0) k = atomicAdd(work + x * kQueueStride, 1u);
            k = __builtin_amdgcn_readfirstlane(k);
            if (k >= len) break;
            for (int j = lane; j < a.nmi_bins * a.nmi_bins; j += 64) hist[j] = 0u;
            __syncthreads();
            int p, i;
            work_item(lo + k, n_items, p, i);
            refine_work<MAXP, LIST>(a, table[p], p, i, ms0, mgh0, sd0, rec0, ms1, mgh1, sd1, rec1, ms2, mgh2, sd2, rec2, ms3, mgh3, sd3, rec3, pats,
                                    items, hist, list, n_px, lane);
            __syncthreads();  // every lane done with the histogram before the next item clears it
        }
    }
}

// The queue counters cleared and the phase table written into the work buffer.
__global__ void k_queue_setup(RefPhases ph, int n_phases, unsigned* work) {
    const int t = threadIdx.x;
    if (t < kQueues) work[t * kQueueStride] = 0u;
    if (t < n_phases) ((RefPhase*)((char*)work + kTableOffset))[t] = ph.p[t];
}

// Cells of the refinement's order: each sheet of the master cut into 32 x 32 Lambert cells
constexpr int kOrderSide = 32;
constexpr int kOrderCells = 2 * kOrderSide * kOrderSide;

// The cell the detector centre's ray lands in on the master at orientation q (item position t),
// numbered by sheet and then along a Morton curve over the cell rows and columns, so that nearby
// cells get nearby numbers.
__device__ __forceinline__ int order_cell(const RefArg& a, float4 q, float3 t) {
    const float fr = 0.5f * (float)(a.h - 1), fc = 0.5f * (float)(a.w - 1);
    const float3 u = normalize3(make_float3(a.c0[0] * fr + a.c1[0] * fc + t.x, a.c0[1] * fr + a.c1[1] * fc + t.y,
                                            a.c0[2] * fr + a.c1[2] * fc + t.z));
    const float3 v = qrot(q, u);
    const float2 uv = sphere_to_square(v);
    const int col = min(max((int)((uv.x + 1.f) * (0.5f * kOrderSide)), 0), kOrderSide - 1);
    const int row = min(max((int)((uv.y + 1.f) * (0.5f * kOrderSide)), 0), kOrderSide - 1);
    int m = 0;
#pragma unroll
    for (int bit = 0; bit < 5; ++bit) m |= (((col >> bit) & 1) << (2 * bit)) | (((row >> bit) & 1) << (2 * bit + 1));
    return (v.z < 0.f ? kOrderSide * kOrderSide : 0) + m;
}

// A counting sort of the items by order_cell, one block: counts per cell in LDS, their exclusive
// prefix, then each item placed at its cell's next slot (the order within a cell is arbitrary;
// every item's record is the same whichever block computes it).
__global__ void __launch_bounds__(1024) k_order(RefArg a, const float4* __restrict__ items, const float* __restrict__ seeds, int seed_stride,
                                                int n_items, int* __restrict__ order) {
    __shared__ unsigned cnt[kOrderCells];
    __shared__ unsigned part[1024];
    const int tid = threadIdx.x;
    for (int i = tid; i < kOrderCells; i += 1024) cnt[i] = 0u;
    __syncthreads();
    auto cell_of = [&](int b) {
        const float* sd = seeds + (size_t)b * seed_stride;
        const float4 it = items[b];
        return order_cell(a, make_float4(sd[0], sd[1], sd[2], sd[3]), make_float3(it.x, it.y, it.z));
    };
    for (int b = tid; b < n_items; b += 1024) atomicAdd(&cnt[cell_of(b)], 1u);
    __syncthreads();
    // exclusive prefix over the cells: each thread sums its run of cells, a scan over the threads,
    // then the run's own offsets
    constexpr int per = kOrderCells / 1024;
    unsigned run = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) run += cnt[tid * per + k];
    part[tid] = run;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned x = tid >= o ? part[tid - o] : 0u;
        __syncthreads();
        part[tid] += x;
        __syncthreads();
    }
    unsigned at = part[tid] - run;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const unsigned c = cnt[tid * per + k];
        cnt[tid * per + k] = at;
        at += c;
    }
    __syncthreads();
    for (int b = tid; b < n_items; b += 1024) order[atomicAdd(&cnt[cell_of(b)], 1u)] = b;
}

// The 43 sums of the six-direction fit for each item's pose (detector.wgsl:433-494 sums +
// sums_reduce): count, sum M, sum M^2, sum M e, then for d = (q^-1 grad M / |u|, c x grad M)
// (a shift of the rays' origin in pixels, a crystal rotation) sum d_i, sum M d_i, sum e d_i and
// the upper triangle of sum d_i d_j. One wave per item over the masked-pixel list; every lane
// ends with the same totals (xor butterflies) and the lanes write them out in parallel.
// rotation_only zeroes the shift's sums, as sums_reduce does for orientation refinement.
constexpr int kSums = 43;

// skip (or NULL): items whose entry is not 0 are left out (their sums untouched).
template <bool LIST>
__global__ void __launch_bounds__(64) k_sums6(RefArg a, const float4* __restrict__ mg, const float* __restrict__ pats, const float* __restrict__ poses,
                                              int rotation_only, const int* __restrict__ skip, float* __restrict__ out) {
    extern __shared__ unsigned lds[];
    unsigned short* list = (unsigned short*)lds;
    const int b = blockIdx.x;
    if (skip && skip[b] != 0) return;
    const int lane = threadIdx.x;
    int n_px = a.h * a.w;
    if constexpr (LIST) n_px = build_list(a, list, lane);
    const unsigned short* lst = LIST ? list : nullptr;
    const float inv_w = 1.f / (float)a.w;
    const float* ps = poses + (size_t)b * 8;
    const float3 t = make_float3(ps[0], ps[1], ps[2]);
    const float* pat = pats + (size_t)(int)ps[3] * a.h * a.w;
    const float4 q = make_float4(ps[4], ps[5], ps[6], ps[7]);
    const float4 qi = make_float4(q.x, -q.y, -q.z, -q.w);
    float s[kSums];
#pragma unroll
    for (int i = 0; i < kSums; ++i) s[i] = 0.f;
    for (int k = lane; k < n_px; k += 64) {
        bool ok;
        const int pix = pass_pixel<LIST>(a, lst, n_px, k, ok);
        if (!ok) continue;
        const int row = LIST ? (int)(((float)pix + 0.5f) * inv_w) : pix / a.w;
        const float fr = (float)row, fc = (float)(pix - row * a.w);
        const float3 u = make_float3(a.c0[0] * fr + a.c1[0] * fc + t.x, a.c0[1] * fr + a.c1[1] * fc + t.y, a.c0[2] * fr + a.c1[2] * fc + t.z);
        const float dist = sqrtf(dot3(u, u));
        const float id = 1.f / dist;
        const float3 c = qrot(q, make_float3(u.x * id, u.y * id, u.z * id));
        const float4 m = sheet_bilinear<float4>(mg, a.side, c);
        const float3 gm = make_float3(m.x, m.y, m.z);
        const float3 ac = cross3(c, gm);
        const float3 w = qrot(qi, gm);
        float d[6] = {w.x * id, w.y * id, w.z * id, ac.x, ac.y, ac.z};
        const float e = pat[pix];
        const float val = m.w;
        s[0] += 1.f;
        s[1] += val;
        s[2] += val * val;
        s[3] += val * e;
        int o = 22;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            s[4 + i] += d[i];
            s[10 + i] += val * d[i];
            s[16 + i] += e * d[i];
#pragma unroll
            for (int j = i; j < 6; ++j) s[o++] += d[i] * d[j];
        }
    }
#pragma unroll
    for (int i = 0; i < kSums; ++i) s[i] = wsum(s[i]);
    // lane i writes sum i (the shift's sums zero for orientation refinement)
#pragma unroll
    for (int i = 0; i < kSums; ++i) {
        if (lane == i) {
            const bool shift = (i >= 4 && i < 7) || (i >= 10 && i < 13) || (i >= 16 && i < 19) || (i >= 22 && i < 37);
            out[(size_t)b * kSums + i] = (rotation_only && shift) ? 0.f : s[i];
        }
    }
}

}  // namespace zb

using namespace zb;

static int ref_frame(const zb_frame* frame, int side, RefArg& a, size_t& lds, const char* who) {
    for (int i = 0; i < 3; ++i) {
        a.c0[i] = frame->c0[i];
        a.c1[i] = frame->c1[i];
        a.c2[i] = frame->c2[i];
        a.mask[i] = frame->mask[i];
    }
    a.h = frame->pat_h;
    a.w = frame->pat_w;
    a.side = side;
    if ((long)a.h * a.w > (1L << 26)) {
        set_error("%s: %d x %d pixels: at most 2^26", who, a.h, a.w);
        return ZB_EINVAL;
    }
    long n_px = 0;
    for (int r = 0; r < a.h; ++r)
        for (int c = 0; c < a.w; ++c) {
            const float dr = (float)r - a.mask[0], dc = (float)c - a.mask[1];
            n_px += (a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2]) ? 1 : 0;
        }
    // the masked pixels as a 16-bit list in LDS while their indices fit 16 bits and the list
    // (with the NMI's histogram, at most 32 x 32 words) stays within 60 KB; else the whole frame
    // is walked with the mask tested per pixel
    const size_t list_bytes = ((size_t)n_px * 2 + 15) / 16 * 16;
    a.use_list = ((long)a.h * a.w <= 65536 && list_bytes <= 60 * 1024) ? 1 : 0;
    lds = a.use_list ? list_bytes : 0;
    return (int)n_px;
}

static int32_t sums_run(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* poses, int32_t n_items,
                        int32_t flags, const int32_t* skip, float* sums, void* stream);

extern "C" int32_t zb_sums(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* poses, int32_t n_items,
                           int32_t flags, float* sums, void* stream) {
    return sums_run(mgrad, side, frame, pats, poses, n_items, flags, nullptr, sums, stream);
}

extern "C" int32_t zb_sums_active(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* poses,
                                  int32_t n_items, int32_t flags, const int32_t* state, float* sums, void* stream) {
    if (n_items > 0 && !state) {
        set_error("zb_sums_active: null state");
        return ZB_EINVAL;
    }
    return sums_run(mgrad, side, frame, pats, poses, n_items, flags, state, sums, stream);
}

static int32_t sums_run(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* poses, int32_t n_items,
                        int32_t flags, const int32_t* skip, float* sums, void* stream) {
    if (!frame || side < 2 || n_items < 0 || frame->pat_h < 2 || frame->pat_w < 2 || (flags & ~ZB_SUMS_ROTATION_ONLY)) {
        set_error("zb_sums: bad arguments");
        return ZB_EINVAL;
    }
    if (n_items == 0) return ZB_OK;
    if (!mgrad || !pats || !poses || !sums) {
        set_error("zb_sums: null pointer");
        return ZB_EINVAL;
    }
    RefArg a = {};
    size_t lds = 0;
    const int n_px = ref_frame(frame, side, a, lds, "zb_sums");
    if (n_px < 0) return n_px;
    if (a.use_list)
        ZB_LAUNCH("k_sums6", k_sums6<true>, dim3(n_items), dim3(64), lds, (hipStream_t)stream, a, (const float4*)mgrad, pats, poses,
                  (flags & ZB_SUMS_ROTATION_ONLY) ? 1 : 0, (const int*)skip, sums);
    else
        ZB_LAUNCH("k_sums6", k_sums6<false>, dim3(n_items), dim3(64), lds, (hipStream_t)stream, a, (const float4*)mgrad, pats, poses,
                  (flags & ZB_SUMS_ROTATION_ONLY) ? 1 : 0, (const int*)skip, sums);
    return check_launch("zb_sums");
}

static void ref_opts(RefArg& a, const zb_refine_opts* o, int out0) {
    a.iters = o->iters;
    a.nmi_bins = o->nmi_bins;
    a.out0 = out0;
    a.max_step = o->max_step;
    a.scale0 = o->scale0;
    a.scale_min = o->scale_min;
    a.scale_max = o->scale_max;
    a.stamp = (float)o->stamp;
}

extern "C" int32_t zb_refine(const float* master, const uint16_t* mgrad_h, int32_t side, const zb_frame* frame, const float* pats, const float* items,
                             const float* seeds, int32_t seed_stride, int32_t n_items, const zb_refine_opts* o, float* records, int32_t out0,
                             void* stream) {
    if (!frame || !o || side < 2 || n_items < 0 || frame->pat_h < 2 || frame->pat_w < 2 || seed_stride < 4 || out0 < 0 || o->iters < 0 ||
        o->nmi_bins < 2 || o->nmi_bins > 32) {
        set_error("zb_refine: bad arguments");
        return ZB_EINVAL;
    }
    if (n_items == 0) return ZB_OK;
    if (!master || !mgrad_h || !pats || !items || !seeds || !records) {
        set_error("zb_refine: null pointer");
        return ZB_EINVAL;
    }
    RefArg a = {};
    size_t list_bytes = 0;
    // the masked pixels (the list each block builds) and the LDS they take
    const int n_px = ref_frame(frame, side, a, list_bytes, "zb_refine");
    if (n_px < 0) return n_px;
    ref_opts(a, o, out0);
    a.seed_stride = seed_stride;
    // the joint histogram, then the list (the histogram's size keeps the list 4-byte aligned)
    const size_t lds = (size_t)o->nmi_bins * o->nmi_bins * 4 + list_bytes;
    if (a.use_list && n_px <= 64 * 48)
        ZB_LAUNCH("k_refine", (k_refine<48, true>), dim3(n_items), dim3(64), lds, (hipStream_t)stream, a, master, (const uint2*)mgrad_h, pats,
                  (const float4*)items, seeds, n_items, (const int*)o->order, records);
    else if (a.use_list)
        ZB_LAUNCH("k_refine", (k_refine<0, true>), dim3(n_items), dim3(64), lds, (hipStream_t)stream, a, master, (const uint2*)mgrad_h, pats,
                  (const float4*)items, seeds, n_items, (const int*)o->order, records);
    else
        ZB_LAUNCH("k_refine", (k_refine<0, false>), dim3(n_items), dim3(64), lds, (hipStream_t)stream, a, master, (const uint2*)mgrad_h, pats,
                  (const float4*)items, seeds, n_items, (const int*)o->order, records);
    return check_launch("zb_refine");
}

// Blocks of k_refine_queue (form `kind`, `lds` bytes of LDS) resident at once on the current
// device: CUs x blocks per CU, queried once per device, form and LDS size (host queries only, so
// that a call under stream capture records nothing else).
static int queue_blocks(int kind, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, size_t>, int> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> g(mu);
    const auto key = std::make_tuple(dev, kind, lds);
    auto it = known.find(key);
    if (it != known.end()) return it->second;
    int cus = 0, per_cu = 0;
    const void* kern = kind == 0 ? (const void*)k_refine_queue<48, true> : kind == 1 ? (const void*)k_refine_queue<0, true>
                                                                                       : (const void*)k_refine_queue<0, false>;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds) != hipSuccess || per_cu < 1 || cus < 1)
        return 0;
    return known[key] = cus * per_cu;
}

extern "C" int32_t zb_refine_phases(const zb_refine_phase* phases, int32_t n_phases, const zb_frame* frame, const float* pats, const float* items,
                                    int32_t n_items, const zb_refine_opts* o, int32_t out0, void* work, int64_t work_bytes, void* stream) {
    if (!frame || !o || n_phases < 1 || n_phases > kMaxRefinePhases || n_items < 0 || frame->pat_h < 2 || frame->pat_w < 2 || out0 < 0 ||
        o->iters < 0 || o->nmi_bins < 2 || o->nmi_bins > 32 || (long)n_items * n_phases >= (1L << 31)) {
        set_error("zb_refine_phases: bad arguments");
        return ZB_EINVAL;
    }
    if (!phases) {
        set_error("zb_refine_phases: null pointer");
        return ZB_EINVAL;
    }
    RefPhases ph = {};
    for (int p = 0; p < n_phases; ++p) {
        const zb_refine_phase& q = phases[p];
        if (q.side < 2 || q.seed_stride < 4) {
            set_error("zb_refine_phases: phase %d: bad arguments", p);
            return ZB_EINVAL;
        }
        if (n_items > 0 && (!q.master || !q.mgrad_h || !q.seeds || !q.records)) {
            set_error("zb_refine_phases: phase %d: null pointer", p);
            return ZB_EINVAL;
        }
        ph.p[p] = RefPhase{q.order, q.side, q.seed_stride};
    }
    if (n_items == 0) return ZB_OK;
    if (!pats || !items || !work) {
        set_error("zb_refine_phases: null pointer");
        return ZB_EINVAL;
    }
    if (work_bytes < ZB_REFINE_QUEUE_BYTES) {
        set_error("zb_refine_phases: work_bytes %lld below %d", (long long)work_bytes, ZB_REFINE_QUEUE_BYTES);
        return ZB_ESPACE;
    }
    if ((uintptr_t)work % 8) {
        set_error("zb_refine_phases: work not 8-byte aligned");
        return ZB_EINVAL;
    }
    RefArg a = {};
    size_t list_bytes = 0;
    const int n_px = ref_frame(frame, phases[0].side, a, list_bytes, "zb_refine_phases");
    if (n_px < 0) return n_px;
    ref_opts(a, o, out0);
    const size_t lds = (size_t)o->nmi_bins * o->nmi_bins * 4 + list_bytes;
    const long total = (long)n_items * n_phases;
    unsigned* queues = (unsigned*)work;
    const hipStream_t st = (hipStream_t)stream;
    // 0: the list and each lane's master values kept between the NMI's passes; 1: the list, the
    // values looked up again (fewer registers: with the queue knob at 2); 2: the whole frame
    const int kind = (a.use_list && n_px <= 64 * 48 && knob(KNOB_REFINE_QUEUE) != 2) ? 0 : a.use_list ? 1 : 2;
    // phases past n_phases repeat the last one's entry (never read)
    for (int p = n_phases; p < kMaxRefinePhases; ++p) ph.p[p] = ph.p[n_phases - 1];
    // the phases' arrays as kernel arguments (phases past n_phases repeat the last: never read)
    const float* ms[kMaxRefinePhases];
    const uint2* mgh[kMaxRefinePhases];
    const float* sd[kMaxRefinePhases];
    float* rec[kMaxRefinePhases];
    for (int p = 0; p < kMaxRefinePhases; ++p) {
        const zb_refine_phase& q = phases[std::min(p, n_phases - 1)];
        ms[p] = q.master;
        mgh[p] = (const uint2*)q.mgrad_h;
        sd[p] = q.seeds;
        rec[p] = q.records;
    }
#define ZB_PHASES ms[0], mgh[0], sd[0], rec[0], ms[1], mgh[1], sd[1], rec[1], ms[2], mgh[2], sd[2], rec[2], ms[3], mgh[3], sd[3], rec[3]
    if (knob(KNOB_REFINE_QUEUE)) {
        // the queue counters cleared and the phase table written into the work buffer
        ZB_LAUNCH("k_queue_setup", k_queue_setup, dim3(1), dim3(64), 0, st, ph, n_phases, queues);
        if (int rc = check_launch("zb_refine_phases/setup")) return rc;
        // as many blocks as stay resident at once (more would only wait for the queues to empty)
        const int resident = queue_blocks(kind, lds);
        if (resident < 1) {
            set_error("zb_refine_phases: device or occupancy query failed");
            return ZB_ELAUNCH;
        }
        const int grid = (int)std::min<long>(total, resident);
        if (kind == 0)
            ZB_LAUNCH("k_refine", (k_refine_queue<48, true>), dim3(grid), dim3(64), lds, st, a, ZB_PHASES, n_phases, pats, (const float4*)items,
                      n_items, queues);
        else if (kind == 1)
            ZB_LAUNCH("k_refine", (k_refine_queue<0, true>), dim3(grid), dim3(64), lds, st, a, ZB_PHASES, n_phases, pats, (const float4*)items,
                      n_items, queues);
        else
            ZB_LAUNCH("k_refine", (k_refine_queue<0, false>), dim3(grid), dim3(64), lds, st, a, ZB_PHASES, n_phases, pats, (const float4*)items,
                      n_items, queues);
    } else {
        if (kind == 0)
            ZB_LAUNCH("k_refine", (k_refine_grid<48, true>), dim3(total), dim3(64), lds, st, a, ZB_PHASES, pats, (const float4*)items, n_items, ph);
        else if (kind == 1)
            ZB_LAUNCH("k_refine", (k_refine_grid<0, true>), dim3(total), dim3(64), lds, st, a, ZB_PHASES, pats, (const float4*)items, n_items, ph);
        else
            ZB_LAUNCH("k_refine", (k_refine_grid<0, false>), dim3(total), dim3(64), lds, st, a, ZB_PHASES, pats, (const float4*)items, n_items,
                      ph);
    }
#undef ZB_PHASES
    return check_launch("zb_refine_phases");
}

extern "C" int32_t zb_refine_order(const zb_frame* frame, const float* items, const float* seeds, int32_t seed_stride, int32_t n_items,
                                   int32_t* order, void* stream) {
    if (!frame || n_items < 0 || seed_stride < 4 || frame->pat_h < 2 || frame->pat_w < 2) {
        set_error("zb_refine_order: bad arguments");
        return ZB_EINVAL;
    }
    if (n_items == 0) return ZB_OK;
    if (!items || !seeds || !order) {
        set_error("zb_refine_order: null pointer");
        return ZB_EINVAL;
    }
    RefArg a = {};
    for (int i = 0; i < 3; ++i) {
        a.c0[i] = frame->c0[i];
        a.c1[i] = frame->c1[i];
        a.c2[i] = frame->c2[i];
        a.mask[i] = frame->mask[i];
    }
    a.h = frame->pat_h;
    a.w = frame->pat_w;
    ZB_LAUNCH("k_order", k_order, dim3(1), dim3(1024), 0, (hipStream_t)stream, a, (const float4*)items, seeds, seed_stride, n_items, order);
    return check_launch("zb_refine_order");
}
