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
#include "common.h"

namespace zb {

struct RefArg {
    float c0[3], c1[3], c2[3];
    float mask[3];
    int h, w, side;
    int iters, nmi_bins, seed_stride, out0;
    float max_step, scale0, scale_min, scale_max;
    float stamp;
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

__device__ Sums rot_sums(const RefArg& a, const float4* __restrict__ mg, const float* __restrict__ pat, float4 q, float3 t, int lane) {
    Sums s;
    s.n = s.m = s.mm = s.me = 0.f;
    s.sd = s.sv = s.se = s.hd = s.ho = make_float3(0.f, 0.f, 0.f);
    const int hw = a.h * a.w;
    int row = lane / a.w, col = lane - (lane / a.w) * a.w;
    const int step_r = 64 / a.w, step_c = 64 - (64 / a.w) * a.w;
    for (int pix = lane; pix < hw; pix += 64) {
        const float fr = (float)row, fc = (float)col;
        const float dr = fr - a.mask[0], dc = fc - a.mask[1];
        if (a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2]) {
            float3 u = make_float3(a.c0[0] * fr + a.c1[0] * fc + t.x, a.c0[1] * fr + a.c1[1] * fc + t.y, a.c0[2] * fr + a.c1[2] * fc + t.z);
            const float3 c = qrot(q, normalize3(u));
            const float4 m = sheet_bilinear<float4>(mg, a.side, c);
            const float3 d = cross3(c, make_float3(m.x, m.y, m.z));
            const float e = pat[pix];
            s.n += 1.f;
            s.m += m.w;
            s.mm += m.w * m.w;
            s.me += m.w * e;
            s.sd = f3add(s.sd, d);
            s.sv = f3add(s.sv, f3s(d, m.w));
            s.se = f3add(s.se, f3s(d, e));
            s.hd = f3add(s.hd, make_float3(d.x * d.x, d.y * d.y, d.z * d.z));
            s.ho = f3add(s.ho, make_float3(d.x * d.y, d.x * d.z, d.y * d.z));
        }
        row += step_r;
        col += step_c;
        if (col >= a.w) {
            col -= a.w;
            row += 1;
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

__global__ void __launch_bounds__(64) k_refine(RefArg a, const float4* __restrict__ mg, const float* __restrict__ pats,
                                               const float4* __restrict__ items, const float* __restrict__ seeds, int n_items,
                                               float* __restrict__ records) {
    __shared__ unsigned hist[32 * 32];
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    const float4 it = items[b];
    const float3 t = make_float3(it.x, it.y, it.z);
    const float* pat = pats + (size_t)(int)it.w * a.h * a.w;
    const float* sd = seeds + (size_t)b * a.seed_stride;
    float4 cur = make_float4(sd[0], sd[1], sd[2], sd[3]);

    // refinement state (detector.wgsl refine_init)
    float4 qbest = cur;
    float best = -2.f;
    float mu = a.scale0;
    float3 gbest = make_float3(0.f, 0.f, 0.f);
    float hbest[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float last_step = 0.f, promised = 0.f, step_size = 0.f, accepted = 0.f;

    for (int itn = 0; itn <= a.iters; ++itn) {
        const Sums s = rot_sums(a, mg, pat, cur, t, lane);
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
    const int hw = a.h * a.w;
    float emin = 3.4e38f, emax = -3.4e38f, smin = 3.4e38f, smax = -3.4e38f;
    for (int i = lane; i < nb * nb; i += 64) hist[i] = 0u;
    for (int pix = lane; pix < hw; pix += 64) {
        const float fr = (float)(pix / a.w), fc = (float)(pix % a.w);
        const float dr = fr - a.mask[0], dc = fc - a.mask[1];
        if (!(a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2])) continue;
        const float e = pat[pix];
        float3 u = make_float3(a.c0[0] * fr + a.c1[0] * fc + t.x, a.c0[1] * fr + a.c1[1] * fc + t.y, a.c0[2] * fr + a.c1[2] * fc + t.z);
        const float sim = sheet_bilinear<float4>(mg, a.side, qrot(qbest, normalize3(u))).w;
        emin = fminf(emin, e);
        emax = fmaxf(emax, e);
        smin = fminf(smin, sim);
        smax = fmaxf(smax, sim);
    }
    for (int o = 1; o < 64; o <<= 1) {
        emin = fminf(emin, __shfl_xor(emin, o, 64));
        emax = fmaxf(emax, __shfl_xor(emax, o, 64));
        smin = fminf(smin, __shfl_xor(smin, o, 64));
        smax = fmaxf(smax, __shfl_xor(smax, o, 64));
    }
    __syncthreads();
    const float ke = (float)nb / fmaxf(emax - emin, 1e-20f);
    const float ks = (float)nb / fmaxf(smax - smin, 1e-20f);
    for (int pix = lane; pix < hw; pix += 64) {
        const float fr = (float)(pix / a.w), fc = (float)(pix % a.w);
        const float dr = fr - a.mask[0], dc = fc - a.mask[1];
        if (!(a.mask[2] <= 0.f || dr * dr + dc * dc <= a.mask[2])) continue;
        const float e = pat[pix];
        float3 u = make_float3(a.c0[0] * fr + a.c1[0] * fc + t.x, a.c0[1] * fr + a.c1[1] * fc + t.y, a.c0[2] * fr + a.c1[2] * fc + t.z);
        const float sim = sheet_bilinear<float4>(mg, a.side, qrot(qbest, normalize3(u))).w;
        const int be = min((int)((e - emin) * ke), nb - 1);
        const int bs = min((int)((sim - smin) * ks), nb - 1);
        atomicAdd(&hist[be * nb + bs], 1u);
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
        float* r = records + (size_t)(a.out0 + b) * 8;
        r[0] = qbest.x;
        r[1] = qbest.y;
        r[2] = qbest.z;
        r[3] = qbest.w;
        r[4] = best;
        r[5] = nmi;
        r[6] = last_step;
        r[7] = a.stamp;
    }
}

}  // namespace zb

using namespace zb;

extern "C" int32_t zb_refine(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* items, const float* seeds,
                             int32_t seed_stride, int32_t n_items, const zb_refine_opts* o, float* records, int32_t out0, void* stream) {
    if (!frame || !o || side < 2 || n_items < 0 || frame->pat_h < 2 || frame->pat_w < 2 || seed_stride < 4 || out0 < 0 || o->iters < 0 ||
        o->nmi_bins < 2 || o->nmi_bins > 32) {
        set_error("zb_refine: bad arguments");
        return ZB_EINVAL;
    }
    if (n_items == 0) return ZB_OK;
    if (!mgrad || !pats || !items || !seeds || !records) {
        set_error("zb_refine: null pointer");
        return ZB_EINVAL;
    }
    RefArg a;
    for (int i = 0; i < 3; ++i) {
        a.c0[i] = frame->c0[i];
        a.c1[i] = frame->c1[i];
        a.c2[i] = frame->c2[i];
        a.mask[i] = frame->mask[i];
    }
    a.h = frame->pat_h;
    a.w = frame->pat_w;
    a.side = side;
    a.iters = o->iters;
    a.nmi_bins = o->nmi_bins;
    a.seed_stride = seed_stride;
    a.out0 = out0;
    a.max_step = o->max_step;
    a.scale0 = o->scale0;
    a.scale_min = o->scale_min;
    a.scale_max = o->scale_max;
    a.stamp = (float)o->stamp;
    ZB_LAUNCH("k_refine", k_refine, dim3(n_items), dim3(64), 0, (hipStream_t)stream, a, (const float4*)mgrad, pats, (const float4*)items, seeds,
                       n_items, records);
    return check_launch("zb_refine");
}
