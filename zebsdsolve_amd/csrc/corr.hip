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
//   k_vgemm   four waves per (m', column group, 128 items), each its own 32 items, looping over
//             every j: the items' coefficients for m' are the MFMA A operand (rows = items, k =
//             (re, im) of one degree), held in registers across all j; the table rows the B
//             operand (columns = V columns, (re, im) interleaved, so one real MFMA per degree
//             gives V as complex pairs), staged once per block through LDS one j ahead.
//             v_mfma_f32_32x32x2_f32 (float32 in, float32 accumulate). V goes to HBM as complex
//             pairs [j][m'][item][c]: per accumulator register the wave writes two items' whole
//             runs of 32 columns (128 bytes, a cache line each), from a scalar base plus a
//             per-lane offset that does not change with j.
//   k_apeak   one lane per (j, item, c) column, four waves a block on one j: the column's L
//             orders into the half-length complex input of the real transform along a, a
//             radix-2 FFT of h = n/2 points in registers (twiddles are compile-time constants,
//             the trivial ones skipped), times the reciprocal denominator (the block's columns,
//             loaded straight into LDS ahead of the spectrum), the column's maximum, the slab's
//             (xor shuffles over its nc lanes); only a slab that reaches the item's best so far
//             looks up its lowest index of that value and offers (value, index) with one 64-bit
//             atomicMax. Or the volume itself (denominators).
//   k_vfinish the peak's six neighbours summed directly from V (already along c), parabolic
//             offsets and the orientation, as so3.hip's k_finish.
//
// Volumes here (rden, and the window's) are in natural order [j][c][a]; V's items are padded to
// a multiple of 64 and its rows to the table's column count, so that no store needs a bounds
// check.
#include "common.h"
#include "tw128.h"
#include "tw256.h"

namespace zb {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Raw buffer resources (gfx9 family: dword 3 = 0x00020000, 32-bit data format): a scalar base
// and a 32-bit per-lane offset, so the per-lane address arithmetic of a long unrolled run of
// loads or stores is one offset register, and the hardware bounds every access by `bytes`.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf(const void* base, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(bytes < 0x7fffffffLL ? bytes : 0x7fffffffLL), 0x00020000);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ float2 buf_f32x2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    return make_float2(__uint_as_float(u.x), __uint_as_float(u.y));
}

// Cache policy of V's stores and loads: non-temporal (nt). V is written once by the GEMM and
// read once by the transform along a, gigabytes apart, so no cache level would hold it for
// the other: streaming it past them measured 3.5 % faster for the GEMM and 8.5 % for the
// transform along a (nickel, L = 64).
constexpr int kStream = 2;
__device__ __forceinline__ void buf_st_f32(float x, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, (int)voff, (int)soff, 0);
}

// items per V row are padded to a multiple of this (a k_vgemm block's items are 4 waves x 32),
// plus a skew of 32
constexpr int kItemPad = 128;

struct VArg {
    int L, n, log2n, nc, log2nc, nb, slots, cols, l_step, pairs, tiles, n_items, ni_pad;
    // rows of b the GEMM computes (jg) and the peak searches (jp): all n/2 + 1, or with b_half
    // j <= n/4 searched and one row more computed for the peak's neighbour along b
    int jg, jp;
    int lp_gs;  // 16-bit products: the first group of 8 slots that takes the hi + lo product
    int vk;     // V's elements: 0 FP32 pairs (cols floats a row), 1 FP16 / 2 BF16 pairs (nc a row)
    // the items' best (value, index) keys, which the correlation's GEMM clears (its blocks of
    // m' = 0 and column group 0, one per item tile) for the peak search after it; null: none
    unsigned long long* best;
};

// The keys of this wave's item tile cleared by the blocks that hold m' = 0 and column group 0.
template <int MB>
__device__ __forceinline__ void clear_best(const VArg& a, int mp, int cg, int tile, int half, int r32) {
    if (a.best && mp == 0 && cg == 0 && half == 0) {
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const int item = (tile * MB + mb) * 32 + r32;
            if (item < a.n_items) a.best[item] = 0ull;
        }
    }
}

// V's complex elements: FP32 (re, im) (8 bytes) or FP16 / BF16 pairs packed in 32 bits
template <int VT>
struct VEl {
    static constexpr unsigned bytes = VT ? 4u : 8u;
    static __device__ __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        if constexpr (VT == 0) {
            const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, kStream);
            return make_float2(__uint_as_float(u.x), __uint_as_float(u.y));
        } else {
            return unpack(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, kStream));
        }
    }
    static __device__ __forceinline__ float2 at(const void* v, size_t i) {
        if constexpr (VT == 0) return reinterpret_cast<const float2*>(v)[i];
        else return unpack(reinterpret_cast<const unsigned*>(v)[i]);
    }
    static __device__ __forceinline__ float2 unpack(unsigned u) {
        if constexpr (VT == 1) {
            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
            const h2 h = __builtin_bit_cast(h2, u);
            return make_float2((float)h[0], (float)h[1]);
        } else {
            return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
        }
    }
    // (re, im) rounded to nearest, ties to even
    static __device__ __forceinline__ unsigned pack(float re, float im) {
        if constexpr (VT == 1) {
            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
            h2 h;
            h[0] = (_Float16)re;
            h[1] = (_Float16)im;
            return __builtin_bit_cast(unsigned, h);
        } else {
            typedef __bf16 b2 __attribute__((ext_vector_type(2)));
            b2 h;
            h[0] = (__bf16)re;
            h[1] = (__bf16)im;
            return __builtin_bit_cast(unsigned, h);
        }
    }
};


// Waves per block: each computes its own tile of 64 items against the same table rows, which
// the block stages once per j through LDS.
constexpr int kWaves = 4;

// One block's GEMM for orders m' whose first live degree slot is at least S0: every slot from S0
// to S - 1 is multiplied (the items' coefficients are zero below their first slot and past the
// last one, and so is the table below l = m'), so no branch stands between the table and the
// MFMA. The table rows of j + 1 go from HBM/L2 to registers while j's products run, and from
// there to the other LDS buffer after them.
template <int MB, int NBLK, int S, int S0>
__device__ __forceinline__ void vgemm_run(const VArg& a, int mp, int tile, int cg, const float* __restrict__ tab, const float (&areg)[MB][S],
                                          float* __restrict__ v, floatx4* lds0, floatx4* lds1) {
    constexpr int NQ = (S - S0) / 4;                     // groups of four slots from S0
    constexpr int F4 = NQ * NBLK * 64;                   // 16-byte pieces of one j's rows
    constexpr int PER = (F4 + 64 * kWaves - 1) / (64 * kWaves);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int r32 = lane & 31;
    const int half = lane >> 5;
    const int ncg = a.cols / (NBLK * 32);
    // the tiled table: per (j, m', cg) [S/4][NBLK][64 lanes][4 slots]; this block's rows start
    // at slot group S0 / 4 and are F4 consecutive 16-byte pieces for each j
    const size_t blk = (size_t)S * NBLK * 64;
    const size_t first = ((size_t)mp * ncg + cg) * blk + (size_t)(S0 / 4) * NBLK * 64 * 4;
    // one j's rows of this block are F4 pieces at tab + first + j jstride floats; the resource
    // is made per j over exactly them, so the hardware's bounds (voffset against the record
    // count) cover every access and no offset passes 32 bits, whatever the table's size (GaN's
    // is 2.4 GB at L = 96)
    const size_t jstride = (size_t)a.L * ncg * blk;
    floatx4 pre[PER];
    // every thread loads (a clamped piece when the rows are fewer than the threads) so that
    // the waits count the same loads on every path
    auto stage_load = [&](int j) {
        const auto trs = buf(tab + first + (size_t)j * jstride, (long long)F4 * 16);
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = min(tid + 64 * kWaves * p, F4 - 1);
            pre[p] = __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(trs, idx * 16, 0, 0));
        }
    };
    // pieces past the rows go to a spare slot after them (no predicate, so the loads above
    // stay where they are issued)
    auto stage_store = [&](floatx4* dst) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = tid + 64 * kWaves * p;
            dst[idx < F4 ? idx : F4] = pre[p];
        }
    };
    // V row of one (j, m'): ni_pad items x cols floats (2 nc, or 32 when nc < 16: the table's
    // zero columns are stored too, so no store needs a predicate)
    const int w = a.cols;
    const int c0 = cg * NBLK * 32;
    const size_t vrow = (size_t)a.ni_pad * w;
    unsigned loff[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) loff[mb] = 4u * (unsigned)(((tile * MB + mb) * 32 + half * 4) * w + c0 + r32);
    // one sum per output tile: the other waves on the SIMD fill the wait of a chain's next
    // product for the previous one
    typedef floatx16 Acc[MB][NBLK][1];
    // one j: prefetch the next j's rows, the products (the coefficients are the A operand, rows =
    // items; the table rows the B operand, columns = V columns), the stores, the staging of the
    // next rows. Two accumulator sets alternate between consecutive j, so the products of j + 1
    // never wait for the stores of j to read their registers.
    auto step = [&](int j, Acc& acc, Acc& other, const floatx4* cur, floatx4* nxt) {
        // the last j stages its own rows again: no branch
        stage_load(min(j + 1, a.jg - 1));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[mb][nb][0][r] = 0.f;
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            floatx4 fa[NBLK];
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb) fa[nb] = cur[(k * NBLK + nb) * 64 + lane];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NBLK; ++nb) {
                        floatx16& c = acc[mb][nb][0];
                        c = __builtin_amdgcn_mfma_f32_32x32x2f32(areg[mb][S0 + 4 * k + i], fa[nb][i], c, 0, 0, 0);
                    }
        }
        // the other set (stored one step ago) stays live across these products, so the two
        // sets get their own registers
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb)
                asm volatile("" ::"a"(other[mb][nb][0]));
        __builtin_amdgcn_sched_barrier(0);
        // C[item][V column]: the lane holds V column r32 of items (r / 4) 8 + half 4 + r % 4, so
        // each store writes two items' whole 128-byte runs of 32 columns (whole cache lines)
        const auto vrs = buf(v + ((size_t)j * a.L + mp) * vrow, (long long)vrow * 4);
        // the stores read the accumulators themselves: their registers are written again only
        // by the products two steps on
#pragma unroll
        for (int nb = 0; nb < NBLK; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r)
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    const floatx16& x = acc[mb][nb][0];
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x[r]), vrs, (int)loff[mb], (int)((((r >> 2) * 8 + (r & 3)) * w + nb * 32) * 4),
                                                          kStream);
                }
        stage_store(nxt);
        __syncthreads();
    };
    stage_load(0);
    stage_store(lds0);
    __syncthreads();
    Acc acc0, acc1;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NBLK; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc1[mb][nb][0][r] = 0.f;
    // an odd count repeats the last j (the same values to the same place)
    for (int j = 0; j < a.jg; j += 2) {
        step(j, acc0, acc1, lds0, lds1);
        step(min(j + 1, a.jg - 1), acc1, acc0, lds1, lds0);
    }
}

// The bucket of four slots the first live slot falls in picks the unrolled variant.
template <int MB, int NBLK, int S, int S0 = 0>
__device__ __forceinline__ void vgemm_bucket(int bucket, const VArg& a, int mp, int tile, int cg, const float* __restrict__ tab,
                                             const float (&areg)[MB][S], float* __restrict__ v, floatx4* lds0, floatx4* lds1) {
    if constexpr (S0 < S) {
        if (bucket == S0 / 4) vgemm_run<MB, NBLK, S, S0>(a, mp, tile, cg, tab, areg, v, lds0, lds1);
        else vgemm_bucket<MB, NBLK, S, S0 + 4>(bucket, a, mp, tile, cg, tab, areg, v, lds0, lds1);
    }
}

// V: [j][m'][ni_pad][cols] floats, complex pairs (re, im) for c < nc.
template <int MB, int NBLK, int S>
__global__ void __launch_bounds__(64 * kWaves) k_vgemm(VArg a, const float* __restrict__ tab, const float2* __restrict__ coef, float* __restrict__ v) {
    extern __shared__ floatx4 vlds[];
    // blocks are dealt to the 8 XCDs round-robin (block id mod 8). The work of an m' falls with
    // m' (fewer live slots), so when L is a multiple of 8 each XCD takes the m' of a snake
    // order (x, 15 - x, 16 + x, 31 - x, ...) and the XCDs get equal work; within an XCD the
    // blocks run m' by m', tiles innermost, so its L2 holds one m's table rows at a time.
    const int ncg = a.cols / (NBLK * 32);
    const int per_m = a.tiles * ncg;
    int mp, rem;
    if ((a.L & 7) == 0) {
        const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
        const int mi = k / per_m;
        rem = k - mi * per_m;
        mp = 8 * mi + ((mi & 1) ? 7 - x : x);
    } else {
        const int logical = xcd_block(blockIdx.x, gridDim.x);
        mp = logical / per_m;
        rem = logical - mp * per_m;
    }
    const int cg = rem / a.tiles;
    const int tile = (rem - cg * a.tiles) * kWaves + (threadIdx.x >> 6);  // this wave's tile of 32 MB items
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int r32 = lane & 31;
    clear_best<MB>(a, mp, cg, tile, half, r32);
    const int s0 = (mp + a.l_step - 1) / a.l_step;
    const int pbase = mp * a.L - (mp * (mp - 1)) / 2 - mp;  // pair(mp, l) = pbase + l
    // B: rows k = (re, im) of one degree slot, columns = items; zero outside [s0, slots) and
    // past the last item
    float areg[MB][S];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int item = (tile * MB + mb) * 32 + r32;
        const bool ok = item < a.n_items;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            float x = 0.f;
            if (ok && s >= s0 && s < a.slots) {
                const float2 g = coef[(size_t)item * a.pairs + pbase + s * a.l_step];
                x = half ? g.y : g.x;
            }
            areg[mb][s] = x;
        }
    }
    constexpr int F4MAX = (S / 4) * NBLK * 64 + 1;  // one j's rows and a spare slot
    vgemm_bucket<MB, NBLK, S>(min(s0, S - 1) >> 2, a, mp, tile, cg, tab, areg, v, vlds, vlds + F4MAX);
}

// The GEMM with 16-bit products (FP16 or BF16 in, FP32 accumulate):
// v_mfma_f32_32x32x16_{f16,bf16}, whose 16 k per lane pair of halves are here the (re, im)
// rows of 8 degree slots, re in lanes 0..31 and im in lanes 32..63 as in the FP32 form. A value
// x is carried as hi = round16(x) and lo = round16(x - hi); groups of 8 slots from lp_gs on take
// hi hi + hi lo + lo hi (three products, the error of a ~22-bit (FP16) or ~16-bit (BF16)
// operand), the ones below it hi hi alone. The table (zb_vtable.qlp): per (j, m', column group)
// the groups in order, each its hi pieces [nblk][64 lanes][8 slots] of 16 bytes, then (from
// lp_gs on) its lo pieces. V is FP32, as in the FP32 form.
template <bool BF>
struct Lp16;
template <>
struct Lp16<false> {
    typedef _Float16 T;
    typedef _Float16 V8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ floatx16 mfma(V8 a, V8 b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
};
template <>
struct Lp16<true> {
    typedef __bf16 T;
    typedef __bf16 V8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ floatx16 mfma(V8 a, V8 b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};

template <int MB, int NBLK, int G8, int GS, int G0, bool BF>
__device__ __forceinline__ void vgemm16_run(const VArg& a, int mp, int tile, int cg, const floatx4* __restrict__ tab,
                                            const typename Lp16<BF>::V8 (&ah)[MB][G8], const typename Lp16<BF>::V8 (&al)[MB][G8 - GS + 1],
                                            unsigned* __restrict__ v, floatx4* lds0, floatx4* lds1) {
    static_assert(NBLK == 2, "the 16-bit form pairs the real and imaginary column blocks");
    typedef typename Lp16<BF>::V8 V8;
    constexpr int NQ = G8 - G0;                        // groups of 8 slots from G0
    constexpr int PB = NBLK * 64;                      // 16-byte pieces of a group's hi (or lo) rows
    // pieces before group g in one (j, m', column group) block
    constexpr auto off = [](int g) { return (g + (g > GS ? g - GS : 0)) * PB; };
    constexpr int F4 = off(G8) - off(G0);
    constexpr int PER = (F4 + 64 * kWaves - 1) / (64 * kWaves);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int r32 = lane & 31;
    const int half = lane >> 5;
    const int ncg = a.cols / (NBLK * 32);
    const size_t blk = (size_t)off(G8);
    const size_t first = ((size_t)mp * ncg + cg) * blk + (size_t)off(G0);
    // one j's rows of this block: F4 16-byte pieces at tab + first + j jstride, a resource per j
    // over exactly them (bounds on every access, offsets within 32 bits at any table size)
    const size_t jstride = (size_t)a.L * ncg * blk;
    floatx4 pre[PER];
    auto stage_load = [&](int j) {
        const auto trs = buf(tab + first + (size_t)j * jstride, (long long)F4 * 16);
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = min(tid + 64 * kWaves * p, F4 - 1);
            pre[p] = __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(trs, idx * 16, 0, 0));
        }
    };
    auto stage_store = [&](floatx4* dst) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int idx = tid + 64 * kWaves * p;
            dst[idx < F4 ? idx : F4] = pre[p];  // past the rows: the spare slot after them
        }
    };
    // V row of one (j, m'): ni_pad items x nc 32-bit (re, im) pairs; the lane's complex column
    const int w = a.nc;
    const size_t vrow = (size_t)a.ni_pad * w;
    unsigned loff[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) loff[mb] = 4u * (unsigned)(((tile * MB + mb) * 32 + half * 4) * w + cg * 32 + r32);
    typedef floatx16 Acc[MB][NBLK][1];
    auto step = [&](int j, Acc& acc, Acc& other, const floatx4* cur, floatx4* nxt) {
        stage_load(min(j + 1, a.jg - 1));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[mb][nb][0][r] = 0.f;
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int g = G0 + k;
            const floatx4* rows = cur + (off(g) - off(G0));
            V8 bh[NBLK];
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb) bh[nb] = __builtin_bit_cast(V8, rows[nb * 64 + lane]);
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int nb = 0; nb < NBLK; ++nb) acc[mb][nb][0] = Lp16<BF>::mfma(ah[mb][g], bh[nb], acc[mb][nb][0]);
            if (g >= GS) {  // known once unrolled
                V8 bl[NBLK];
#pragma unroll
                for (int nb = 0; nb < NBLK; ++nb) bl[nb] = __builtin_bit_cast(V8, rows[PB + nb * 64 + lane]);
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NBLK; ++nb) {
                        acc[mb][nb][0] = Lp16<BF>::mfma(ah[mb][g], bl[nb], acc[mb][nb][0]);
                        acc[mb][nb][0] = Lp16<BF>::mfma(al[mb][g >= GS ? g - GS : 0], bh[nb], acc[mb][nb][0]);
                    }
            }
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < NBLK; ++nb)
                asm volatile("" ::"a"(other[mb][nb][0]));
        __builtin_amdgcn_sched_barrier(0);
        // C[item][column]: block 0 the real parts and block 1 the imaginary parts of complex
        // column cg 32 + r32, packed as one 32-bit pair: each store writes two items' whole
        // 128-byte runs of 32 pairs
        const auto vrs = buf(v + ((size_t)j * a.L + mp) * vrow, (long long)vrow * 4);
#pragma unroll
        for (int r = 0; r < 16; ++r)
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const unsigned u = VEl<BF ? 2 : 1>::pack(acc[mb][0][0][r], acc[mb][1][0][r]);
                __builtin_amdgcn_raw_buffer_store_b32(u, vrs, (int)loff[mb], (int)((((r >> 2) * 8 + (r & 3)) * w) * 4), kStream);
            }
        stage_store(nxt);
        __syncthreads();
    };
    stage_load(0);
    stage_store(lds0);
    __syncthreads();
    Acc acc0, acc1;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NBLK; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc1[mb][nb][0][r] = 0.f;
    for (int j = 0; j < a.jg; j += 2) {
        step(j, acc0, acc1, lds0, lds1);
        step(min(j + 1, a.jg - 1), acc1, acc0, lds1, lds0);
    }
}

template <int MB, int NBLK, int G8, int GS, bool BF, int G0 = 0>
__device__ __forceinline__ void vgemm16_bucket(int bucket, const VArg& a, int mp, int tile, int cg, const floatx4* __restrict__ tab,
                                               const typename Lp16<BF>::V8 (&ah)[MB][G8], const typename Lp16<BF>::V8 (&al)[MB][G8 - GS + 1],
                                               unsigned* __restrict__ v, floatx4* lds0, floatx4* lds1) {
    if constexpr (G0 < G8) {
        if (bucket == G0) vgemm16_run<MB, NBLK, G8, GS, G0, BF>(a, mp, tile, cg, tab, ah, al, v, lds0, lds1);
        else vgemm16_bucket<MB, NBLK, G8, GS, BF, G0 + 1>(bucket, a, mp, tile, cg, tab, ah, al, v, lds0, lds1);
    }
}

// k_vgemm with 16-bit products: the same blocks, item tiles and V; G8 groups of 8 slots, those
// from GS on with hi + lo products.
template <int MB, int NBLK, int G8, int GS, bool BF>
__global__ void __launch_bounds__(64 * kWaves) k_vgemm16(VArg a, const floatx4* __restrict__ tab, const float2* __restrict__ coef, unsigned* __restrict__ v) {
    typedef typename Lp16<BF>::T T;
    typedef typename Lp16<BF>::V8 V8;
    extern __shared__ floatx4 vlds[];
    const int ncg = a.cols / (NBLK * 32);
    const int per_m = a.tiles * ncg;
    int mp, rem;
    if ((a.L & 7) == 0) {
        const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
        const int mi = k / per_m;
        rem = k - mi * per_m;
        mp = 8 * mi + ((mi & 1) ? 7 - x : x);
    } else {
        const int logical = xcd_block(blockIdx.x, gridDim.x);
        mp = logical / per_m;
        rem = logical - mp * per_m;
    }
    const int cg = rem / a.tiles;
    const int tile = (rem - cg * a.tiles) * kWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int r32 = lane & 31;
    clear_best<MB>(a, mp, cg, tile, half, r32);
    const int s0 = (mp + a.l_step - 1) / a.l_step;
    const int pbase = mp * a.L - (mp * (mp - 1)) / 2 - mp;
    // A: row = item, the 8 k of this lane half = component `half` of slots 8 g .. 8 g + 7, as
    // hi and lo; zero outside [s0, slots) and past the last item
    V8 ah[MB][G8], al[MB][G8 - GS + 1];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int item = (tile * MB + mb) * 32 + r32;
        const bool ok = item < a.n_items;
#pragma unroll
        for (int g = 0; g < G8; ++g) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int s = 8 * g + i;
                float x = 0.f;
                if (ok && s >= s0 && s < a.slots) {
                    const float2 c = coef[(size_t)item * a.pairs + pbase + s * a.l_step];
                    x = half ? c.y : c.x;
                }
                const T h = (T)x;
                ah[mb][g][i] = h;
                if (g >= GS) al[mb][g >= GS ? g - GS : 0][i] = (T)(x - (float)h);
            }
        }
    }
    // two buffers of one j's rows from group 0 (hi, and lo from GS on) and a spare slot
    constexpr int pieces = (2 * G8 - GS) * NBLK * 64 + 1;
    vgemm16_bucket<MB, NBLK, G8, GS, BF>(min(s0, 8 * G8 - 1) >> 3, a, mp, tile, cg, tab, ah, al, v, vlds, vlds + pieces);
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

// v exp(-2 pi i k / 128) for a k known once the loops are unrolled: the multiplications by 1,
// -i, -1, i and the odd eighth roots of unity reduce to moves and one product per part
__device__ __forceinline__ float2 twmul(float2 v, int k) {
    k &= 127;
    constexpr float r = 0.70710678118654752f;
    if (k == 0) return v;
    if (k == 32) return make_float2(v.y, -v.x);
    if (k == 64) return make_float2(-v.x, -v.y);
    if (k == 96) return make_float2(-v.y, v.x);
    if (k == 16) return make_float2((v.x + v.y) * r, (v.y - v.x) * r);
    if (k == 48) return make_float2((v.y - v.x) * r, -(v.x + v.y) * r);
    if (k == 80) return make_float2(-(v.x + v.y) * r, (v.x - v.y) * r);
    if (k == 112) return make_float2((v.x - v.y) * r, (v.x + v.y) * r);
    return cmul(v, make_float2(kTwCos[k], -kTwSin[k]));
}

// Z_q = (X_q + X_{q+h}) + i w_q (X_q - X_{q+h}) with X_{q+h} = conj X_{h-q}, given x = X_q,
// yc = conj X_{h-q} and w_q = exp(-2 pi i k / 128)
__device__ __forceinline__ float2 zterm(float2 x, float2 yc, int k) {
    const float2 t = twmul(csub(x, yc), k);
    return make_float2(x.x + yc.x - t.y, x.y + yc.y + t.x);
}

// ONEVOL: every block's items share one denominator volume (one volume for the call, or volumes
// of whole blocks), whose columns for the block's j then go through LDS.
template <int H, int LH, int MODE, bool ONEVOL, int VT = 0>
__global__ void __launch_bounds__(256) k_apeak(VArg a, const void* __restrict__ v, const float* __restrict__ rden, int rden_div,
                                               unsigned long long* __restrict__ best, float2* __restrict__ vol) {
    // MODE 0 with one volume: the denominators of the block's j, float4 slot (a4 nc + c) holding
    // column c's values 4 a4 .. 4 a4 + 3 (adjacent columns in adjacent slots for the reads)
    extern __shared__ float4 rlds[];
    constexpr int N = 2 * H;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int sl = lane >> a.log2nc;
    const int c = lane & (a.nc - 1);
    // slabs (j, item b) over the padded items, items innermost: a block's slabs share j (ni_pad
    // is a multiple of the slabs per block) and its waves read adjacent items' rows
    const int per_wave = 64 >> a.log2nc;
    const int nw = (int)(blockDim.x >> 6);
    const int per_block = per_wave * nw;
    const long slab0 = (long)blockIdx.x * per_block;
    const int j = (int)(slab0 / a.ni_pad);
    const int b0 = (int)(slab0 - (long)j * a.ni_pad);
    // a block wholly in the padding has nothing to do
    if (b0 >= a.n_items) return;
    const int b = b0 + (tid >> 6) * per_wave + sl;
    const bool live = b < a.n_items;
    constexpr bool one_vol = MODE == 0 && ONEVOL;
    if constexpr (one_vol) {
        // straight from memory into LDS (no registers), 64 slots an instruction, issued before the
        // spectrum's loads
        const float4* src = reinterpret_cast<const float4*>(rden + ((size_t)(b0 / rden_div) * a.nb + j) * a.nc * N);
        const int slots = (a.nc * N) >> 2;
        for (int k = tid >> 6; k * 64 < slots; k += nw) {
            const int sidx = k * 64 + lane;
            const int a4 = sidx >> a.log2nc, cc = sidx & (a.nc - 1);
            if (sidx < slots)
                __builtin_amdgcn_global_load_lds((const void*)(src + cc * (N / 4) + a4), (void __attribute__((address_space(3)))*)(rlds + k * 64),
                                                 16, 0, 0);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    const int wc = VT ? a.nc : a.cols / 2;                  // complex pairs per item row
    constexpr unsigned EB = VEl<VT>::bytes;
    const unsigned rs = (unsigned)(a.ni_pad * wc) * EB;     // bytes per m' row
    const auto vrs = buf((const char*)v + (size_t)j * a.L * rs, (long long)a.L * rs);
    const unsigned vl = (unsigned)(b * wc + c) * EB;
    float2 z[H];
#pragma unroll
    for (int q = 0; q < H; ++q) {
        if constexpr (VT == 0) {
            z[q] = q < a.L ? VEl<VT>::load(vrs, vl, (unsigned)q * rs) : make_float2(0.f, 0.f);
        } else {
            // no branch around a load whose value is converted (the conversion would wait for
            // it there): rows q >= L lie past the resource's L rows and load as zeros, the whole
            // offset in the per-lane register, which the bounds check covers
            z[q] = VEl<VT>::load(vrs, vl + (unsigned)q * rs, 0u);
        }
    }
    // the real transform along a (C = Re V_0 + 2 Re sum_{m' > 0} V_m' e^{-i m' a}) from a complex
    // one of half its length, built in place pairwise (q, h - q); X_0 is taken real
    constexpr int KS = 64 / H;  // w_q = exp(-2 pi i q / n) = exp(-2 pi i (q KS) / 128)
    // For the pair (q, h - q) with s = X_q + conj X_{h-q}, d = X_q - conj X_{h-q}, t = w_q d:
    // Z_q = s + i t and, as w_{h-q} = -conj w_q, Z_{h-q} = conj s + i conj t (one product a pair).
    z[0] = make_float2(z[0].x, z[0].x);
#pragma unroll
    for (int q = 1; q < H / 2; ++q) {
        const float2 x = z[q], y = z[H - q];
        const float2 sm = make_float2(x.x + y.x, x.y - y.y);
        const float2 t = twmul(make_float2(x.x - y.x, x.y + y.y), q * KS);
        z[q] = make_float2(sm.x - t.y, sm.y + t.x);
        z[H - q] = make_float2(sm.x + t.y, t.x - sm.y);
    }
    z[H / 2] = zterm(z[H / 2], make_float2(z[H / 2].x, -z[H / 2].y), (H / 2) * KS);
    // radix-2 decimation in frequency: natural in, bit-reversed out
#pragma unroll
    for (int hh = H / 2; hh >= 1; hh >>= 1) {
#pragma unroll
        for (int blk = 0; blk < H; blk += 2 * hh) {
#pragma unroll
            for (int jj = 0; jj < hh; ++jj) {
                const float2 u = z[blk + jj], w = z[blk + jj + hh];
                z[blk + jj] = cadd(u, w);
                z[blk + jj + hh] = twmul(csub(u, w), jj * (64 / hh));
            }
        }
    }
    // z[p] = (C[2f], C[2f + 1]), f = bitrev(p)
    if (MODE == 1) {
        if (!live) return;
        float2* out = vol + (((size_t)b * a.nb + j) * a.nc + c) * H;
#pragma unroll
        for (int p = 0; p < H; ++p) out[brev_c<LH>(p)] = z[p];
        return;
    }
    float m = -3.4e38f;
    if constexpr (one_vol) {
        // the loads into LDS were issued first: waiting for them waits for no spectrum load
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const float* rd = reinterpret_cast<const float*>(rlds) + c * 4;
#pragma unroll
        for (int p = 0; p < H; ++p) {
            const int f2 = 2 * brev_c<LH>(p);  // values f2, f2 + 1: slot f2 / 4, offset f2 % 4
            const float2 r = *reinterpret_cast<const float2*>(rd + (f2 >> 2) * a.nc * 4 + (f2 & 3));
            z[p] = make_float2(z[p].x * r.x, z[p].y * r.y);
            m = fmaxf(m, fmaxf(z[p].x, z[p].y));
        }
    } else {
        // padding lanes (b >= n_items) read item 0's volume and are masked below
        const float2* rd = reinterpret_cast<const float2*>(rden) + (((size_t)((live ? b : 0) / rden_div) * a.nb + j) * a.nc + c) * H;
#pragma unroll
        for (int p = 0; p < H; ++p) {
            const float2 r = rd[brev_c<LH>(p)];
            z[p] = make_float2(z[p].x * r.x, z[p].y * r.y);
            m = fmaxf(m, fmaxf(z[p].x, z[p].y));
        }
    }
    if (!live) m = -3.4e38f;
    float smax = m;
    for (int o = 1; o < a.nc; o <<= 1) smax = fmaxf(smax, __shfl_xor(smax, o, 64));
    // a slab below the item's best so far (a stale read is never above the true best) cannot
    // hold the argmax; one that reaches it finds its lowest index of that value
    const unsigned long long cur = live ? __hip_atomic_load(best + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
    const bool need = live && (cur == 0ull || smax >= from_ord((unsigned)(cur >> 32)));
    unsigned idx = 0xffffffffu;
    if (need && m == smax) {
        const unsigned base = (unsigned)(j * a.nc + c) * (unsigned)a.n;
#pragma unroll
        for (int p = 0; p < H; ++p) {
            const unsigned f2 = 2u * (unsigned)brev_c<LH>(p);
            if (z[p].x == smax) idx = min(idx, base + f2);
            if (z[p].y == smax) idx = min(idx, base + f2 + 1);
        }
    }
    for (int o = 1; o < a.nc; o <<= 1) idx = min(idx, (unsigned)__shfl_xor((int)idx, o, 64));
    if (need && c == 0) {
        const unsigned long long key = ((unsigned long long)ord_bits(smax) << 32) | (unsigned long long)(0xffffffffu - idx);
        atomicMax(best + b, key);
    }
}

// a per-lane choice of two values (by value: a conditional over two lvalues would select an
// address and push the register array out to scratch memory)
__device__ __forceinline__ float2 pick(bool p, float2 x, float2 y) { return make_float2(p ? x.x : y.x, p ? x.y : y.y); }
__device__ __forceinline__ float2 tw128c(int k) { return make_float2(kTwCos[k & 127], -kTwSin[k & 127]); }
__device__ __forceinline__ float2 tw256c(int k) { return make_float2(kTw256Cos[k], -kTw256Sin[k]); }

// k_apeak for n = 256 (a 128-point half-length transform), whose 128 complex values a column
// would need do not fit one lane's registers: two lanes per column, lane (hf, c), hf = lane / 32,
// register r holding the column's order q = 2 r + hf. The in-place build pairs q with 128 - q,
// which sits in the same lane (register 64 - r for hf 0, 63 - r for hf 1); each half then runs
// the 64-point transform of its own orders (E of the even ones, O of the odd ones) in
// registers, and the one stage that mixes the halves, X[k] = E[k] + w_k O[k] and
// X[k + 64] = E[k] - w_k O[k], trades values once (v_permlane32_swap) on the way to the peak
// search. One wave per (j, 32 columns, item), four items a block on one j and column group,
// whose 32 denominator columns go straight into LDS ahead of the spectrum loads (ONEVOL).
// MODE 0: the peak search of k_apeak; MODE 1: the volume, natural order.
template <int MODE, bool ONEVOL, int VT = 0>
__global__ void __launch_bounds__(256, 2) k_apeak128(VArg a, const void* __restrict__ v, const float* __restrict__ rden, int rden_div,
                                                  unsigned long long* __restrict__ best, float2* __restrict__ vol) {
    constexpr int Q = 64, N = 256;
    extern __shared__ float4 rlds[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool hf = lane >= 32;
    const int cl = lane & 31;
    const int ncw = a.nc >> 5;
    const int groups = (a.n_items + 3) >> 2;
    const int id = blockIdx.x;
    const int g = id % groups;
    const int rest = id / groups;
    const int cw = rest % ncw;
    const int j = rest / ncw;
    const int b0 = g * 4;
    const int b = b0 + (tid >> 6);
    const bool live = b < a.n_items;
    const int c = cw * 32 + cl;
    constexpr bool one_vol = MODE == 0 && ONEVOL;
    if constexpr (one_vol) {
        // float4 slot (a4 32 + column) of the group's 32 columns, 64 slots an instruction
        const float4* src = reinterpret_cast<const float4*>(rden + (((size_t)(b0 / rden_div) * a.nb + j) * a.nc + cw * 32) * N);
#pragma unroll
        for (int k = tid >> 6; k < 32; k += 4) {
            const int sidx = k * 64 + lane;
            const int a4 = sidx >> 5, cc = sidx & 31;
            __builtin_amdgcn_global_load_lds((const void*)(src + cc * (N / 4) + a4), (void __attribute__((address_space(3)))*)(rlds + k * 64), 16, 0,
                                             0);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    const int wc = VT ? a.nc : a.cols / 2;
    constexpr unsigned EB = VEl<VT>::bytes;
    const unsigned rs = (unsigned)(a.ni_pad * wc) * EB;
    const auto vrs = buf((const char*)v + (size_t)j * a.L * rs, (long long)a.L * rs);
    // orders q >= L lie past the resource's L rows and load as zeros (no lane-divergent branch);
    // the whole offset goes in the per-lane register, which the bounds check covers
    const unsigned vl = (unsigned)((live ? b : 0) * wc + c) * EB + (hf ? rs : 0u);
    float2 z[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) z[r] = VEl<VT>::load(vrs, vl + (unsigned)r * 2u * rs, 0u);
    // the half-length input: with s = X_q + conj X_{128-q}, d = X_q - conj X_{128-q} and
    // t = w_q d (w_q = exp(-2 pi i q / 256)), Z_q = s + i t and Z_{128-q} = conj s + i conj t.
    // X_0 is taken real (X_128 = 0 for L <= 128); order 64 (hf 0, r 32) pairs with itself.
    auto pair = [](float2 x, float2 y, float2 w, float2& lo, float2& hi) {
        const float2 sm = make_float2(x.x + y.x, x.y - y.y);
        const float2 t = cmul(make_float2(x.x - y.x, x.y + y.y), w);
        lo = make_float2(sm.x - t.y, sm.y + t.x);
        hi = make_float2(sm.x + t.y, t.x - sm.y);
    };
    {
        float2 lo, hi;
        pair(z[0], z[Q - 1], tw256c(1), lo, hi);  // hf 1: orders 1 and 127
        const float x0 = z[0].x;
        z[0] = pick(hf, lo, make_float2(x0, x0));
        z[Q - 1] = pick(hf, hi, z[Q - 1]);
    }
#pragma unroll
    for (int r = 1; r < Q / 2; ++r) {
        // hf 0: orders 2 r and 128 - 2 r (register 64 - r); hf 1: 2 r + 1 and 127 - 2 r (63 - r)
        const float2 y = pick(hf, z[Q - 1 - r], z[Q - r]);
        float2 lo, hi;
        pair(z[r], y, pick(hf, tw256c(2 * r + 1), tw128c(r)), lo, hi);
        z[r] = lo;
        z[Q - 1 - r] = pick(hf, hi, z[Q - 1 - r]);
        z[Q - r] = pick(hf, z[Q - r], hi);
    }
    z[Q / 2] = pick(hf, z[Q / 2], zterm(z[Q / 2], make_float2(z[Q / 2].x, -z[Q / 2].y), 32));
    // the 64-point transforms within each half: natural in, bit-reversed out
#pragma unroll
    for (int hh = Q / 2; hh >= 1; hh >>= 1) {
#pragma unroll
        for (int blk = 0; blk < Q; blk += 2 * hh) {
#pragma unroll
            for (int jj = 0; jj < hh; ++jj) {
                const float2 u = z[blk + jj], w = z[blk + jj + hh];
                z[blk + jj] = cadd(u, w);
                z[blk + jj + hh] = twmul(csub(u, w), jj * (64 / hh));
            }
        }
    }
    // z[r] = E[k] (hf 0) or O[k] (hf 1), k = bitrev6(r); registers r and r + 32 hold k and k + 1.
    // Swapping lanes 32..63 of register r with lanes 0..31 of register r + 32 leaves E[k], O[k]
    // in the low half and E[k + 1], O[k + 1] in the high half: each lane combines its own k'.
#pragma unroll
    for (int r = 0; r < Q / 2; ++r) {
        const int k = brev_c<6>(r);
        const auto sx = __builtin_amdgcn_permlane32_swap(__float_as_uint(z[r].x), __float_as_uint(z[r + Q / 2].x), false, false);
        const auto sy = __builtin_amdgcn_permlane32_swap(__float_as_uint(z[r].y), __float_as_uint(z[r + Q / 2].y), false, false);
        const float2 e = make_float2(__uint_as_float(sx[0]), __uint_as_float(sy[0]));
        const float2 o = make_float2(__uint_as_float(sx[1]), __uint_as_float(sy[1]));
        const float2 t = cmul(o, pick(hf, tw128c(k + 1), tw128c(k)));
        z[r] = cadd(e, t);
        z[r + Q / 2] = csub(e, t);
    }
    // register r of half hf: X[K + hf] = (C[2 (K + hf)], C[2 (K + hf) + 1]) with
    // K = bitrev6(r mod 32) + 64 (r / 32), always even
    auto kk = [](int r) { return brev_c<6>(r & 31) + 64 * (r >> 5); };
    if (MODE == 1) {
        if (!live) return;
        float2* out = vol + (((size_t)b * a.nb + j) * a.nc + c) * (N / 2) + (hf ? 1 : 0);
#pragma unroll
        for (int r = 0; r < Q; ++r) out[kk(r)] = z[r];
        return;
    }
    float m = -3.4e38f;
    if constexpr (one_vol) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        // values 2 (K + hf), +1: float4 slot (K / 2) 32 + column, floats 2 hf and 2 hf + 1
        const float* rd = reinterpret_cast<const float*>(rlds) + cl * 4 + (hf ? 2 : 0);
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            const float2 w = *reinterpret_cast<const float2*>(rd + (kk(r) >> 1) * 128);
            z[r] = make_float2(z[r].x * w.x, z[r].y * w.y);
            m = fmaxf(m, fmaxf(z[r].x, z[r].y));
        }
    } else {
        const float2* rd = reinterpret_cast<const float2*>(rden + (((size_t)((live ? b : 0) / rden_div) * a.nb + j) * a.nc + c) * N) + (hf ? 1 : 0);
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            const float2 w = rd[kk(r)];
            z[r] = make_float2(z[r].x * w.x, z[r].y * w.y);
            m = fmaxf(m, fmaxf(z[r].x, z[r].y));
        }
    }
    if (!live) return;  // wave-uniform (after the block's barrier)
    float smax = m;
    for (int o = 1; o < 64; o <<= 1) smax = fmaxf(smax, __shfl_xor(smax, o, 64));
    const unsigned long long cur = __hip_atomic_load(best + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!(cur == 0ull || smax >= from_ord((unsigned)(cur >> 32)))) return;  // wave-uniform
    unsigned idx = 0xffffffffu;
    if (m == smax) {
        const unsigned base = (unsigned)(j * a.nc + c) * (unsigned)N + (hf ? 2u : 0u);
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            const unsigned f2 = 2u * (unsigned)kk(r);
            if (z[r].x == smax) idx = min(idx, base + f2);
            if (z[r].y == smax) idx = min(idx, base + f2 + 1);
        }
    }
    for (int o = 1; o < 64; o <<= 1) idx = min(idx, (unsigned)__shfl_xor((int)idx, o, 64));
    if (lane == 0) {
        const unsigned long long key = ((unsigned long long)ord_bits(smax) << 32) | (unsigned long long)(0xffffffffu - idx);
        atomicMax(best + b, key);
    }
}

__device__ __forceinline__ float vertex_v(float lo, float mid, float hi) {
    const float den = lo - 2.f * mid + hi;
    if (den >= -1e-12f) return 0.f;
    return fminf(fmaxf(0.5f * (lo - hi) / den, -0.5f), 0.5f);
}

template <int VT>
__global__ void __launch_bounds__(64) k_vfinish(VArg a, const void* __restrict__ v, const float2* __restrict__ tw, const float* __restrict__ rden,
                                                int rden_div, const unsigned long long* __restrict__ best, float* __restrict__ peaks) {
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
    const int wc = VT ? a.nc : a.cols / 2;
    auto V = [&](int jj, int mp, int cc) { return VEl<VT>::at(v, (((size_t)jj * a.L + mp) * a.ni_pad + b) * wc + cc); };
    auto term = [&](float2 x, int mp, int at) {
        const float2 e = tw[((mp * at) & (n - 1)) * ts];
        return x.x * e.x - x.y * e.y;
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
    const float* rd = rden + (size_t)(b / rden_div) * a.nb * nc * n;
    auto R = [&](int jj, int cc, int at) { return rd[((size_t)jj * nc + cc) * n + at]; };
    const float da = vertex_v(s[0] * R(j, iq, iam), v0, s[1] * R(j, iq, iap));
    const float dc = vertex_v(s[2] * R(j, iqm, ia), v0, s[3] * R(j, iqp, ia));
    float dj = 0.f;
    if (inner) dj = vertex_v(s[4] * R(j - 1, iq, ia), v0, s[5] * R(j + 1, iq, ia));
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
__global__ void __launch_bounds__(256) k_vrden(const float* __restrict__ vf, const float* __restrict__ vf2, const float* __restrict__ stats,
                                               long per_item, long total, float* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const float ar = fmaxf(stats[(i / per_item) * 4], 1e-12f);
    const float m = vf[i];
    const float x = vf2[i] - m * m / ar;
    out[i] = x > 1e-6f * ar ? 1.f / sqrtf(x) : 0.f;
}

}  // namespace zb

using namespace zb;

static int ilog2v(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return (1 << r) == v ? r : -1;
}

static int v_arg(const zb_vtable* t, int n_items, VArg& a, const char* who) {
    if (!t || t->L < 2 || ilog2v(t->n) < 4 || t->n < 2 * t->L - 1 || ilog2v(t->nc) < 1 || t->nc > (t->n == 256 ? 128 : 64) || t->n / t->nc > 4 ||
        !(t->l_step == 1 || t->l_step == 2) || t->slots != (t->L + t->l_step - 1) / t->l_step || t->cols != 32 * ((2 * t->nc + 31) / 32) ||
        t->n / 2 > 128 || (t->n == 256 && t->nc < 32) || !t->q) {
        set_error("%s: bad table (L %d, n %d, nc %d, slots %d, cols %d; this path takes n <= 256, nc <= 64, or nc <= 128 at n = 256)", who,
                  t ? t->L : 0, t ? t->n : 0,
                  t ? t->nc : 0, t ? t->slots : 0, t ? t->cols : 0);
        return ZB_EINVAL;
    }
    if (t->slots > (t->n == 256 && !t->lp_kind ? 96 : 64)) {
        set_error("%s: %d degree slots: at most 64 (96 at n = 256 with FP32 products)", who, t->slots);
        return ZB_EINVAL;
    }
    if (t->lp_kind && (t->lp_kind < 1 || t->lp_kind > 2 || !t->qlp || t->lp_split < 0 || t->lp_split % 16 || t->nc < 32 || t->nc > 64)) {
        set_error("%s: bad 16-bit table (kind %d, split slot %d, nc %d; kinds 1 FP16, 2 BF16, the split a multiple of 16, 32 <= nc <= 64)", who,
                  t->lp_kind, t->lp_split, t->nc);
        return ZB_EINVAL;
    }
    a.lp_gs = t->lp_kind ? t->lp_split / 8 : 0;
    a.vk = t->lp_kind;
    a.best = nullptr;
    a.L = t->L;
    a.n = t->n;
    a.log2n = ilog2v(t->n);
    a.nc = t->nc;
    a.log2nc = ilog2v(t->nc);
    a.nb = t->n / 2 + 1;
    if (t->b_half < 0 || t->b_half > 1) {
        set_error("%s: b_half %d (0 or 1)", who, t->b_half);
        return ZB_EINVAL;
    }
    a.jp = t->b_half ? t->n / 4 + 1 : a.nb;
    a.jg = t->b_half ? t->n / 4 + 2 : a.nb;
    a.slots = t->slots;
    a.cols = t->cols;
    a.l_step = t->l_step;
    a.pairs = t->L * (t->L + 1) / 2;
    a.n_items = n_items;
    // the items of a row, padded to whole blocks, and 32 more so that consecutive rows (one per
    // (j, m')) do not sit a power of two apart in HBM
    a.ni_pad = (n_items + kItemPad - 1) / kItemPad * kItemPad + 32;
    a.tiles = 1;
    // the transform along a addresses one j's rows of V with 32-bit byte offsets (at n = 256 up
    // to row 127, past the L loaded ones, where the bounds check returns zeros)
    const int64_t rows = a.n == 256 ? 128 : a.L;
    if ((int64_t)a.ni_pad * (a.cols / 2) * 8 * rows > 0x7fffffffLL) {
        set_error("%s: %d items in one call: at most %lld at this size", who, n_items,
                  (long long)(0x7fffffffLL / ((int64_t)(a.cols / 2) * 8 * rows) - kItemPad - 32));
        return ZB_EINVAL;
    }
    return ZB_OK;
}

static int64_t al256(int64_t v) { return (v + 255) / 256 * 256; }
// V for the padded items: cols floats per item row (complex pairs, then the table's zero columns)
static int64_t v_bytes(const VArg& a) { return (int64_t)a.ni_pad * a.jg * a.L * (a.vk ? 2 * a.nc : a.cols) * (a.vk ? 2 : 4); }
// V and the per-item best; the denominators add two volumes
static int64_t vcorr_ws(const VArg& a, int volumes) {
    return al256(v_bytes(a)) + al256((int64_t)a.n_items * 8) + (volumes ? 2 * al256((int64_t)a.n_items * a.nb * a.n * a.nc * 4) : 0);
}

extern "C" int64_t zb_vcorr_workspace(const zb_vtable* t, int32_t n_items, int32_t denominators) {
    VArg a;
    if (n_items < 0) {
        set_error("zb_vcorr_workspace: n_items %d", n_items);
        return ZB_EINVAL;
    }
    if (v_arg(t, n_items, a, "zb_vcorr_workspace")) return ZB_EINVAL;
    if (denominators) a.jg = a.jp = a.nb;  // the window's volumes take every row of b
    return vcorr_ws(a, denominators != 0);
}

template <int MB, int NBLK, int S>
static void launch_vgemm_t(const VArg& a, const float* q, const float* alm, float* v, hipStream_t s) {
    // a block takes kWaves tiles of 32 MB items (ni_pad is a multiple of that)
    const int tiles = (a.n_items + kWaves * 32 * MB - 1) / (kWaves * 32 * MB);
    VArg b = a;
    b.tiles = tiles;
    const int ncg = a.cols / (NBLK * 32);
    const size_t lds = 2 * ((size_t)(S / 4) * NBLK * 64 + 1) * 16;  // two buffers of one j's rows (and a spare slot)
    ZB_LAUNCH("k_vgemm", (k_vgemm<MB, NBLK, S>), dim3((unsigned)(tiles * a.L * ncg)), dim3(64 * kWaves), lds, s, b, q, (const float2*)alm, v);
}

template <int NBLK, int G8, int GS, bool BF>
static void launch_vgemm16_t(const VArg& a, const void* q, const float* alm, void* v, hipStream_t s) {
    const int tiles = (a.n_items + kWaves * 32 - 1) / (kWaves * 32);
    VArg b = a;
    b.tiles = tiles;
    const int ncg = a.cols / (NBLK * 32);
    const size_t lds = 2 * ((size_t)(2 * G8 - GS) * NBLK * 64 + 1) * 16;
    ZB_LAUNCH("k_vgemm", (k_vgemm16<1, NBLK, G8, GS, BF>), dim3((unsigned)(tiles * a.L * ncg)), dim3(64 * kWaves), lds, s, b, (const floatx4*)q,
              (const float2*)alm, (unsigned*)v);
}

// The 16-bit table's tiling: G8 groups of 8 slots (an even count), nblk column blocks of 32.
static void vgemm16_shape(int nc, int slots, int& cols, int& g8, int& nblk) {
    cols = 32 * ((2 * nc + 31) / 32);
    g8 = 2 * ((slots + 15) / 16);
    nblk = cols == 32 ? 1 : 2;
}

template <int NBLK, int G8, bool BF, int GS = 0>
static void launch_vgemm16_gs(const VArg& a, const void* q, const float* alm, void* v, hipStream_t s) {
    if constexpr (GS < G8) {
        if (a.lp_gs > GS) return launch_vgemm16_gs<NBLK, G8, BF, GS + 2>(a, q, alm, v, s);
    }
    launch_vgemm16_t<NBLK, G8, (GS < G8 ? GS : G8), BF>(a, q, alm, v, s);
}

template <int NBLK, bool BF>
static void launch_vgemm16_g(int g8, const VArg& a, const void* q, const float* alm, void* v, hipStream_t s) {
    if (g8 == 2) launch_vgemm16_gs<NBLK, 2, BF>(a, q, alm, v, s);
    else if (g8 == 4) launch_vgemm16_gs<NBLK, 4, BF>(a, q, alm, v, s);
    else if (g8 == 6) launch_vgemm16_gs<NBLK, 6, BF>(a, q, alm, v, s);
    else launch_vgemm16_gs<NBLK, 8, BF>(a, q, alm, v, s);
}

// V as 16-bit pairs of the products' type (the real and imaginary column blocks of a group
// together: nblk 2, nc >= 32)
static int launch_vgemm16(const VArg& a, int kind, const void* q, const float* alm, void* v, hipStream_t s) {
    int cols, g8, nblk;
    vgemm16_shape(a.nc, a.slots, cols, g8, nblk);
    if (kind == 2) launch_vgemm16_g<2, true>(g8, a, q, alm, v, s);
    else launch_vgemm16_g<2, false>(g8, a, q, alm, v, s);
    return check_launch("k_vgemm");
}

extern "C" int32_t zb_vtable16_shape(int32_t nc, int32_t slots, int32_t* cols, int32_t* groups, int32_t* nblk) {
    if (nc < 2 || nc > 64 || (nc & (nc - 1)) || slots < 1 || slots > 64 || !cols || !groups || !nblk) {
        set_error("zb_vtable16_shape: nc %d, slots %d", nc, slots);
        return ZB_EINVAL;
    }
    int c, g, b;
    vgemm16_shape(nc, slots, c, g, b);
    *cols = c;
    *groups = g;
    *nblk = b;
    return ZB_OK;
}

// The choice of (S, NBLK) fixes the table's tiling (zb_vtable_shape).
static void vgemm_shape(int nc, int slots, int& cols, int& S, int& nblk) {
    cols = 32 * ((2 * nc + 31) / 32);
    S = slots <= 8 ? 8 : slots <= 16 ? 16 : slots <= 32 ? 32 : slots <= 48 ? 48 : slots <= 64 ? 64 : 96;
    nblk = (cols == 32 || S >= 64) ? 1 : 2;
}

static int launch_vgemm(const VArg& a, const float* q, const float* alm, float* v, hipStream_t s) {
    int cols, S, nblk;
    vgemm_shape(a.nc, a.slots, cols, S, nblk);
    // one tile of 32 items per wave: the coefficients (S registers), two sets of accumulators
    // (2 NBLK 16) and the staged table rows leave room for two or three waves per SIMD
#define ZB_VG(NB, SS) launch_vgemm_t<1, NB, SS>(a, q, alm, v, s)
    if (S == 96) {
        ZB_VG(1, 96);
    } else if (S == 64) {
        ZB_VG(1, 64);
    } else if (nblk == 1) {
        if (S == 8) ZB_VG(1, 8); else if (S == 16) ZB_VG(1, 16); else if (S == 32) ZB_VG(1, 32); else ZB_VG(1, 48);
    } else {
        if (S == 8) ZB_VG(2, 8); else if (S == 16) ZB_VG(2, 16); else if (S == 32) ZB_VG(2, 32); else ZB_VG(2, 48);
    }
#undef ZB_VG
    return check_launch("k_vgemm");
}

extern "C" int32_t zb_vtable_shape(int32_t nc, int32_t slots, int32_t* cols, int32_t* S, int32_t* nblk) {
    if (nc < 2 || nc > 128 || (nc & (nc - 1)) || slots < 1 || slots > 96 || !cols || !S || !nblk) {
        set_error("zb_vtable_shape: nc %d, slots %d", nc, slots);
        return ZB_EINVAL;
    }
    int c, s, b;
    vgemm_shape(nc, slots, c, s, b);
    *cols = c;
    *S = s;
    *nblk = b;
    return ZB_OK;
}

template <int MODE, int VT = 0>
static int launch_apeak(const VArg& a, const void* v, const float* rden, int rden_div, unsigned long long* best, float* vol, hipStream_t s) {
    if constexpr (MODE == 0 && VT == 0) {
        if (a.vk == 1) return launch_apeak<0, 1>(a, v, rden, rden_div, best, vol, s);
        if (a.vk == 2) return launch_apeak<0, 2>(a, v, rden, rden_div, best, vol, s);
    }
    if (a.n == 256) {
        // two lanes per column: one wave per (j, 32 columns, item), four items a block
        const bool onevol = MODE == 0 && (rden_div >= a.n_items || rden_div % 4 == 0);
        const dim3 grid((unsigned)((long)a.jp * (a.nc / 32) * ((a.n_items + 3) / 4)));
        const size_t lds = onevol ? (size_t)32 * 256 * sizeof(float) : 0;
        if (onevol)
            ZB_LAUNCH("k_apeak", (k_apeak128<MODE, true, VT>), grid, dim3(256), lds, s, a, v, rden, rden_div, best, (float2*)vol);
        else
            ZB_LAUNCH("k_apeak", (k_apeak128<MODE, false, VT>), grid, dim3(256), lds, s, a, v, rden, rden_div, best, (float2*)vol);
        return check_launch("k_apeak");
    }
    const int per_wave = 64 >> a.log2nc;
    // up to four waves a block, at most 32 slabs (ni_pad is a multiple of 32)
    const int waves = per_wave >= 32 ? 1 : (32 / per_wave < 4 ? 32 / per_wave : 4);
    const int per_block = per_wave * waves;
    const long slabs = (long)a.ni_pad * a.jp;
    const dim3 grid((unsigned)(slabs / per_block));
    const int H = a.n / 2;
    // blocks start at multiples of per_block, so volumes of a multiple of per_block items never
    // split a block
    const bool onevol = MODE == 0 && (rden_div >= a.n_items || rden_div % per_block == 0);
    const size_t lds = onevol ? (size_t)a.nc * a.n * sizeof(float) : 0;
#define ZB_AP(HH, LL)                                                                                                              \
    do {                                                                                                                           \
        if (onevol)                                                                                                                \
            ZB_LAUNCH("k_apeak", (k_apeak<HH, LL, MODE, true, VT>), grid, dim3(64 * waves), lds, s, a, v, rden, rden_div, best,        \
                      (float2*)vol);                                                                                               \
        else                                                                                                                       \
            ZB_LAUNCH("k_apeak", (k_apeak<HH, LL, MODE, false, VT>), grid, dim3(64 * waves), lds, s, a, v, rden, rden_div, best,       \
                      (float2*)vol);                                                                                               \
    } while (0)
    if constexpr (VT == 0) {
        if (H == 8) ZB_AP(8, 3); else if (H == 16) ZB_AP(16, 4); else if (H == 32) ZB_AP(32, 5); else ZB_AP(64, 6);
    } else {
        ZB_AP(64, 6);  // 16-bit V: nc >= 32, so n >= 128
    }
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
    if (work_bytes < vcorr_ws(a, 0)) {
        set_error("zb_vcorr_correlate: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)vcorr_ws(a, 0));
        return ZB_ESPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    float* v = (float*)work;
    unsigned long long* best = (unsigned long long*)((char*)work + al256(v_bytes(a)));
    a.best = best;  // cleared by the GEMM, before the peak search's atomics
    if (int rc = t->lp_kind ? launch_vgemm16(a, t->lp_kind, t->qlp, alm, v, s) : launch_vgemm(a, t->q, alm, v, s)) return rc;
    if (int rc = launch_apeak<0>(a, v, rden, rden_div, best, nullptr, s)) return rc;
    if (a.vk == 0)
        ZB_LAUNCH("k_vfinish", k_vfinish<0>, dim3(n_items), dim3(64), 0, s, a, (const void*)v, (const float2*)twiddle, rden, rden_div,
                  (const unsigned long long*)best, peaks);
    else if (a.vk == 1)
        ZB_LAUNCH("k_vfinish", k_vfinish<1>, dim3(n_items), dim3(64), 0, s, a, (const void*)v, (const float2*)twiddle, rden, rden_div,
                  (const unsigned long long*)best, peaks);
    else
        ZB_LAUNCH("k_vfinish", k_vfinish<2>, dim3(n_items), dim3(64), 0, s, a, (const void*)v, (const float2*)twiddle, rden, rden_div,
                  (const unsigned long long*)best, peaks);
    return check_launch("zb_vcorr_correlate/finish");
}

extern "C" int32_t zb_vcorr_denominators(const zb_vtable* t_f, const zb_vtable* t_f2, const float* w_alm, const float* w_stats, int32_t n_windows,
                                         void* work, int64_t work_bytes, float* rden, void* stream) {
    VArg a, a2;
    if (int rc = v_arg(t_f, n_windows, a, "zb_vcorr_denominators")) return rc;
    if (int rc = v_arg(t_f2, n_windows, a2, "zb_vcorr_denominators")) return rc;
    a.vk = a2.vk = 0;  // the window's volumes from the FP32 tables q
    a.jg = a.jp = a.nb;  // and over every row of b: the window has no symmetry
    a2.jg = a2.jp = a2.nb;
    if (a.L != a2.L || a.n != a2.n || a.nc != a2.nc || a.l_step != a2.l_step || n_windows < 0 ||
        (n_windows > 0 && (!w_alm || !w_stats || !work || !rden))) {
        set_error("zb_vcorr_denominators: bad arguments");
        return ZB_EINVAL;
    }
    if (n_windows == 0) return ZB_OK;
    if (work_bytes < vcorr_ws(a, 1)) {
        set_error("zb_vcorr_denominators: workspace %lld bytes, need %lld", (long long)work_bytes, (long long)vcorr_ws(a, 1));
        return ZB_ESPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    float* v = (float*)work;
    char* after = (char*)work + al256(v_bytes(a)) + al256((int64_t)n_windows * 8);
    const int64_t vb = al256((int64_t)n_windows * a.nb * a.n * a.nc * 4);
    float* vf = (float*)after;
    float* vf2 = (float*)(after + vb);
    if (int rc = launch_vgemm(a, t_f->q, w_alm, v, s)) return rc;
    if (int rc = launch_apeak<1>(a, v, nullptr, 1, nullptr, vf, s)) return rc;
    if (int rc = launch_vgemm(a2, t_f2->q, w_alm, v, s)) return rc;
    if (int rc = launch_apeak<1>(a, v, nullptr, 1, nullptr, vf2, s)) return rc;
    const long per_item = (long)a.nb * a.n * a.nc;
    const long total = per_item * n_windows;
    ZB_LAUNCH("k_vrden", k_vrden, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, vf, vf2, w_stats, per_item, total, rden);
    return check_launch("zb_vcorr_denominators/rden");
}
