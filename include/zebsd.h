/*
 * zebsd.h — C ABI of the MI355X (gfx950) indexing hot path: libzebsd_hip.so.
 *
 * Stateless kernel entry points over caller-owned device memory. Every pointer here that is
 * not a struct of parameters is a device pointer (hipMalloc'd, or a torch tensor's data_ptr);
 * the structs themselves live in host memory and are read during the call. Every call enqueues
 * work on `stream` (a hipStream_t, NULL = the null stream) and returns without waiting.
 *
 * Return value: 0 (or a size, for the *_workspace queries) on success; a negative code on
 * failure, with the reason kept per thread and read by zb_last_error. This replaces the
 * reference's engine-held error text (src/exports.zig:22-40, `ze_error`).
 *
 * The reference binds its path through an engine object (src/exports.zig, src/native_api.zig,
 * python/zebsdsolve/_native.py:95-174). Its kernels are WGSL entry points recorded by
 * src/spherical.zig and src/detector.zig; each entry point below names the recording function
 * and the WGSL entry points it replaces. The engine-level surface (ze.Solver, ze.index) is
 * mirrored above this ABI in Python (zebsdsolve_amd/__init__.py).
 */
#ifndef ZEBSD_H
#define ZEBSD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define ZB_OK 0
#define ZB_EINVAL -1   /* a bad argument: size, shape, null pointer */
#define ZB_ESPACE -2   /* workspace too small */
#define ZB_ELAUNCH -3  /* a kernel launch failed */

/* Flags of zb_sh_coefficients (src/spherical.zig:96-98). */
#define ZB_FLAG_CENTER 1 /* subtract the window mean */
#define ZB_FLAG_NORM 2   /* scale to unit harmonic norm */
#define ZB_FLAG_WINDOW 4 /* project the window itself (ones) instead of a pattern */
/* The Legendre sums with 16-bit products and FP32 accumulation (the SH-analysis GEMM of the
 * precision form; no counterpart in the reference): FP16 or BF16 on the matrix cores, each value as
 * hi = round16(x) and lo = round16(x - hi), hi lo + lo hi added for the degrees from the split
 * degree (bits 16..23 of the flags: ZB_FLAG_SPLIT(l)) on. Below n = 32 the FP32 sums are used. */
#define ZB_FLAG_FP16 8
#define ZB_FLAG_BF16 16
#define ZB_FLAG_SPLIT_SHIFT 16
#define ZB_FLAG_SPLIT(l) (((l) & 0xff) << ZB_FLAG_SPLIT_SHIFT)

/* The detector's attitude and circular mask in pixels of the patterns worked on
 * (spherical.Frame, src/spherical.zig:80-90). */
typedef struct zb_frame {
    float c0[3]; /* detector row direction in the sample frame */
    float c1[3]; /* column direction */
    float c2[3]; /* normal */
    int32_t pat_h, pat_w;
    float mask[3]; /* centre row, centre column, squared radius (<= 0: no mask) */
} zb_frame;

/* What depends only on the bandwidth L (spherical.Tables, src/spherical.zig:106-190). Device
 * pointers. n = fft_size(L): the sphere grid has n rings of 2n points, the orientation grid n
 * points per turn. */
typedef struct zb_band {
    int32_t L;
    int32_t n;
    const float* ring;    /* [n][2]: cos(theta), quadrature weight; Gauss-Legendre, so ring n - 1 - y mirrors ring y
                             (cos theta negated, the same weight), which the sampling relies on */
    const float* wplm;    /* [n/2][L(L+1)/2]: w_y P_l^m(cos theta_y) 2 pi / (2n) */
    const float* dq;      /* [L][L][L]: d^l_{k,m}(pi/2) at (m L + l) L + k */
    const float* twiddle; /* [1024][2]: exp(-2 pi i t / 1024) */
} zb_band;

/* What a phase adds at a bandwidth (spherical.PhaseGpu, src/spherical.zig:192-242). */
typedef struct zb_phase {
    int32_t fold;   /* 1, 2 or 4: the master's orders m are multiples of it */
    int32_t nc;     /* n / fold: points of the c axis */
    int32_t mu_max; /* (L - 1) / fold */
    int32_t l_step; /* 2 when the master has no odd l, else 1 */
    const float* w_f;  /* [2 mu_max + 1][L][L][2]: i^{-m} f_lm d^l_{km}(pi/2), k innermost */
    const float* w_f2; /* the same for the master's square */
} zb_phase;

/* A phase's correlation as a GEMM table at a bandwidth: the reference's transforms along b and c
 * of the spectrum (so3.wgsl spectrum, a_peak_m) folded into a pattern-independent table,
 *   Q[j][m'][c][l] = sum_mu f_{l, fold mu} d^l_{m', fold mu}(2 pi j / n) e^{-2 pi i mu c / nc},
 * so that V[j][m'][c] = sum_{l >= m'} Q[j][m'][c][l] conj(g_{l m'}). As matrix-core operands,
 * per (j, m') and degree slot s (l = s l_step) two rows of `cols` floats: row 0 holds
 * (Re Q_c, Im Q_c) at columns (2c, 2c + 1), row 1 holds (Im Q_c, -Re Q_c), columns past 2 nc
 * zero. With (cols, S, nblk) from zb_vtable_shape, q stores them tiled for 16-byte loads: per
 * (j, m', column group g < cols / (32 nblk)), [S/4][nblk][64][4] floats, element
 * [t][b][32 r + x][i] = row r, slot 4t + i, column 32 (g nblk + b) + x (zero for slots past the
 * last). Built once per phase and bandwidth (zebsdsolve_amd/tables.py VTableHost). No
 * counterpart in the reference's tables; it replaces W_f of spherical.PhaseGpu
 * (src/spherical.zig:192-242) on this path.
 *
 * With lp_kind 1 (FP16) or 2 (BF16) (nc >= 32), zb_vcorr_correlate multiplies in 16 bits with
 * FP32 accumulation and keeps V as pairs of the same 16-bit type (the transform along a, the
 * peak and its neighbours in FP32), from qlp: the same rows with each value x as
 * hi = round16(x) and lo = round16(x - hi), tiled by (cols, G, 2) from zb_vtable16_shape: per
 * (j, m', column group g < nc / 32), the G groups of 8 slots in order, each [2][64][8] hi
 * values, element [b][32 r + x][i] = row r, slot 8t + i, column 2 (32 g + x) + b (b = 0 the
 * real, 1 the imaginary part of complex column 32 g + x), followed, for groups t >= lp_split / 8,
 * by the same for lo. Those groups take hi hi + hi lo + lo hi (the item's coefficients split
 * the same way), the others hi hi. Then q is used only by zb_vcorr_denominators (the window's
 * volumes stay FP32). lp_kind 0: FP32 products and V from q. */
typedef struct zb_vtable {
    int32_t L;
    int32_t n;      /* fft_size(L), at most 256 on this path */
    int32_t nc;     /* n / fold: at most 64, or 128 at n = 256 */
    int32_t l_step; /* 2 when the master has no odd l, else 1 */
    int32_t slots;  /* ceil(L / l_step) */
    int32_t cols;
    const float* q; /* tiled, (n/2 + 1) L S 2 cols floats */
    int32_t lp_kind;  /* 0 FP32 products, 1 FP16, 2 BF16 */
    int32_t lp_split; /* first slot (a multiple of 8) of the groups with hi + lo products */
    const void* qlp;  /* the 16-bit table (lp_kind != 0), 2-byte values */
    /* 1: the master has a 2-fold axis along y, so its correlation obeys C(a, b, c) =
     * C(a + pi, pi - b, pi - c) and every orientation has a representative with b <= pi / 2:
     * zb_vcorr_correlate searches the rows j <= n / 4 only (and computes one more, the peak's
     * neighbour along b). Ties keep the lowest index, as over the whole range: a cell past
     * n / 4 has its partner at a lower j. 0: every row j <= n / 2. zb_vcorr_denominators ignores
     * it (the window has no such symmetry). */
    int32_t b_half;
} zb_vtable;

/* The tiling of zb_vtable.q for nc points along c (up to 128) and `slots` degree slots (up to
 * 96): cols, S (slots per wave, the first of 8, 16, 32, 48, 64, 96 >= slots) and nblk (column
 * blocks of 32 per wave: 1 when cols = 32 or S >= 64, else 2). */
int32_t zb_vtable_shape(int32_t nc, int32_t slots, int32_t* cols, int32_t* S, int32_t* nblk);

/* The tiling of zb_vtable.qlp: cols, groups (of 8 slots: 2 ceil(slots / 16)) and nblk (1 when
 * cols = 32, else 2). */
int32_t zb_vtable16_shape(int32_t nc, int32_t slots, int32_t* cols, int32_t* groups, int32_t* nblk);

/* Gauss-Newton refinement options (detector.RefineOptions, src/detector.zig:57-70). */
typedef struct zb_refine_opts {
    int32_t iters;    /* Gauss-Newton iterations (0: score the seeds only) */
    float max_step;   /* largest rotation of one iteration, radians */
    float scale0;     /* curvature scale a refinement starts with */
    float scale_min;
    float scale_max;
    int32_t nmi_bins; /* histogram levels of the normalized mutual information, 2..32 */
    uint32_t stamp;   /* written into each record's last slot */
    /* a permutation of the n_items items (device memory, zb_refine_order's), or NULL: with it the
     * items are taken in that order, each XCD a contiguous run of it, so that items refined side
     * by side look up nearby regions of the master; the records are the same either way */
    const int32_t* order;
} zb_refine_opts;

/* Library version string and the last error of this thread (length written, without NUL). */
const char* zb_version(void);
int32_t zb_last_error(char* buf, int32_t cap);

/* Patterns -> float32, zero mean and unit norm under the mask, zero outside it.
 * raw: n patterns of h x w, pixel 0 u8, 1 u16, 2 f32; out: [n][h][w] f32.
 * Replaces Prep.load + the `normalize` filter (src/prep.zig:143-204, 263-272; prep.wgsl:57-70,
 * 399-443). */
int32_t zb_normalize(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], float* out,
                     void* stream);

/* Bytes of workspace zb_sh_coefficients needs for n_items. */
int64_t zb_sh_workspace(const zb_band* band, int32_t n_items);

/* Harmonic coefficients of items: each pattern resampled onto the Gauss-Legendre sphere grid
 * from its detector position, ring spectra, Legendre sums, norm.
 * items: [n_items][4] f32 = (t in pixels, pattern index); pats: [*][h][w] f32 (unused with
 * ZB_FLAG_WINDOW, may be NULL); alm: [n_items][L(L+1)/2][2] f32 out; stats: [n_items][4] out
 * (window area, window mean, harmonic norm, a_00).
 * Replaces Spherical.coefficients (src/spherical.zig:319-361: sphere.wgsl sphere_sums, rings,
 * legendre, sh_norm). */
int32_t zb_sh_coefficients(const zb_band* band, const zb_frame* frame, const float* pats, const float* items,
                           int32_t n_items, int32_t flags, void* work, int64_t work_bytes, float* alm, float* stats,
                           void* stream);

/* zb_sh_coefficients without a workspace argument, for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Its workspace is one static device array the library holds on
 * each device it runs on until the process ends: ZB_SH_SCRATCH_BYTES (64 MiB), enough for
 * 1 023 items at L = 64 (65 536 bytes an item and 8 more), 4 094 at L = 32 and 341 at L = 96; more items
 * fail with ZB_ESPACE and the limit in the message. It allocates nothing and clears nothing (every
 * scratch byte it reads it has written in the same call). Calls that overlap on one device share
 * that array and race: serialize them (one stream, or events between streams). */
#define ZB_SH_SCRATCH_BYTES (64LL * 1024 * 1024)
int32_t zb_sh_coefficients_noalloc(const zb_band* band, const zb_frame* frame, const float* pats, const float* items, int32_t n_items,
                                   int32_t flags, float* alm, float* stats, void* stream);

/* Bytes of workspace zb_so3_correlate / zb_so3_denominators need for n_items. */
int64_t zb_so3_workspace(const zb_band* band, const zb_phase* phase, int32_t n_items);

/* Best orientation of each item over SO(3): the Wigner-d contraction of its coefficients with
 * the phase's weights, transforms along b, c and a, times the reciprocal denominator volume of
 * item b / rden_div, the argmax (ties: lowest index) and its parabolic sub-cell peak.
 * alm: [n_items][pairs][2]; rden: [*][n/2+1][nc][n] f32; peaks: [n_items][8] out =
 * quaternion (w, x, y, z), score, a, b, c.
 * Replaces Spherical.correlate (src/spherical.zig:401-432: so3.wgsl spectrum, a_peak_m,
 * peak_reduce, peak_finish_m). */
int32_t zb_so3_correlate(const zb_band* band, const zb_phase* phase, const float* alm, int32_t n_items,
                         const float* rden, int32_t rden_div, void* work, int64_t work_bytes, float* peaks,
                         void* stream);

/* Reciprocal denominator volumes of the normalized correlation for n_windows window
 * coefficient sets (zb_sh_coefficients with ZB_FLAG_WINDOW): 1 / sqrt(<w,f^2> - <w,f>^2 / area),
 * 0 where that is <= 1e-6 area. w_alm: [n][pairs][2]; w_stats: [n][4]; rden: [n][n/2+1][nc][n] out.
 * Replaces Spherical.denominators (src/spherical.zig:446-464: so3.wgsl a_volume, rden_build). */
int32_t zb_so3_denominators(const zb_band* band, const zb_phase* phase, const float* w_alm, const float* w_stats,
                            int32_t n_windows, void* work, int64_t work_bytes, float* rden, void* stream);

/* Bytes of workspace zb_vcorr_correlate (denominators = 0) or zb_vcorr_denominators (1) needs
 * for n_items. */
int64_t zb_vcorr_workspace(const zb_vtable* table, int32_t n_items, int32_t denominators);

/* zb_so3_correlate on the matrix cores: V = the table times each item's coefficients (float32
 * MFMA), then the real transform along a, times the reciprocal denominator, argmax (ties: lowest
 * index (j nc + c) n + a) and the parabolic sub-cell peak, as zb_so3_correlate.
 * twiddle: [1024][2] (zb_band.twiddle); rden: [*][n/2+1][nc][n] f32 in natural order (a
 * innermost), as zb_vcorr_denominators writes it; peaks: [n_items][8] out, as zb_so3_correlate.
 * Replaces Spherical.correlate (src/spherical.zig:401-432: so3.wgsl spectrum, a_peak_m,
 * peak_reduce, peak_finish_m). */
int32_t zb_vcorr_correlate(const zb_vtable* table, const float* twiddle, const float* alm, int32_t n_items, const float* rden,
                           int32_t rden_div, void* work, int64_t work_bytes, float* peaks, void* stream);

/* zb_so3_denominators on the matrix cores, from the tables of the master (t_f) and of its
 * square (t_f2): rden [n_windows][n/2+1][nc][n] out, natural order (a innermost).
 * Replaces Spherical.denominators (src/spherical.zig:446-464: so3.wgsl a_volume, rden_build). */
int32_t zb_vcorr_denominators(const zb_vtable* t_f, const zb_vtable* t_f2, const float* w_alm, const float* w_stats,
                              int32_t n_windows, void* work, int64_t work_bytes, float* rden, void* stream);

/* Gauss-Newton refinement of each item's orientation on the detector, then its NMI, into one
 * record per item: records[(out0 + b) * 8 ..] = best q (w, x, y, z), NCC, NMI, last improving
 * step (radians), stamp. master: [2][side][side] f32, the (smoothed) master whose gradient
 * field zb_master_gradient made; mgrad_h: that field in half precision (zb_master_pack), which
 * the Gauss-Newton passes read; the recorded NCC and NMI are those of the float32 master at the
 * best orientation. items: [n][4] (t, pattern index); seeds: orientation of item b at
 * seeds[b * seed_stride .. + 4]; pats: normalized patterns [*][h][w].
 * Replaces Detector.seed + Detector.refine (src/detector.zig:178-212: detector.wgsl seed,
 * refine_init, sums_rot, refine_rot_step, nmi, collect), fused into one launch. */
int32_t zb_refine(const float* master, const uint16_t* mgrad_h, int32_t side, const zb_frame* frame, const float* pats,
                  const float* items, const float* seeds, int32_t seed_stride, int32_t n_items, const zb_refine_opts* opts,
                  float* records, int32_t out0, void* stream);

/* One phase of zb_refine_phases: zb_refine's per-phase arguments. */
typedef struct zb_refine_phase {
    const float* master;      /* [2][side][side] f32 */
    const uint16_t* mgrad_h;  /* the master's half-precision gradient field */
    int32_t side;
    const float* seeds;       /* item b's orientation at seeds[b * seed_stride .. + 4] */
    int32_t seed_stride;
    const int32_t* order;     /* zb_refine_order's permutation of the items for this phase, or NULL */
    float* records;           /* [*][8]: item b's record at (out0 + b) * 8 */
} zb_refine_phase;

/* Bytes of device workspace zb_refine_phases needs (its queue counters). */
#define ZB_REFINE_QUEUE_BYTES 1024

/* zb_refine for n_phases (1..4) phases over the same items, frame and patterns, in one launch:
 * a grid of as many blocks as stay resident pulls the n_phases x n_items refinements from
 * device-side queues (one per XCD, over contiguous runs of the phase-major, per-phase ordered
 * items), each block building the masked-pixel list once. Every record is bit-identical to
 * zb_refine's for the same phase (opts.order is ignored: each phase brings its own). work:
 * ZB_REFINE_QUEUE_BYTES of device memory for the persistent form (the knob "refine_queue"),
 * cleared by a kernel on `stream` at the start of the call (calls that overlap on one work
 * buffer race); the block-per-work-item form leaves it untouched.
 * Replaces the per-phase loop of Detector.refine dispatches (src/engine.zig:2085-2119). */
int32_t zb_refine_phases(const zb_refine_phase* phases, int32_t n_phases, const zb_frame* frame, const float* pats,
                         const float* items, int32_t n_items, const zb_refine_opts* opts, int32_t out0, void* work,
                         int64_t work_bytes, void* stream);

/* The refinement's item order for zb_refine_opts.order: the n_items items sorted by the master
 * region their detector's centre ray lands in at their seed orientation (seeds[b * seed_stride
 * ..], as zb_refine takes them; items: [n][4] (t, pattern index)), as cells of 32 x 32 a Lambert
 * sheet along a Morton curve; order: [n_items] out. One block of 1 024 threads. No counterpart in
 * the reference, whose refinement dispatches take items in batch order. */
int32_t zb_refine_order(const zb_frame* frame, const float* items, const float* seeds, int32_t seed_stride, int32_t n_items, int32_t* order,
                        void* stream);

/* Flag of zb_sums: zero the sums of the shift directions (orientation refinement's view). */
#define ZB_SUMS_ROTATION_ONLY 2

/* The 43 sums of the six-direction fit of each item's pose against its pattern: count, sum M,
 * sum M^2, sum M e, then for d = (q^-1 grad M / |u|, c x grad M) (a shift of the rays' origin
 * along the sample's x, y, z in pixels; a crystal rotation about x, y, z in radians) sum d_i,
 * sum M d_i, sum e d_i (i < 6) and the upper triangle of sum d_i d_j, row by row (21).
 * NCC, its gradient and Gauss-Newton matrix over the six directions follow from them
 * (planefit.zig:177-201 localOf; zebsdsolve_amd/fit.py). poses: [n][8] = (t, pattern index, q);
 * sums: [n][43] out.
 * Replaces Detector.sumsOf (src/detector.zig:144-159: detector.wgsl sums, sums_reduce), with
 * the bilinear master lookup. */
int32_t zb_sums(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* poses, int32_t n_items,
                int32_t flags, float* sums, void* stream);

/* zb_sums for the items still moving: an item whose state [n] (zb_fit6_step's) is not 0 is left
 * out and its sums are not written. No counterpart in the reference (whose fit re-scores every
 * pattern each step); a fit to convergence then costs about the steps its slowest items need. */
int32_t zb_sums_active(const float* mgrad, int32_t side, const zb_frame* frame, const float* pats, const float* poses, int32_t n_items,
                       int32_t flags, const int32_t* state, float* sums, void* stream);

/* The six-direction fit to convergence, one Levenberg-Marquardt step at a time, in FP64 on the
 * device (one thread an item): zb_fit6_propose turns each item's sums (zb_sums at its pose) into
 * the NCC, gradient and Gauss-Newton matrix (planefit.zig:177-201 localOf), damps the matrix by
 * lambda on its diagonal plus 1e-9 of its trace (planefit.zig:229-247 damped; the last
 * `unknowns` of the six directions: 3 the rotation alone, 6 the shift too), solves it by
 * elimination with partial pivoting (planefit.zig:203-227 solve), limits the step to max_shift
 * pixels and max_rot radians, and writes the moved poses to trial ([n][8], as poses) and
 * (NCC, step length) to cur ([n][2] f64). zb_fit6_accept, given zb_sums at the trial poses,
 * keeps a trial whose NCC rose (pose and sums copied, lambda / 10) or the old pose
 * (lambda x 10); state [n] i32: 0 fitting, 1 converged (the step was below tol), 2 flat (no
 * NCC; the pose is left alone). lambda: [n] f64, set by the caller before the first step. */
int32_t zb_fit6_propose(const float* poses, const float* sums, const double* lambda, int32_t n, int32_t unknowns, double max_rot,
                        double max_shift, float* trial, double* cur, int32_t* state, void* stream);
int32_t zb_fit6_accept(const float* trial, const float* trial_sums, int32_t n, double tol, const double* cur, float* poses, float* sums,
                       double* lambda, int32_t* state, void* stream);

/* The same fit from each item's local model instead of its sums, so that the gradient can come
 * from elsewhere (zebsdsolve_amd/fit.py drives it with the gradient of its autograd function):
 * local [n][29] f64 = (NCC, ok (1.0 or 0.0), g[6], the Gauss-Newton matrix's upper triangle row
 * by row [21]) over the six directions of zb_sums. zb_local6 writes it from the sums
 * (planefit.zig:177-201 localOf; ok 0 and the rest 0 where the master is flat under the mask);
 * zb_fit6_step proposes the damped step from it as zb_fit6_propose does from the sums
 * (planefit.zig:203-247 damped, solve); zb_fit6_keep keeps a trial whose local NCC rose (pose
 * and local copied, lambda / 10), else the old pose (lambda x 10), as zb_fit6_accept. Items
 * that have converged (state 1) stay put: zb_fit6_step leaves their trial at the pose and
 * zb_fit6_keep leaves them alone, so their trial need not be scored (zb_sums_active). */
int32_t zb_local6(const float* sums, int32_t n, double* local, void* stream);
int32_t zb_fit6_step(const float* poses, const double* local, const double* lambda, int32_t n, int32_t unknowns, double max_rot, double max_shift,
                     float* trial, double* cur, int32_t* state, void* stream);
int32_t zb_fit6_keep(const float* trial, const double* trial_local, int32_t n, double tol, const double* cur, float* poses, double* local,
                     double* lambda, int32_t* state, void* stream);

/* The pattern filters before the path. background: [h][w] f32 subtracted from every pattern
 * (the static background, e.g. the mean of patterns over the map), or NULL; dynamic: 0 none,
 * 1 subtract, 2 divide by a Gaussian blur of each pattern of deviation dynamic_sigma pixels
 * (the reference's dynamic_sigma setting times the pattern width). */
typedef struct zb_prep_opts {
    const float* background;
    int32_t dynamic;
    float dynamic_sigma;
    /* CLAHE (settings clahe, clahe_grid 1..16, clahe_bins 2..256, clahe_clip > 0): the image
     * stretched to [0, 1] and equalized over grid x grid tiles */
    int32_t clahe;
    int32_t clahe_grid;
    int32_t clahe_bins;
    float clahe_clip;
    /* band pass (settings bandpass, bp_fine >= 0, bp_coarse > bp_fine, pixels): the Gaussian of
     * bp_fine (the image itself when bp_fine <= 0.05) minus the Gaussian of bp_coarse */
    int32_t bandpass;
    float bp_fine;
    float bp_coarse;
} zb_prep_opts;

/* zb_normalize after the filters, fused: raw patterns in, the static background subtracted,
 * the dynamic background subtracted or divided by, zero mean and unit norm under the mask out.
 * The blur is the reference's: direct separable taps below 12 pixels of deviation, sums over
 * boxes from 12. One block per pattern with the image in LDS: fails (ZB_EINVAL) for patterns
 * whose image and blur buffers exceed the device's LDS per block.
 * CLAHE and the band pass follow, in the reference's order; CLAHE's histograms take
 * grid^2 bins words of the LDS, and the band pass parks its coarse Gaussian in `out`.
 * Replaces Prep.run + Prep.filter (src/prep.zig:143-204; prep.wgsl load, background, blur_rows,
 * blur_cols, box_sum, coarse_rows, coarse_cols, spread_rows, spread_cols, combine, stretch,
 * clahe_hist, clahe_cdf, clahe_apply, normalize). */
int32_t zb_prep(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* opts,
                float* out, void* stream);

/* Bytes of workspace zb_prep_work needs for n patterns h x w with these filters: 0 where one
 * block's LDS holds the image and its scratch (zb_prep's case), else a slot of image and
 * scratch for each of up to 1 024 blocks. */
int64_t zb_prep_workspace(int32_t n, int32_t h, int32_t w, const zb_prep_opts* opts);

/* zb_prep for patterns of any size (the reference's 240 x 320 frames with its default filters):
 * where LDS holds the image, zb_prep itself; past it, the same arithmetic in the same order over
 * a slot of global memory a block (work: zb_prep_workspace bytes, 8-byte aligned; a smaller one
 * takes fewer blocks, down to one slot), each block taking patterns in turn. The results are bit-identical
 * to zb_prep's where both run (the knob "prep_global" forces the global form). Replaces the
 * same Prep.run + Prep.filter as zb_prep. */
int32_t zb_prep_work(const void* raw, int32_t pixel, int32_t n, int32_t h, int32_t w, const float mask[3], const zb_prep_opts* opts,
                     void* work, int64_t work_bytes, float* out, void* stream);

/* CLAHE of n h x w float images in place, in global memory (for images larger than zb_prep's
 * LDS: a master's two sheets): each image stretched to [0, 1] by its extremes, then equalized
 * over grid x grid tiles (1..64) of bins levels (2..256) with the histograms clipped at clip x
 * the mean count, as zb_prep's CLAHE. work: zb_clahe_workspace bytes (the tables).
 * Replaces Prep.filter with only CLAHE on, as bakeMaster runs it on the sheets
 * (src/engine.zig:1741-1743; prep.wgsl stretch, clahe_hist, clahe_cdf, clahe_apply). */
int64_t zb_clahe_workspace(int32_t n, int32_t grid, int32_t bins);
int32_t zb_clahe(float* img, int32_t n, int32_t h, int32_t w, int32_t grid, int32_t bins, float clip, void* work, int64_t work_bytes,
                 void* stream);

/* The mean of a [2][side][side] master over its point group: each texel the average of the
 * sheets (bilinear) at op . d over ops [n_ops][3][3] (row-major, device memory); dst != src.
 * Replaces Prep.symmetrize (src/prep.zig:241-249; mastersph.wgsl:111-122). */
int32_t zb_master_symmetrize(const float* src, float* dst, int32_t side, const float* ops, int32_t n_ops, void* stream);

/* Simulated patterns: the master (bilinear, [2][side][side] f32) in direction q . ray of every
 * pixel. poses: [n][8] = (t, pattern index, q); out: [n][h][w].
 * Replaces Detector.project (src/detector.zig:137-142: detector.wgsl project). */
int32_t zb_project(const float* master, int32_t side, const zb_frame* frame, const float* poses, int32_t n,
                   float* out, void* stream);

/* One pass of Gaussian smoothing on the sphere of a [2][side][side] master: each texel the
 * weighted mean over a 9 x 9 lattice of directions in its tangent plane, step 0.6 sigma_rad.
 * Replaces one `sphere_blur` dispatch of Prep.sphereBlur (src/prep.zig:215-237;
 * mastersph.wgsl:87-109). */
int32_t zb_master_blur(const float* src, float* dst, int32_t side, float sigma_rad, void* stream);

/* The master's gradient on the sphere at every texel by central differences one texel apart,
 * with its value in .w: grad [2][side][side][4].
 * Replaces Prep.gradient (src/prep.zig:253-260; mastersph.wgsl:124-138). */
int32_t zb_master_gradient(const float* sheets, int32_t side, float* grad, void* stream);

/* The gradient field in half precision, each value rounded to nearest even, each texel stored
 * with the texel below it: grad_h [2][side][side][2][4] f16 (16 bytes a texel: [0] the texel,
 * [1] the next row's at the same column; the last row of a sheet repeats itself), for
 * zb_refine's Gauss-Newton passes: a bilinear lookup is two adjacent 16-byte loads. The reference
 * keeps only the float32 field (mastersph.wgsl:124-138); this copy halves the refinement's
 * vector-memory instructions and the cache lines a lookup touches. */
int32_t zb_master_pack(const float* grad, int32_t side, uint16_t* grad_h, void* stream);

/* Kernel timing (events around every launch while on). The reference times its dispatches the
 * same way (Gpu.setProfile / profileReport, src/gpu/gpu.zig:287-313); its C ABI has no entry
 * point for it. zb_profile(1) clears and starts, zb_profile(0) stops; zb_profile_report waits
 * for the recorded launches and writes one line "kernel count total_ms" per kernel. */
int32_t zb_profile(int32_t on);
int32_t zb_profile_report(char* buf, int32_t cap);
/* Only the kernels named in `names` (comma-separated, e.g. "k_refine,k_vgemm") are timed while
 * profiling is on; NULL or "" times every kernel (the default). Fewer events in the stream: the
 * other launches run as they do unprofiled. No counterpart in the reference. */
int32_t zb_profile_select(const char* names);

/* Launch-shape knobs, process-wide, read at each launch; none changes any result bit, only how
 * the work is cut into blocks (the choices the host's tuner makes by timing). zb_knob_set
 * returns the previous value (ZB_EINVAL for an unknown name or a value out of range),
 * zb_knob_get the current one.
 *   "refine_queue" 0..2 (0): zb_refine_phases as one block per work item (0) or as a persistent
 *                            grid draining device-side queues (1; 2: with the NMI's second pass
 *                            looking the master up again rather than keeping its values, fewer
 *                            registers)
 *   "rings_group"  0..64 (0): ring pairs per block of the sphere stage's ring transforms, a power
 *                            of two (0: 16 KB of LDS worth; at most n / 2 and 64 KB)
 *   "prep_global"  0..1 (0): zb_prep_work's filters through global memory even where LDS holds
 *                            the image (1)
 * The reference fixes the ring lines a workgroup transforms at TW_LEN / n
 * (src/spherical.zig:305, 376) and has no counterpart of these. */
int32_t zb_knob_set(const char* name, int32_t value);
int32_t zb_knob_get(const char* name);

#ifdef __cplusplus
}
#endif

#endif
