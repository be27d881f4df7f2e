// The six-direction fit of each item's pose to convergence: Levenberg-Marquardt steps in FP64 on
// the device, one thread an item, from the 43 sums zb_sums gives (the NCC, its gradient and
// Gauss-Newton matrix over a shift of the rays' origin and a crystal rotation:
// planefit.zig:177-201 localOf), damped as planefit.zig:229-247 damped (Marquardt's lambda on
// the diagonal, 1e-9 of the trace), solved by elimination with partial pivoting
// (planefit.zig:203-227 solve). A step proposes the moved poses; zb_sums scores them; the
// acceptance keeps a pose whose NCC rose (lambda / 10) and otherwise keeps the old one
// (lambda x 10).
#include "common.h"

namespace zb {

constexpr int kSums = 43;

// NCC, gradient g[6] and matrix H[36] from the sums; false where the master is flat under the
// mask (planefit.zig:177-201)
__device__ bool local6(const float* s, double& ncc, double* g, double* H) {
    const double n = fmax((double)s[0], 1.0);
    const double mean = (double)s[1] / n;
    const double var_n = (double)s[2] - n * mean * mean;
    if (!(var_n > 1e-12)) return false;
    const double inv = 1.0 / sqrt(var_n);
    ncc = (double)s[3] * inv;
    double sd[6], au[6];
    for (int i = 0; i < 6; ++i) {
        sd[i] = (double)s[4 + i];
        au[i] = ((double)s[10 + i] - mean * sd[i]) * inv;
        g[i] = ((double)s[16 + i] - ncc * au[i]) * inv;
    }
    int o = 22;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            const double v = ((double)s[o] - sd[i] * sd[j] / n - au[i] * au[j]) * inv * inv;
            H[i * 6 + j] = v;
            H[j * 6 + i] = v;
            ++o;
        }
    return true;
}

// x = A^-1 b for the last k of the six unknowns, A damped by lambda; false on a zero pivot
__device__ bool damped_solve(const double* H, const double* g, int k, double lambda, double* x) {
    const int o = 6 - k;
    double A[36], b[6];
    double tr = 0.0;
    for (int i = 0; i < k; ++i) {
        b[i] = g[o + i];
        for (int j = 0; j < k; ++j) A[i * 6 + j] = H[(o + i) * 6 + o + j];
        tr += A[i * 6 + i];
    }
    for (int i = 0; i < k; ++i) A[i * 6 + i] = A[i * 6 + i] * (1.0 + lambda) + 1e-9 * tr + 1e-30;
    for (int c = 0; c < k; ++c) {
        int piv = c;
        for (int r = c + 1; r < k; ++r)
            if (fabs(A[r * 6 + c]) > fabs(A[piv * 6 + c])) piv = r;
        if (fabs(A[piv * 6 + c]) < 1e-300) return false;
        if (piv != c) {
            for (int q = 0; q < k; ++q) {
                const double t = A[c * 6 + q];
                A[c * 6 + q] = A[piv * 6 + q];
                A[piv * 6 + q] = t;
            }
            const double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
        }
        for (int r = c + 1; r < k; ++r) {
            const double f = A[r * 6 + c] / A[c * 6 + c];
            if (f == 0.0) continue;
            for (int q = c; q < k; ++q) A[r * 6 + q] -= f * A[c * 6 + q];
            b[r] -= f * b[c];
        }
    }
    for (int i = 0; i < 6; ++i) x[i] = 0.0;
    for (int r = k - 1; r >= 0; --r) {
        double acc = b[r];
        for (int q = r + 1; q < k; ++q) acc -= A[r * 6 + q] * x[o + q];
        x[o + r] = acc / A[r * 6 + r];
    }
    return true;
}

struct Fit6Arg {
    int n, unknowns;
    double max_rot, max_shift, tol;
};

// state: 0 fitting, 1 converged (its last step below tol), 2 flat (no NCC to fit)
__global__ void __launch_bounds__(256) k_fit6_propose(const float* __restrict__ poses, const float* __restrict__ sums,
                                                      const double* __restrict__ lambda, Fit6Arg a, float* __restrict__ trial,
                                                      double* __restrict__ cur, int* __restrict__ state) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float* p = poses + (size_t)i * 8;
    float* t = trial + (size_t)i * 8;
    double ncc, g[6], H[36], x[6];
    const bool ok = local6(sums + (size_t)i * kSums, ncc, g, H);
    if (!ok || state[i] == 2) {
        state[i] = 2;
        for (int k = 0; k < 8; ++k) t[k] = p[k];
        cur[2 * i] = 0.0;
        cur[2 * i + 1] = 0.0;
        return;
    }
    // the Newton step on the model NCC + g . x - x . H x / 2
    if (!damped_solve(H, g, a.unknowns, lambda[i], x))
        for (int k = 0; k < 6; ++k) x[k] = 0.0;
    // at most max_shift pixels and max_rot radians a step
    const double sh = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double th = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
    const double ks = sh > a.max_shift ? a.max_shift / sh : 1.0;
    const double kr = th > a.max_rot ? a.max_rot / th : 1.0;
    for (int k = 0; k < 3; ++k) {
        x[k] *= ks;
        x[3 + k] *= kr;
    }
    // the pose moved: t + shift; exp(rotation) q, unit with w >= 0 (src/rot.zig:60-64, 192-198)
    const double th2 = x[3] * x[3] + x[4] * x[4] + x[5] * x[5];
    const double h = sqrt(th2);
    const double s = h < 1e-6 ? 0.5 - th2 / 48.0 : sin(0.5 * h) / h;
    const double ew = cos(0.5 * h), ex = s * x[3], ey = s * x[4], ez = s * x[5];
    const double qw = p[4], qx = p[5], qy = p[6], qz = p[7];
    double nw = ew * qw - ex * qx - ey * qy - ez * qz;
    double nx = ew * qx + ex * qw + ey * qz - ez * qy;
    double ny = ew * qy - ex * qz + ey * qw + ez * qx;
    double nz = ew * qz + ex * qy - ey * qx + ez * qw;
    const double nn = sqrt(nw * nw + nx * nx + ny * ny + nz * nz);
    const double sg = (nw < 0.0 ? -1.0 : 1.0) / (nn > 0.0 ? nn : 1.0);
    t[0] = (float)((double)p[0] + x[0]);
    t[1] = (float)((double)p[1] + x[1]);
    t[2] = (float)((double)p[2] + x[2]);
    t[3] = p[3];
    t[4] = (float)(nw * sg);
    t[5] = (float)(nx * sg);
    t[6] = (float)(ny * sg);
    t[7] = (float)(nz * sg);
    cur[2 * i] = ncc;
    cur[2 * i + 1] = sqrt(sh * ks * sh * ks + th2);
}

__global__ void __launch_bounds__(256) k_fit6_accept(const float* __restrict__ trial, const float* __restrict__ tsums, Fit6Arg a,
                                                     const double* __restrict__ cur, float* __restrict__ poses, float* __restrict__ sums,
                                                     double* __restrict__ lambda, int* __restrict__ state) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n || state[i] == 2) return;
    double ncc, g[6], H[36];
    const float* ts = tsums + (size_t)i * kSums;
    if (local6(ts, ncc, g, H) && ncc > cur[2 * i]) {
        for (int k = 0; k < 8; ++k) poses[(size_t)i * 8 + k] = trial[(size_t)i * 8 + k];
        for (int k = 0; k < kSums; ++k) sums[(size_t)i * kSums + k] = ts[k];
        lambda[i] = fmax(lambda[i] * 0.1, 1e-9);
    } else {
        lambda[i] = fmin(lambda[i] * 10.0, 1e9);
    }
    state[i] = cur[2 * i + 1] < a.tol ? 1 : 0;
}

// The same fit driven from each item's local model instead of its sums: local [n][kLocal] =
// (NCC, ok (1 or 0), g[6], H's upper triangle row by row [21]), as zb_local6 writes it from the
// sums (or a caller assembles it, e.g. with the gradient from an autograd pass over zb_local6).
constexpr int kLocal = 29;

__device__ __forceinline__ bool unpack_local(const double* l, double& ncc, double* g, double* H) {
    ncc = l[0];
    for (int k = 0; k < 6; ++k) g[k] = l[2 + k];
    int o = 8;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            H[i * 6 + j] = l[o];
            H[j * 6 + i] = l[o];
            ++o;
        }
    return l[1] > 0.5;
}

__global__ void __launch_bounds__(256) k_local6(const float* __restrict__ sums, int n, double* __restrict__ local) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double ncc = 0.0, g[6] = {0, 0, 0, 0, 0, 0}, H[36];
    for (int k = 0; k < 36; ++k) H[k] = 0.0;
    const bool ok = local6(sums + (size_t)i * kSums, ncc, g, H);
    double* l = local + (size_t)i * kLocal;
    l[0] = ok ? ncc : 0.0;
    l[1] = ok ? 1.0 : 0.0;
    for (int k = 0; k < 6; ++k) l[2 + k] = ok ? g[k] : 0.0;
    int o = 8;
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) l[o++] = ok ? H[a * 6 + b] : 0.0;
}

// the damped step of one item from its model, limited and applied: trial pose, cur = (NCC, step)
__device__ void fit6_move(const float* p, double ncc, const double* g, const double* H, double lambda, const Fit6Arg& a, float* t, double* cur) {
    double x[6];
    if (!damped_solve(H, g, a.unknowns, lambda, x))
        for (int k = 0; k < 6; ++k) x[k] = 0.0;
    const double sh = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double th = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
    const double ks = sh > a.max_shift ? a.max_shift / sh : 1.0;
    const double kr = th > a.max_rot ? a.max_rot / th : 1.0;
    for (int k = 0; k < 3; ++k) {
        x[k] *= ks;
        x[3 + k] *= kr;
    }
    const double th2 = x[3] * x[3] + x[4] * x[4] + x[5] * x[5];
    const double h = sqrt(th2);
    const double s = h < 1e-6 ? 0.5 - th2 / 48.0 : sin(0.5 * h) / h;
    const double ew = cos(0.5 * h), ex = s * x[3], ey = s * x[4], ez = s * x[5];
    const double qw = p[4], qx = p[5], qy = p[6], qz = p[7];
    double nw = ew * qw - ex * qx - ey * qy - ez * qz;
    double nx = ew * qx + ex * qw + ey * qz - ez * qy;
    double ny = ew * qy - ex * qz + ey * qw + ez * qx;
    double nz = ew * qz + ex * qy - ey * qx + ez * qw;
    const double nn = sqrt(nw * nw + nx * nx + ny * ny + nz * nz);
    const double sg = (nw < 0.0 ? -1.0 : 1.0) / (nn > 0.0 ? nn : 1.0);
    t[0] = (float)((double)p[0] + x[0]);
    t[1] = (float)((double)p[1] + x[1]);
    t[2] = (float)((double)p[2] + x[2]);
    t[3] = p[3];
    t[4] = (float)(nw * sg);
    t[5] = (float)(nx * sg);
    t[6] = (float)(ny * sg);
    t[7] = (float)(nz * sg);
    cur[0] = ncc;
    cur[1] = sqrt(sh * ks * sh * ks + th2);
}

__global__ void __launch_bounds__(256) k_fit6_step(const float* __restrict__ poses, const double* __restrict__ local,
                                                   const double* __restrict__ lambda, Fit6Arg a, float* __restrict__ trial,
                                                   double* __restrict__ cur, int* __restrict__ state) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float* p = poses + (size_t)i * 8;
    float* t = trial + (size_t)i * 8;
    double ncc, g[6], H[36];
    if (state[i] == 1) {
        // converged: it stays where it is (its trial is not scored)
        for (int k = 0; k < 8; ++k) t[k] = p[k];
        return;
    }
    if (!unpack_local(local + (size_t)i * kLocal, ncc, g, H) || state[i] == 2) {
        state[i] = 2;
        for (int k = 0; k < 8; ++k) t[k] = p[k];
        cur[2 * i] = 0.0;
        cur[2 * i + 1] = 0.0;
        return;
    }
    fit6_move(p, ncc, g, H, lambda[i], a, t, cur + 2 * i);
}

__global__ void __launch_bounds__(256) k_fit6_keep(const float* __restrict__ trial, const double* __restrict__ tlocal, Fit6Arg a,
                                                   const double* __restrict__ cur, float* __restrict__ poses, double* __restrict__ local,
                                                   double* __restrict__ lambda, int* __restrict__ state) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n || state[i] != 0) return;  // flat, or converged before this step
    const double* tl = tlocal + (size_t)i * kLocal;
    if (tl[1] > 0.5 && tl[0] > cur[2 * i]) {
        for (int k = 0; k < 8; ++k) poses[(size_t)i * 8 + k] = trial[(size_t)i * 8 + k];
        for (int k = 0; k < kLocal; ++k) local[(size_t)i * kLocal + k] = tl[k];
        lambda[i] = fmax(lambda[i] * 0.1, 1e-9);
    } else {
        lambda[i] = fmin(lambda[i] * 10.0, 1e9);
    }
    state[i] = cur[2 * i + 1] < a.tol ? 1 : 0;
}

}  // namespace zb

using namespace zb;

static int fit6_arg(int32_t n, int32_t unknowns, double max_rot, double max_shift, double tol, Fit6Arg& a, const char* who) {
    if (n < 0 || !(unknowns == 3 || unknowns == 6) || !(max_rot > 0.0) || !(max_shift > 0.0) || !(tol >= 0.0)) {
        set_error("%s: bad arguments (n %d, unknowns %d)", who, n, unknowns);
        return ZB_EINVAL;
    }
    a.n = n;
    a.unknowns = unknowns;
    a.max_rot = max_rot;
    a.max_shift = max_shift;
    a.tol = tol;
    return ZB_OK;
}

extern "C" int32_t zb_fit6_propose(const float* poses, const float* sums, const double* lambda, int32_t n, int32_t unknowns, double max_rot,
                                   double max_shift, float* trial, double* cur, int32_t* state, void* stream) {
    Fit6Arg a;
    if (int rc = fit6_arg(n, unknowns, max_rot, max_shift, 0.0, a, "zb_fit6_propose")) return rc;
    if (n == 0) return ZB_OK;
    if (!poses || !sums || !lambda || !trial || !cur || !state) {
        set_error("zb_fit6_propose: null pointer");
        return ZB_EINVAL;
    }
    ZB_LAUNCH("k_fit6_propose", k_fit6_propose, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, poses, sums, lambda, a, trial, cur,
              (int*)state);
    return check_launch("zb_fit6_propose");
}

extern "C" int32_t zb_fit6_accept(const float* trial, const float* trial_sums, int32_t n, double tol, const double* cur, float* poses, float* sums,
                                  double* lambda, int32_t* state, void* stream) {
    Fit6Arg a;
    if (int rc = fit6_arg(n, 6, 1.0, 1.0, tol, a, "zb_fit6_accept")) return rc;
    if (n == 0) return ZB_OK;
    if (!trial || !trial_sums || !cur || !poses || !sums || !lambda || !state) {
        set_error("zb_fit6_accept: null pointer");
        return ZB_EINVAL;
    }
    ZB_LAUNCH("k_fit6_accept", k_fit6_accept, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, trial, trial_sums, a, cur, poses, sums,
              lambda, (int*)state);
    return check_launch("zb_fit6_accept");
}

extern "C" int32_t zb_local6(const float* sums, int32_t n, double* local, void* stream) {
    if (n < 0 || (n > 0 && (!sums || !local))) {
        set_error("zb_local6: bad arguments");
        return ZB_EINVAL;
    }
    if (n == 0) return ZB_OK;
    ZB_LAUNCH("k_local6", k_local6, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, sums, n, local);
    return check_launch("zb_local6");
}

extern "C" int32_t zb_fit6_step(const float* poses, const double* local, const double* lambda, int32_t n, int32_t unknowns, double max_rot,
                                double max_shift, float* trial, double* cur, int32_t* state, void* stream) {
    Fit6Arg a;
    if (int rc = fit6_arg(n, unknowns, max_rot, max_shift, 0.0, a, "zb_fit6_step")) return rc;
    if (n == 0) return ZB_OK;
    if (!poses || !local || !lambda || !trial || !cur || !state) {
        set_error("zb_fit6_step: null pointer");
        return ZB_EINVAL;
    }
    ZB_LAUNCH("k_fit6_step", k_fit6_step, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, poses, local, lambda, a, trial, cur,
              (int*)state);
    return check_launch("zb_fit6_step");
}

extern "C" int32_t zb_fit6_keep(const float* trial, const double* trial_local, int32_t n, double tol, const double* cur, float* poses,
                                double* local, double* lambda, int32_t* state, void* stream) {
    Fit6Arg a;
    if (int rc = fit6_arg(n, 6, 1.0, 1.0, tol, a, "zb_fit6_keep")) return rc;
    if (n == 0) return ZB_OK;
    if (!trial || !trial_local || !cur || !poses || !local || !lambda || !state) {
        set_error("zb_fit6_keep: null pointer");
        return ZB_EINVAL;
    }
    ZB_LAUNCH("k_fit6_keep", k_fit6_keep, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, trial, trial_local, a, cur, poses, local,
              lambda, (int*)state);
    return check_launch("zb_fit6_keep");
}
