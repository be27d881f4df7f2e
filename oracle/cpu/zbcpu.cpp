// zbcpu.cpp — the indexing path as a float32 C++17 / OpenMP program on the host cores: the CPU
// restatement of the reference engine that the throughput is compared with (not the reference
// binary, which needs Zig, wgpu-native and a Vulkan device, none of them here).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's CPU baseline, never by the
// product (zebsdsolve_amd/), which runs on the GPU alone.
//
// Per pattern, in the reference's order and with its algorithm (one OpenMP thread a pattern):
//   prep.wgsl normalize            mean and norm under the circular mask
//   sphere.wgsl sample, sphere_sums, window_mean (67-151)
//                                   ray-plane intersection and bilinear lookup at every point
//                                   of the Gauss-Legendre grid, the window's quadrature mean
//   sphere.wgsl rings (153-207)     ring y and its mirror as one complex line, FFT along phi,
//                                   split into the two rings' orders, kept as (N + S, N - S)
//   sphere.wgsl legendre, sh_norm   a_lm = sum_y wplm[y][p] (N + S or N - S by the parity of
//                                   l + m), unit harmonic norm
//   so3.wgsl spectrum (140-277)     T[k, m', m] = i^{m'} sum_l W[mu][l][k] conj(g_{l m'})
//                                   d^l_{k m'}(pi/2) for k >= 0, T[-k] = (-1)^{m'+m} T[k], the
//                                   FFT along k (b) keeping j <= n / 2
//   so3.wgsl a_peak_m (396-593)     the FFT along the master's orders (c), the real transform
//                                   along m' (a), times the reciprocal denominator, the argmax
//                                   (ties: the lowest index (j nc + c) n + a)
//   so3.wgsl peak_finish (665-796)  parabolic offsets from the six neighbours, clamped to half a
//                                   step (b only inside 0 < j < n / 2), quaternion conj(S)
//   detector.wgsl refine_* (318-659) Gauss-Newton on the orientation: the 19 sums over the
//                                   masked pixels from the float32 gradient field, the curvature
//                                   rescaled by gain against promise, steps clamped to max_step
//   detector.wgsl nmi (739-832)     normalized mutual information with the master, 16 levels
// and the window's denominator volume 1 / sqrt(<w,f^2> - <w,f>^2 / area) (so3.wgsl:595-626)
// from the same transforms. FFTs are iterative radix-2 (decimation in time, bit-reversed load).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<float> cf;

extern "C" {

struct zbc_band {
    int32_t L, n;
    const float* ring;  // [n][2] cos theta, weight (Gauss-Legendre, theta ascending)
    const float* wplm;  // [n/2][pairs] w_y P_l^m(x_y) 2 pi / (2n)
    const float* dq;    // [m][l][k] d^l_{k,m}(pi/2)
};

struct zbc_phase {
    int32_t fold, nc, mu_max, l_step;
    const float* w;  // [2 mu_max + 1][L][L][2] i^{-m} f_lm d^l_{km}(pi/2), k innermost
};

struct zbc_frame {
    float c0[3], c1[3], c2[3];
    int32_t h, w;
    float mask[3];  // centre row, centre column, squared radius (<= 0: none)
};

struct zbc_master {
    int32_t side;
    const float* sheets;  // [2][side][side]
    const float* grad;    // [2][side][side][4] (gradient, value)
};

struct zbc_opts {
    int32_t iters;
    float max_step, scale0, scale_min, scale_max;
    int32_t nmi_bins;
};

}  // extern "C"

namespace {

constexpr float kPi = 3.14159265358979f;
constexpr float kTau = 6.28318530717959f;

inline int pair_of(int L, int m, int l) { return m * L - (m * (m - 1)) / 2 + (l - m); }

// ── FFT ──

struct Fft {
    int n = 0, lg = 0;
    std::vector<cf> tw;    // exp(-2 pi i t / n), t < n / 2
    std::vector<int> rev;  // bit reversal
    explicit Fft(int n_) : n(n_) {
        while ((1 << lg) < n) ++lg;
        tw.resize(n / 2);
        for (int t = 0; t < n / 2; ++t) tw[t] = cf(std::cos(kTau * t / n), -std::sin(kTau * t / n));
        rev.resize(n);
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
            rev[i] = r;
        }
    }
    // forward transform (e^{-i}) of x in place
    void run(cf* x) const {
        for (int i = 0; i < n; ++i)
            if (i < rev[i]) std::swap(x[i], x[rev[i]]);
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < half; ++k) {
                    const cf t = tw[k * step] * x[i + k + half];
                    x[i + k + half] = x[i + k] - t;
                    x[i + k] += t;
                }
        }
    }
};

// ── the master on the sphere (common.wgsl sphere_to_square; detector.wgsl master_at) ──

inline void sphere_to_square(float vx, float vy, float vz, float& u, float& w) {
    const float ax = std::fabs(vx), ay = std::fabs(vy);
    const float xn = std::sqrt(std::max(kPi * 0.5f * (ax * ax + ay * ay) / (1.f + std::fabs(vz)), 0.f));
    float a, b;
    if (ax >= ay) {
        a = xn;
        b = (4.f * xn / kPi) * std::atan2(ay, std::max(ax, 1e-30f));
    } else {
        a = (4.f * xn / kPi) * std::atan2(ax, std::max(ay, 1e-30f));
        b = xn;
    }
    const float sc = 1.f / std::sqrt(kPi / 2.f);
    u = a * sc * (vx < 0.f ? -1.f : 1.f);
    w = b * sc * (vy < 0.f ? -1.f : 1.f);
}

// bilinear lookup of `ch` channels of [north | south] sheets in direction v
inline void sheet_at(const float* S, int side, int ch, float vx, float vy, float vz, float* out) {
    float u, w;
    sphere_to_square(vx, vy, vz, u, w);
    const float last = (float)(side - 1);
    const float fj = std::min(std::max((u + 1.f) * 0.5f * last, 0.f), last);
    const float fi = std::min(std::max((w + 1.f) * 0.5f * last, 0.f), last);
    const int c0 = std::min((int)fj, side - 2), r0 = std::min((int)fi, side - 2);
    const size_t o = (size_t)r0 * side + c0 + (vz < 0.f ? (size_t)side * side : 0);
    const float tj = fj - c0, ti = fi - r0;
    for (int c = 0; c < ch; ++c) {
        const float a = S[o * ch + c], b = S[(o + 1) * ch + c], d = S[(o + side) * ch + c], e = S[(o + side + 1) * ch + c];
        const float top = a * (1.f - tj) + b * tj, bot = d * (1.f - tj) + e * tj;
        out[c] = top * (1.f - ti) + bot * ti;
    }
}

inline void qmul(const float* a, const float* b, float* o) {
    const float w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    const float x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    const float y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    const float z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w, o[1] = x, o[2] = y, o[3] = z;
}

inline void qrot(const float* q, const float* v, float* o) {
    const float ux = q[1], uy = q[2], uz = q[3];
    const float tx = 2.f * (uy * v[2] - uz * v[1]), ty = 2.f * (uz * v[0] - ux * v[2]), tz = 2.f * (ux * v[1] - uy * v[0]);
    o[0] = v[0] + q[0] * tx + (uy * tz - uz * ty);
    o[1] = v[1] + q[0] * ty + (uz * tx - ux * tz);
    o[2] = v[2] + q[0] * tz + (ux * ty - uy * tx);
}

inline bool in_mask(const zbc_frame& f, float r, float c) {
    if (f.mask[2] <= 0.f) return true;
    const float dr = r - f.mask[0], dc = c - f.mask[1];
    return dr * dr + dc * dc <= f.mask[2];
}

// ── per-thread work buffers ──

struct Work {
    int L = 0, n = 0, nc = 0, pairs = 0, nb = 0;
    Fft fphi{16}, fn{16}, fc{16};
    std::vector<cf> line, spec, alm, T, U, V, z;
    std::vector<float> vol, val, win, dirs;
    void size(const zbc_band& b, const zbc_phase& p) {
        if (L == b.L && n == b.n && nc == p.nc) return;
        L = b.L, n = b.n, nc = p.nc, pairs = L * (L + 1) / 2, nb = n / 2 + 1;
        fphi = Fft(2 * n), fn = Fft(n), fc = Fft(nc);
        line.assign(2 * n, cf());
        spec.assign((size_t)(n / 2) * L * 2, cf());
        alm.assign(pairs, cf());
        T.assign(n, cf());
        U.assign((size_t)nb * L * nc, cf());
        V.assign(nc, cf());
        z.assign(n, cf());
        vol.assign((size_t)nb * nc * n, 0.f);
        val.assign((size_t)2 * n * n, 0.f);
        win.assign((size_t)2 * n * n, 0.f);
        // the grid's directions (sin theta cos phi, sin theta sin phi, cos theta)
        dirs.resize((size_t)2 * n * n * 3);
        for (int y = 0; y < n; ++y) {
            const float x = b.ring[2 * y], st = std::sqrt(std::max(1.f - x * x, 0.f));
            for (int j = 0; j < 2 * n; ++j) {
                const float ph = kTau * j / (2 * n);
                float* d = &dirs[((size_t)y * 2 * n + j) * 3];
                d[0] = st * std::cos(ph), d[1] = st * std::sin(ph), d[2] = x;
            }
        }
    }
};

// ── the sphere stage: coefficients of one item (pattern or window) ──

// returns the window's area; alm normalized when `norm`
float coefficients(const zbc_band& b, const zbc_frame& f, const float* pat, const float* t, bool window, bool center, bool norm, Work& W,
                   float* stats) {
    const int n = b.n, nphi = 2 * n, L = b.L;
    const float ct = f.c2[0] * t[0] + f.c2[1] * t[1] + f.c2[2] * t[2];
    double sv = 0, sw = 0;
    for (int y = 0; y < n; ++y) {
        const float wy = b.ring[2 * y + 1];
        for (int j = 0; j < nphi; ++j) {
            const float* d = &W.dirs[((size_t)y * nphi + j) * 3];
            const float nd = d[0] * f.c2[0] + d[1] * f.c2[1] + d[2] * f.c2[2];
            float v = 0.f, in = 0.f;
            if (nd > 1e-4f) {
                const float s = ct / nd;
                const float hv[3] = {s * d[0] - t[0], s * d[1] - t[1], s * d[2] - t[2]};
                const float row = hv[0] * f.c0[0] + hv[1] * f.c0[1] + hv[2] * f.c0[2];
                const float col = hv[0] * f.c1[0] + hv[1] * f.c1[1] + hv[2] * f.c1[2];
                if (row >= 0.f && row <= f.h - 1 && col >= 0.f && col <= f.w - 1 && in_mask(f, row, col)) {
                    in = 1.f;
                    if (!window) {
                        const int r0 = std::min((int)row, f.h - 2), q0 = std::min((int)col, f.w - 2);
                        const float fr = row - r0, fq = col - q0;
                        const float* P = pat + (size_t)r0 * f.w + q0;
                        const float top = P[0] * (1.f - fq) + P[1] * fq, bot = P[f.w] * (1.f - fq) + P[f.w + 1] * fq;
                        v = top * (1.f - fr) + bot * fr;
                    } else {
                        v = 1.f;
                    }
                }
            }
            W.val[(size_t)y * nphi + j] = v;
            W.win[(size_t)y * nphi + j] = in;
            sv += (double)v * in * wy;
            sw += (double)in * wy;
        }
    }
    const float mean = center ? (float)(sv / std::max(sw, 1e-20)) : 0.f;
    // rings: y and its mirror n - 1 - y as one line N + i S
    const int nh = n / 2;
    for (int y = 0; y < nh; ++y) {
        const int ym = n - 1 - y;
        for (int j = 0; j < nphi; ++j) {
            const float a = (W.val[(size_t)y * nphi + j] - mean) * W.win[(size_t)y * nphi + j];
            const float c = (W.val[(size_t)ym * nphi + j] - mean) * W.win[(size_t)ym * nphi + j];
            W.line[j] = cf(a, c);
        }
        W.fphi.run(W.line.data());
        for (int m = 0; m < L; ++m) {
            const cf zp = W.line[m], zm = std::conj(W.line[(nphi - m) % nphi]);
            const cf N = 0.5f * (zp + zm), S = cf(0.f, -0.5f) * (zp - zm);
            W.spec[((size_t)y * L + m) * 2] = N + S;
            W.spec[((size_t)y * L + m) * 2 + 1] = N - S;
        }
    }
    // Legendre sums
    double nrm2 = 0;
    for (int m = 0; m < L; ++m)
        for (int l = m; l < L; ++l) {
            const int p = pair_of(L, m, l);
            const int par = (l + m) & 1;
            cf acc(0.f, 0.f);
            for (int y = 0; y < nh; ++y) acc += b.wplm[(size_t)y * W.pairs + p] * W.spec[((size_t)y * L + m) * 2 + par];
            W.alm[p] = acc;
            nrm2 += (m == 0 ? 1.0 : 2.0) * std::norm(acc);
        }
    const float nrm = (float)std::sqrt(std::max(nrm2, 1e-30));
    if (norm)
        for (auto& a : W.alm) a /= nrm;
    const float area = (float)(sw * kTau / nphi);
    if (stats) stats[0] = area, stats[1] = mean, stats[2] = nrm, stats[3] = W.alm[0].real();
    return area;
}

// ── the correlation: spectrum, transforms along b, c and a; the volume into W.vol ──

void volume(const zbc_band& b, const zbc_phase& ph, const cf* g, Work& W) {
    const int L = b.L, n = b.n, nc = ph.nc, nb = n / 2 + 1, mu_max = ph.mu_max, fold = ph.fold;
    const cf* Wt = reinterpret_cast<const cf*>(ph.w);
    std::fill(W.U.begin(), W.U.end(), cf());
    std::vector<cf>& T = W.T;
    // spectrum and its transform along k (b), per (m', mu)
    for (int mp = 0; mp < L; ++mp) {
        const float* dqm = b.dq + (size_t)mp * L * L;
        for (int mu = -mu_max; mu <= mu_max; ++mu) {
            const int am = std::abs(mu) * fold;
            const int lo = std::max(mp, am);
            if (lo >= L) continue;
            std::fill(T.begin(), T.end(), cf());
            float* Tf = reinterpret_cast<float*>(T.data());
            const float* wrow = reinterpret_cast<const float*>(Wt + (size_t)(mu + mu_max) * L * L);
            // the master has no odd l when l_step is 2: only its degrees are summed
            int l0 = lo;
            if (ph.l_step == 2 && (l0 & 1)) ++l0;
            for (int l = l0; l < L; l += ph.l_step) {
                const cf gl = g[pair_of(L, mp, l)];
                const float gr = gl.real(), gi = -gl.imag();  // conj(g)
                const float* wl = wrow + (size_t)l * L * 2;
                const float* dl = dqm + (size_t)l * L;
                for (int k = 0; k <= l; ++k) {
                    const float cr = gr * dl[k], ci = gi * dl[k];
                    Tf[2 * k] += wl[2 * k] * cr - wl[2 * k + 1] * ci;
                    Tf[2 * k + 1] += wl[2 * k] * ci + wl[2 * k + 1] * cr;
                }
            }
            // i^{m'}, then T[-k] = (-1)^{m'+m} T[k] at n - k
            static const cf ip[4] = {cf(1, 0), cf(0, 1), cf(-1, 0), cf(0, -1)};
            const cf s = ip[mp & 3];
            const bool odd = ((mp + am) & 1) != 0;
            for (int k = 0; k < L; ++k) T[k] *= s;
            for (int k = 1; k < L; ++k) T[n - k] = odd ? -T[k] : T[k];
            for (int k = L; k <= n - L; ++k) T[k] = cf();
            W.fn.run(T.data());
            const int iq = ((mu % nc) + nc) % nc;
            for (int j = 0; j < nb; ++j) W.U[((size_t)j * L + mp) * nc + iq] = T[j];
        }
    }
    // along c, then along a (real: C(a) = Re V[0] + 2 Re sum_{m'>0} V[m'] e^{-i m' a})
    std::vector<cf> Vj((size_t)L * nc);
    for (int j = 0; j < nb; ++j) {
        for (int mp = 0; mp < L; ++mp) {
            cf* u = &W.U[((size_t)j * L + mp) * nc];
            W.fc.run(u);
        }
        for (int c = 0; c < nc; ++c) {
            for (int mp = 0; mp < L; ++mp) W.z[mp] = (mp == 0 ? 1.f : 2.f) * W.U[((size_t)j * L + mp) * nc + c];
            W.z[0] = cf(W.z[0].real(), 0.f);
            for (int k = L; k < n; ++k) W.z[k] = cf();
            W.fn.run(W.z.data());
            float* out = &W.vol[((size_t)j * nc + c) * n];
            for (int a = 0; a < n; ++a) out[a] = W.z[a].real();
        }
    }
}

inline float vertex(float lo, float mid, float hi) {
    const float den = lo - 2.f * mid + hi;
    if (den >= -1e-12f) return 0.f;
    return std::min(std::max(0.5f * (lo - hi) / den, -0.5f), 0.5f);
}

// argmax of vol x rden and the sub-cell peak: q (w, x, y, z), score
void peak(const zbc_band& b, const zbc_phase& ph, const float* rden, Work& W, float* q, float* score) {
    const int n = b.n, nc = ph.nc, nb = n / 2 + 1;
    const size_t N = (size_t)nb * nc * n;
    for (size_t i = 0; i < N; ++i) W.vol[i] *= rden[i];
    size_t best = 0;
    for (size_t i = 1; i < N; ++i)
        if (W.vol[i] > W.vol[best]) best = i;
    const float v0 = W.vol[best];
    const int j = (int)(best / ((size_t)nc * n)), c = (int)((best / n) % nc), a = (int)(best % n);
    auto S = [&](int jj, int cc, int aa) { return W.vol[((size_t)jj * nc + cc) * n + aa]; };
    const float da = vertex(S(j, c, (a + n - 1) % n), v0, S(j, c, (a + 1) % n));
    const float dc = vertex(S(j, (c + nc - 1) % nc, a), v0, S(j, (c + 1) % nc, a));
    const float dj = (j > 0 && j < n / 2) ? vertex(S(j - 1, c, a), v0, S(j + 1, c, a)) : 0.f;
    const float step = kTau / n;
    const float A = (a + da) * step, B = (j + dj) * step, C = (c + dc) * step;
    const float qa[4] = {std::cos(0.5f * A), 0.f, 0.f, std::sin(0.5f * A)};
    const float qb[4] = {std::cos(0.5f * B), 0.f, std::sin(0.5f * B), 0.f};
    const float qc[4] = {std::cos(0.5f * C), 0.f, 0.f, std::sin(0.5f * C)};
    float t[4], r[4];
    qmul(qb, qc, t);
    qmul(qa, t, r);
    float sg = r[0] < 0.f ? -1.f : 1.f;  // conj, then w >= 0
    q[0] = sg * r[0], q[1] = -sg * r[1], q[2] = -sg * r[2], q[3] = -sg * r[3];
    *score = v0;
}

// ── refinement and NMI on the detector ──

struct Pix {
    std::vector<float> u;  // unit rays [k][3]
    std::vector<float> e;  // pattern values
};

void pixels(const zbc_frame& f, const float* pat, const float* t, Pix& P) {
    P.u.clear();
    P.e.clear();
    for (int r = 0; r < f.h; ++r)
        for (int c = 0; c < f.w; ++c) {
            if (!in_mask(f, (float)r, (float)c)) continue;
            float u[3];
            for (int i = 0; i < 3; ++i) u[i] = f.c0[i] * r + f.c1[i] * c + t[i];
            const float inv = 1.f / std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            P.u.push_back(u[0] * inv), P.u.push_back(u[1] * inv), P.u.push_back(u[2] * inv);
            P.e.push_back(pat[(size_t)r * f.w + c]);
        }
}

// NCC, gradient and Gauss-Newton matrix over the rotation from the 19 sums at q
float ncc_at(const zbc_master& M, const Pix& P, const float* q, float* grad, float* hess) {
    double n = 0, sm = 0, smm = 0, sme = 0, sd[3] = {0, 0, 0}, sv[3] = {0, 0, 0}, se[3] = {0, 0, 0}, dd[9] = {0};
    const size_t K = P.e.size();
    for (size_t k = 0; k < K; ++k) {
        float cr[3], m[4];
        qrot(q, &P.u[3 * k], cr);
        sheet_at(M.grad, M.side, 4, cr[0], cr[1], cr[2], m);
        const float d[3] = {cr[1] * m[2] - cr[2] * m[1], cr[2] * m[0] - cr[0] * m[2], cr[0] * m[1] - cr[1] * m[0]};
        const float e = P.e[k], v = m[3];
        n += 1, sm += v, smm += v * v, sme += v * e;
        for (int i = 0; i < 3; ++i) {
            sd[i] += d[i], sv[i] += v * d[i], se[i] += e * d[i];
            for (int j2 = 0; j2 < 3; ++j2) dd[3 * i + j2] += d[i] * d[j2];
        }
    }
    n = std::max(n, 1.0);
    const double mean = sm / n, var_n = smm - n * mean * mean;
    if (!(var_n > 1e-12)) {
        std::fill(grad, grad + 3, 0.f);
        std::fill(hess, hess + 9, 0.f);
        return -2.f;
    }
    const double inv = 1.0 / std::sqrt(var_n), ncc = sme * inv;
    double au[3];
    for (int i = 0; i < 3; ++i) au[i] = (sv[i] - mean * sd[i]) * inv;
    for (int i = 0; i < 3; ++i) grad[i] = (float)((se[i] - ncc * au[i]) * inv);
    for (int i = 0; i < 3; ++i)
        for (int j2 = 0; j2 < 3; ++j2) hess[3 * i + j2] = (float)((dd[3 * i + j2] - sd[i] * sd[j2] / n - au[i] * au[j2]) * inv * inv);
    return (float)ncc;
}

inline void qexp(const float* p, float* q) {
    const float th = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const float k = th <= 1e-3f ? 0.5f - th * th / 48.f : std::sin(0.5f * th) / th;
    q[0] = std::cos(0.5f * th), q[1] = k * p[0], q[2] = k * p[1], q[3] = k * p[2];
}

// solve H x = g (3 x 3, Cramer's rule); false when singular
bool solve3(const float* H, const float* g, float* x) {
    const double a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], gg = H[6], h = H[7], i = H[8];
    const double det = a * (e * i - f * h) - b * (d * i - f * gg) + c * (d * h - e * gg);
    if (std::fabs(det) < 1e-30) return false;
    const double inv = 1.0 / det;
    x[0] = (float)((g[0] * (e * i - f * h) - b * (g[1] * i - f * g[2]) + c * (g[1] * h - e * g[2])) * inv);
    x[1] = (float)((a * (g[1] * i - f * g[2]) - g[0] * (d * i - f * gg) + c * (d * g[2] - g[1] * gg)) * inv);
    x[2] = (float)((a * (e * g[2] - g[1] * h) - b * (d * g[2] - g[1] * gg) + g[0] * (d * h - e * gg)) * inv);
    return true;
}

void refine(const zbc_master& M, const zbc_opts& o, const Pix& P, const float* q0, float* qbest, float* best_out, float* last_out) {
    float cur[4] = {q0[0], q0[1], q0[2], q0[3]};
    float qb[4] = {cur[0], cur[1], cur[2], cur[3]};
    float best = -2.f, mu = o.scale0, gb[3] = {0, 0, 0}, hb[9] = {0};
    float last = 0.f, promised = 0.f, step_size = 0.f, accepted = 0.f;
    for (int it = 0; it <= o.iters; ++it) {
        float g[3], h[9];
        const float ncc = ncc_at(M, P, cur, g, h);
        if (ncc > best) {
            if (accepted > 0.5f && promised > 1e-6f) {
                const float rho = (ncc - best) / promised;
                mu = std::min(std::max(mu * (2.f - rho), mu / 3.f), mu * 3.f);
            }
            std::copy(cur, cur + 4, qb);
            best = ncc;
            std::copy(g, g + 3, gb);
            std::copy(h, h + 9, hb);
            last = step_size;
            accepted += 1.f;
        } else {
            mu *= 3.f;
        }
        mu = std::min(std::max(mu, o.scale_min), o.scale_max);
        if (it == o.iters) break;
        float H[9];
        for (int i = 0; i < 9; ++i) H[i] = hb[i] * mu;
        const float tr = std::max(H[0] + H[4] + H[8], 1e-20f);
        H[0] += 1e-6f * tr, H[4] += 1e-6f * tr, H[8] += 1e-6f * tr;
        float st[3] = {0, 0, 0};
        if (!solve3(H, gb, st)) st[0] = st[1] = st[2] = 0.f;
        float ln = std::sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
        if (ln > o.max_step) {
            for (float& s : st) s *= o.max_step / ln;
            ln = o.max_step;
        }
        float Hs[3];
        for (int i = 0; i < 3; ++i) Hs[i] = H[3 * i] * st[0] + H[3 * i + 1] * st[1] + H[3 * i + 2] * st[2];
        promised = gb[0] * st[0] + gb[1] * st[1] + gb[2] * st[2] - 0.5f * (st[0] * Hs[0] + st[1] * Hs[1] + st[2] * Hs[2]);
        step_size = ln;
        float dq[4], q[4];
        qexp(st, dq);
        qmul(dq, qb, q);
        const float qn = 1.f / std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int i = 0; i < 4; ++i) cur[i] = q[i] * qn;
    }
    std::copy(qb, qb + 4, qbest);
    *best_out = best;
    *last_out = last;
}

float nmi(const zbc_master& M, const Pix& P, const float* q, int bins) {
    const size_t K = P.e.size();
    if (K == 0) return 0.f;
    std::vector<float> s(K);
    for (size_t k = 0; k < K; ++k) {
        float cr[3];
        qrot(q, &P.u[3 * k], cr);
        sheet_at(M.sheets, M.side, 1, cr[0], cr[1], cr[2], &s[k]);
    }
    const auto [emin, emax] = std::minmax_element(P.e.begin(), P.e.end());
    const auto [smin, smax] = std::minmax_element(s.begin(), s.end());
    const float ke = bins / std::max(*emax - *emin, 1e-20f), ks = bins / std::max(*smax - *smin, 1e-20f);
    std::vector<double> H((size_t)bins * bins, 0.0);
    for (size_t k = 0; k < K; ++k) {
        const int be = std::min((int)((P.e[k] - *emin) * ke), bins - 1);
        const int bs = std::min((int)((s[k] - *smin) * ks), bins - 1);
        H[(size_t)be * bins + bs] += 1.0;
    }
    auto ent = [&](const std::vector<double>& x) {
        double h = 0;
        for (double v : x)
            if (v > 0) h -= (v / K) * std::log(v / K);
        return h;
    };
    std::vector<double> ra(bins, 0.0), rb(bins, 0.0);
    for (int i = 0; i < bins; ++i)
        for (int j2 = 0; j2 < bins; ++j2) ra[i] += H[(size_t)i * bins + j2], rb[j2] += H[(size_t)i * bins + j2];
    const double hj = ent(H), hm = ent(ra) + ent(rb);
    return hm > 1e-12 ? (float)(2.0 * (hm - hj) / hm) : 0.f;
}

}  // namespace

extern "C" {

int32_t zbc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// The window's reciprocal denominator volume [n/2 + 1][nc][n] for detector position t:
// 1 / sqrt(<w,f^2> - <w,f>^2 / area), 0 where that is <= 1e-6 area (so3.wgsl:595-626).
int32_t zbc_rden(const zbc_band* b, const zbc_phase* pf, const zbc_phase* pf2, const zbc_frame* f, const float* t, float* rden) {
    if (!b || !pf || !pf2 || !f || !t || !rden) return -1;
    Work W;
    W.size(*b, *pf);
    float stats[4];
    const float area = coefficients(*b, *f, nullptr, t, true, false, false, W, stats);
    const std::vector<cf> walm = W.alm;
    volume(*b, *pf, walm.data(), W);
    const std::vector<float> v1 = W.vol;
    volume(*b, *pf2, walm.data(), W);
    const float a = std::max(area, 1e-12f);
    for (size_t i = 0; i < v1.size(); ++i) {
        const float v = W.vol[i] - v1[i] * v1[i] / a;
        rden[i] = v > 1e-6f * a ? 1.f / std::sqrt(v) : 0.f;
    }
    return 0;
}

// Index n items: records [n][8] = best q (w, x, y, z), NCC, NMI, last improving step, 0; peaks
// [n][8] (q, score, 0, 0, 0; may be NULL). pats: raw float32 [*][h][w]; items [n][4] = (t, pattern
// index); rden from zbc_rden. threads <= 0: OpenMP's default.
int32_t zbc_index(const zbc_band* b, const zbc_phase* ph, const zbc_frame* f, const zbc_master* M, const zbc_opts* o, const float* pats,
                  const float* items, int32_t n, const float* rden, float* records, float* peaks, int32_t threads) {
    if (!b || !ph || !f || !M || !o || !pats || !items || !rden || !records || n < 0) return -1;
#ifdef _OPENMP
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Work W;
        W.size(*b, *ph);
        Pix P;
        std::vector<float> pn((size_t)f->h * f->w);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < n; ++i) {
            const float* t = items + 4 * (size_t)i;
            const float* raw = pats + (size_t)t[3] * f->h * f->w;
            // normalize under the mask (prep.wgsl:399-443)
            double s = 0, cnt = 0;
            for (int r = 0; r < f->h; ++r)
                for (int c = 0; c < f->w; ++c)
                    if (in_mask(*f, (float)r, (float)c)) s += raw[(size_t)r * f->w + c], cnt += 1;
            const float mean = (float)(s / std::max(cnt, 1.0));
            double ss = 0;
            for (int r = 0; r < f->h; ++r)
                for (int c = 0; c < f->w; ++c) {
                    const size_t k = (size_t)r * f->w + c;
                    const float v = in_mask(*f, (float)r, (float)c) ? raw[k] - mean : 0.f;
                    pn[k] = v;
                    ss += (double)v * v;
                }
            const float inv = ss > 0 ? (float)(1.0 / std::sqrt(ss)) : 0.f;
            for (float& v : pn) v *= inv;
            coefficients(*b, *f, pn.data(), t, false, true, true, W, nullptr);
            const std::vector<cf> g = W.alm;
            volume(*b, *ph, g.data(), W);
            float q[4], score;
            peak(*b, *ph, rden, W, q, &score);
            if (peaks) {
                float* pk = peaks + 8 * (size_t)i;
                std::copy(q, q + 4, pk);
                pk[4] = score, pk[5] = pk[6] = pk[7] = 0.f;
            }
            pixels(*f, pn.data(), t, P);
            float qb[4], best, last;
            refine(*M, *o, P, q, qb, &best, &last);
            float* rec = records + 8 * (size_t)i;
            std::copy(qb, qb + 4, rec);
            rec[4] = best;
            rec[5] = nmi(*M, P, qb, o->nmi_bins);
            rec[6] = last;
            rec[7] = 0.f;
        }
    }
    return 0;
}

}  // extern "C"
