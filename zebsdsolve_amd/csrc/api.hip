// Error reporting and version of the C ABI (include/zebsd.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace zb {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: launch failed: %s", what, hipGetErrorString(e));
        return ZB_ELAUNCH;
    }
    return ZB_OK;
}

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

static std::mutex g_prof_mu;
static std::atomic<bool> g_prof{false};
static std::vector<ProfRec> g_recs;
static std::vector<hipEvent_t> g_pool;

static hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

bool prof_on() { return g_prof.load(std::memory_order_relaxed); }

static std::vector<std::string> g_prof_only;  // kernels to time (empty: all), under g_prof_mu

bool prof_wants(const char* name) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    if (g_prof_only.empty()) return true;
    for (auto& k : g_prof_only)
        if (k == name) return true;
    return false;
}

void prof_begin(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    ProfRec r{name, take_event(), take_event()};
    if (r.a) (void)hipEventRecord(r.a, s);
    g_recs.push_back(r);
}

void prof_end(hipStream_t s) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    if (!g_recs.empty() && g_recs.back().b) (void)hipEventRecord(g_recs.back().b, s);
}

// Launch-shape knobs (the choices an autotuner makes): none changes any result bit.
struct Knob {
    const char* name;
    int lo, hi, def;
    std::atomic<int> v;
};
static Knob g_knobs[] = {
    {"refine_queue", 0, 2, 0, {0}},  // zb_refine_phases: 0 a block per work item, 1 the persistent queue grid, 2 that
                                     // grid with the NMI's master values looked up again (fewer registers)
    {"rings_group", 0, 64, 0, {0}},  // ring pairs per k_rings block (0: 2048 / (2n), at most n / 2)
    {"prep_global", 0, 1, 0, {0}},   // zb_prep_work: 1 the filters through global memory even where LDS holds the image
};
static_assert(sizeof g_knobs / sizeof g_knobs[0] == KNOB_COUNT, "a knob per id");

int knob(int id) { return g_knobs[id].v.load(std::memory_order_relaxed); }

static Knob* find_knob(const char* name) {
    if (!name) return nullptr;
    for (auto& k : g_knobs)
        if (!strcmp(k.name, name)) return &k;
    return nullptr;
}

}  // namespace zb

// Sets a knob; returns its previous value, or ZB_EINVAL for an unknown name or a value out of
// range. Process-wide, read at each launch.
extern "C" int32_t zb_knob_set(const char* name, int32_t value) {
    zb::Knob* k = zb::find_knob(name);
    if (!k) {
        zb::set_error("zb_knob_set: no knob %s", name ? name : "(null)");
        return ZB_EINVAL;
    }
    if (value < k->lo || value > k->hi) {
        zb::set_error("zb_knob_set: %s = %d outside %d..%d", k->name, value, k->lo, k->hi);
        return ZB_EINVAL;
    }
    return k->v.exchange(value);
}

extern "C" int32_t zb_knob_get(const char* name) {
    zb::Knob* k = zb::find_knob(name);
    if (!k) {
        zb::set_error("zb_knob_get: no knob %s", name ? name : "(null)");
        return ZB_EINVAL;
    }
    return k->v.load();
}

extern "C" const char* zb_version(void) { return "0.1.0"; }

// Kernel timing, the counterpart of the reference's Gpu.setProfile / profileReport
// (src/gpu/gpu.zig:287-313; python/zebsdsolve/__init__.py:944-951). No counterpart in the C ABI
// of the reference. on = 1 starts recording (and clears what was recorded), 0 stops.
extern "C" int32_t zb_profile(int32_t on) {
    std::lock_guard<std::mutex> g(zb::g_prof_mu);
    for (auto& r : zb::g_recs) {
        if (r.a) zb::g_pool.push_back(r.a);
        if (r.b) zb::g_pool.push_back(r.b);
    }
    zb::g_recs.clear();
    zb::g_prof.store(on != 0);
    return ZB_OK;
}

// Only the kernels named in `names` (comma-separated) are timed while profiling is on; NULL or
// "" times every kernel. Fewer events in the stream: the others run as they do unprofiled.
extern "C" int32_t zb_profile_select(const char* names) {
    std::lock_guard<std::mutex> g(zb::g_prof_mu);
    zb::g_prof_only.clear();
    if (!names) return ZB_OK;
    std::string cur;
    for (const char* c = names;; ++c) {
        if (*c == ',' || *c == 0) {
            if (!cur.empty()) zb::g_prof_only.push_back(cur);
            cur.clear();
            if (!*c) break;
        } else if (*c != ' ') {
            cur += *c;
        }
    }
    return ZB_OK;
}

// Waits for the recorded launches and writes "name count total_ms\n" per kernel name into buf
// (the bytes needed are returned; the text is cut at cap - 1).
extern "C" int32_t zb_profile_report(char* buf, int32_t cap) {
    std::lock_guard<std::mutex> g(zb::g_prof_mu);
    std::map<std::string, std::pair<long, double>> acc;
    for (auto& r : zb::g_recs) {
        float ms = 0.f;
        if (r.a && r.b && hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& e = acc[r.name];
            e.first += 1;
            e.second += ms;
        }
    }
    std::string out;
    char line[256];
    for (auto& kv : acc) {
        snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        out += line;
    }
    if (buf && cap > 0) {
        size_t k = out.size() < (size_t)(cap - 1) ? out.size() : (size_t)(cap - 1);
        memcpy(buf, out.data(), k);
        buf[k] = 0;
    }
    return (int32_t)out.size();
}

extern "C" int32_t zb_last_error(char* buf, int32_t cap) {
    int32_t n = (int32_t)strlen(zb::g_err);
    if (buf && cap > 0) {
        int32_t k = n < cap - 1 ? n : cap - 1;
        memcpy(buf, zb::g_err, k);
        buf[k] = 0;
    }
    return n;
}
