// vgi_runtime.cpp — error reporting for the C ABI (thread-local last error).
#include <atomic>
#include "vgi_runtime.h"

#include <string.h>

#include <mutex>
#include <string>
#include <vector>

namespace vgi {

static thread_local char g_last_error[512] = "";

vgi_status set_error(vgi_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

vgi_status check_hip(hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory) return set_error(VGI_E_NOMEM, "%s: %s", what, hipGetErrorString(e));
    return set_error(VGI_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

vgi_status check_launch(const char* kernel) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(VGI_E_HIP, "launch of %s failed: %s", kernel, hipGetErrorString(e));
    return VGI_OK;
}

// ---------------------------------------------------------------------------
// per-pass profiler
// ---------------------------------------------------------------------------
namespace {
struct Pending {
    int id;
    hipEvent_t a, b;
    double bytes;
};
struct Accum {
    std::string name;
    double ms = 0;
    uint64_t count = 0;
    double bytes = 0;
    double trace[4] = {0, 0, 0, 0};  // closest rays, any-hit rays, nodes, tris
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::string g_prof_only;  // record only this pass name (empty = all)
std::vector<Accum> g_prof_acc;
std::vector<Pending> g_prof_pending;
std::vector<hipEvent_t> g_prof_pool;

hipEvent_t take_event() {
    if (!g_prof_pool.empty()) {
        hipEvent_t e = g_prof_pool.back();
        g_prof_pool.pop_back();
        return e;
    }
    // timing-only events: no system-scope fence (cache writeback and
    // invalidation) on record, which would otherwise add to the measured
    // interval and slow the work that follows
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}
int name_id(const char* name) {
    for (size_t i = 0; i < g_prof_acc.size(); ++i)
        if (g_prof_acc[i].name == name) return (int)i;
    g_prof_acc.push_back(Accum{name});
    return (int)g_prof_acc.size() - 1;
}
void drain_locked() {
    for (auto& p : g_prof_pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_prof_acc[p.id].ms += ms;
            g_prof_acc[p.id].count += 1;
            g_prof_acc[p.id].bytes += p.bytes;
        }
        g_prof_pool.push_back(p.a);
        g_prof_pool.push_back(p.b);
    }
    g_prof_pending.clear();
}
}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t s, double bytes) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_on) return;
    if (!g_prof_only.empty() && g_prof_only != name) return;
    a_ = take_event();
    if (!a_) return;
    if (hipEventRecord(a_, s) != hipSuccess) {
        g_prof_pool.push_back(a_);
        a_ = nullptr;
        return;
    }
    id_ = name_id(name);
    s_ = s;
    bytes_ = bytes;
}

ProfScope::~ProfScope() {
    if (!a_) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    hipEvent_t b = take_event();
    if (!b || hipEventRecord(b, s_) != hipSuccess) {
        g_prof_pool.push_back(a_);
        if (b) g_prof_pool.push_back(b);
        return;
    }
    g_prof_pending.push_back(Pending{id_, a_, b, bytes_});
}

// ---------------------------------------------------------------------------
// traversal statistics
// ---------------------------------------------------------------------------
namespace {
std::mutex g_stats_mu;
unsigned long long* g_stats_dev = nullptr;
constexpr size_t kStatsWords = 4u * (size_t)kTraceCounterSlots;
}  // namespace

unsigned long long* trace_counters() {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    return g_stats_dev;
}

void trace_stats_collect(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    if (!g_stats_dev) return;
    unsigned long long host[kStatsWords];
    if (hipStreamSynchronize(s) != hipSuccess) return;
    if (hipMemcpy(host, g_stats_dev, sizeof(host), hipMemcpyDeviceToHost) != hipSuccess) return;
    double sum[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < kStatsWords; ++i) sum[i & 3u] += (double)host[i];
    memset(host, 0, sizeof(host));
    (void)hipMemcpy(g_stats_dev, host, sizeof(host), hipMemcpyHostToDevice);
    std::lock_guard<std::mutex> lp(g_prof_mu);
    if (!g_prof_on) return;
    Accum& a = g_prof_acc[name_id(name)];
    for (int k = 0; k < 4; ++k) a.trace[k] += sum[k];
}

}  // namespace vgi

extern "C" vgi_status vgi_trace_stats_enable(uint32_t enable) {
    std::lock_guard<std::mutex> lk(vgi::g_stats_mu);
    if (enable && !vgi::g_stats_dev) {
        void* p = nullptr;
        VGI_TRY_HIP(hipMalloc(&p, vgi::kStatsWords * sizeof(unsigned long long)));
        VGI_TRY_HIP(hipMemset(p, 0, vgi::kStatsWords * sizeof(unsigned long long)));
        vgi::g_stats_dev = (unsigned long long*)p;
    } else if (!enable && vgi::g_stats_dev) {
        VGI_TRY_HIP(hipDeviceSynchronize());
        VGI_TRY_HIP(hipFree(vgi::g_stats_dev));
        vgi::g_stats_dev = nullptr;
    }
    return VGI_OK;
}

extern "C" vgi_status vgi_profiler_enable(uint32_t enable) {
    std::lock_guard<std::mutex> lk(vgi::g_prof_mu);
    vgi::g_prof_on = enable != 0;
    return VGI_OK;
}

namespace vgi {
static std::atomic<uint32_t> g_atrous_block{16}, g_trace_tile{8};
uint32_t tuning_atrous_block() { return g_atrous_block.load(std::memory_order_relaxed); }
uint32_t tuning_trace_tile_width() { return g_trace_tile.load(std::memory_order_relaxed); }
}  // namespace vgi

extern "C" vgi_status vgi_tuning_set(const vgi_tuning* t) {
    const uint32_t ab = t && t->atrous_block ? t->atrous_block : 16u;
    const uint32_t tw = t && t->trace_tile_width ? t->trace_tile_width : 8u;
    VGI_REQUIRE(ab == 8 || ab == 16 || ab == 32, "vgi_tuning_set: atrous_block %u is not 8, 16 or 32", ab);
    VGI_REQUIRE(tw == 8 || tw == 16 || tw == 32, "vgi_tuning_set: trace_tile_width %u is not 8, 16 or 32", tw);
    vgi::g_atrous_block.store(ab, std::memory_order_relaxed);
    vgi::g_trace_tile.store(tw, std::memory_order_relaxed);
    return VGI_OK;
}

extern "C" vgi_status vgi_tuning_get(vgi_tuning* t) {
    VGI_REQUIRE(t, "vgi_tuning_get: null argument");
    t->atrous_block = vgi::tuning_atrous_block();
    t->trace_tile_width = vgi::tuning_trace_tile_width();
    return VGI_OK;
}

extern "C" vgi_status vgi_profiler_only(const char* name) {
    std::lock_guard<std::mutex> lk(vgi::g_prof_mu);
    vgi::g_prof_only = name ? name : "";
    return VGI_OK;
}

extern "C" vgi_status vgi_profiler_reset(void) {
    std::lock_guard<std::mutex> lk(vgi::g_prof_mu);
    vgi::drain_locked();
    vgi::g_prof_acc.clear();
    return VGI_OK;
}

extern "C" int32_t vgi_profiler_read(vgi_profile_entry* out, int32_t cap) {
    std::lock_guard<std::mutex> lk(vgi::g_prof_mu);
    vgi::drain_locked();
    int32_t n = 0;
    for (auto& a : vgi::g_prof_acc) {
        if (out && n < cap) {
            memset(&out[n], 0, sizeof(out[n]));
            snprintf(out[n].name, sizeof(out[n].name), "%s", a.name.c_str());
            out[n].total_ms = a.ms;
            out[n].count = a.count;
            out[n].bytes = a.bytes;
            out[n].rays_closest = a.trace[0];
            out[n].rays_any = a.trace[1];
            out[n].nodes = a.trace[2];
            out[n].tris = a.trace[3];
        }
        ++n;
    }
    return n;
}

extern "C" const char* vgi_last_error(void) { return vgi::g_last_error; }

extern "C" const char* vgi_version(void) { return "vgi 0.1 (gfx950)"; }
