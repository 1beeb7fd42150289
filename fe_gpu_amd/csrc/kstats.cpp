#include "kstats.hpp"

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/fehip.h"
#include "common.hpp"

namespace fehip {
namespace {

struct Pending {
    int id;
    hipEvent_t a, b;
    double bytes, flops;
};

struct Acc {
    int64_t launches = 0;
    double ms = 0.0, bytes = 0.0, flops = 0.0;
};

std::mutex g_mu;
bool g_on = false;
std::vector<Pending> g_pending;
hipEvent_t g_open[kStatCount];
Acc g_acc[kStatCount];
const char* g_names[kStatCount] = {"apply_iter", "apply_data", "finalize",    "reduce_runs", "gram",
                                   "segsum",     "apply_iter32", "apply_resid", "couplings"};

void resolve_locked() {
    for (auto& p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_acc[p.id].ms += ms;
            g_acc[p.id].launches += 1;
            g_acc[p.id].bytes += p.bytes;
            g_acc[p.id].flops += p.flops;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_pending.clear();
}

}  // namespace

bool kstats_enabled() { return g_on; }

void kstat_begin(int id, hipStream_t s) {
    if (!g_on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    std::lock_guard<std::mutex> l(g_mu);
    g_open[id] = e;
}

void kstat_end(int id, hipStream_t s, double bytes, double flops) {
    if (!g_on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    std::lock_guard<std::mutex> l(g_mu);
    g_pending.push_back({id, g_open[id], e, bytes, flops});
    if (g_pending.size() > 4096) resolve_locked();
}

}  // namespace fehip

using namespace fehip;

extern "C" {

int fehip_kstats_enable(int on) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    g_on = on != 0;
    return set_ok();
}

int fehip_kstats_reset(void) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    for (auto& a : g_acc) a = Acc();
    return set_ok();
}

int fehip_kstats_get(int id, char* name, int name_len, int64_t* launches, double* total_ms, double* bytes,
                     double* flops) {
    if (id < 0 || id >= kStatCount) return set_error("fehip_kstats_get: id out of range");
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    if (name && name_len > 0) {
        std::strncpy(name, g_names[id], name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (launches) *launches = g_acc[id].launches;
    if (total_ms) *total_ms = g_acc[id].ms;
    if (bytes) *bytes = g_acc[id].bytes;
    if (flops) *flops = g_acc[id].flops;
    return set_ok();
}

}  // extern "C"
