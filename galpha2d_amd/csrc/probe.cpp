// probe.cpp -- optional per-kernel timing: HIP events recorded on the launch stream around every
// launch of one named kernel (used by bench.py for the roofline figure).  Off by default; the
// check on every launch is one relaxed load.
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {
std::atomic<int> g_on{0};
std::mutex g_mu;
std::string g_name;
std::vector<hipEvent_t> g_ev;   // pairs
size_t g_used = 0;
}

void ga_probe_begin(const char *k, hipStream_t st)
{
    if (!g_on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_name != k) return;
    if (g_used + 2 > g_ev.size()) {
        for (int i = 0; i < 64; i++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            g_ev.push_back(e);
        }
    }
    (void)hipEventRecord(g_ev[g_used], st);
}

void ga_probe_end(const char *k, hipStream_t st)
{
    if (!g_on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_name != k) return;
    (void)hipEventRecord(g_ev[g_used + 1], st);
    g_used += 2;
}

extern "C" int galphaKernelProbe(const char *kernel_name)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_used = 0;
    if (!kernel_name) {
        g_on = 0;
        g_name.clear();
        return 0;
    }
    g_name = kernel_name;
    g_on = 1;
    return 0;
}

extern "C" int galphaKernelProbeResult(int *launches, double *total_ms)
{
    std::lock_guard<std::mutex> lk(g_mu);
    double tot = 0;
    for (size_t i = 0; i + 1 < g_used; i += 2) {
        if (hipEventSynchronize(g_ev[i + 1]) != hipSuccess) return -1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, g_ev[i], g_ev[i + 1]) != hipSuccess) return -1;
        tot += ms;
    }
    if (launches) *launches = (int)(g_used / 2);
    if (total_ms) *total_ms = tot;
    return 0;
}
