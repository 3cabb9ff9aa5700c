// Library-wide entry points: ABI version and the per-thread error string.

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "td3_common.h"

namespace {
thread_local char g_error[1024] = "";

// Kernel timing: when enabled, every timed launch site brackets its kernel with two
// HIP events on the launch stream; td3_timing_get adds the elapsed times per name.
struct Timed {
    const char* name;
    hipEvent_t start, stop;
};
std::mutex g_timing_mutex;
bool g_timing = false;
std::vector<Timed> g_timed;
}  // namespace

bool td3_timing_on() { return g_timing; }

void* td3_timing_begin(hipStream_t stream) {
    if (!g_timing) return nullptr;
    hipEvent_t* ev = new hipEvent_t[2];
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
        delete[] ev;
        return nullptr;
    }
    hipEventRecord(ev[0], stream);
    return ev;
}

void td3_timing_end(void* token, const char* name, hipStream_t stream) {
    if (token == nullptr) return;
    hipEvent_t* ev = (hipEvent_t*)token;
    hipEventRecord(ev[1], stream);
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    g_timed.push_back({name, ev[0], ev[1]});
    delete[] ev;
}

extern "C" int td3_timing_enable(int on) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    for (auto& t : g_timed) {
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    g_timed.clear();
    g_timing = on != 0;
    return TD3_OK;
}

extern "C" int td3_timing_get(const char* name, double* total_ms, int64_t* launches) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    double total = 0.0;
    int64_t count = 0;
    for (auto& t : g_timed) {
        if (strcmp(t.name, name) != 0) continue;
        TD3_HIP(hipEventSynchronize(t.stop));
        float ms = 0.0f;
        TD3_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        total += ms;
        ++count;
    }
    *total_ms = total;
    *launches = count;
    return TD3_OK;
}

void td3_set_error(const char* fmt, ...) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
}

extern "C" int td3_abi_version(void) { return TD3_ABI_VERSION; }

extern "C" const char* td3_last_error(void) { return g_error; }
