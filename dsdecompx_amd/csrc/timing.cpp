// timing.cpp -- per-kernel HIP-event timing behind dsdx_timing_* (C ABI).
#include "timing.hpp"

#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "dsdx_hip.h"

namespace dsdx::timing {
namespace {

const char* const kNames[kKernelCount] = {"migrate_wrap", "halo_classify", "halo_scan",
                                          "halo_scatter", "reverse_halo", "reverse_exchange",
                                          "forward_exchange", "halo_refresh", "ghost_owners",
                                          "migrate_classify", "halo_copy",
                                          "grid_migrate_count", "grid_migrate_scatter",
                                          "grid_exchange_counts", "halo_pack", "unpack_records",
                                          "pack_reverse", "accumulate_reverse", "loopback_exchange"};

struct Pending {
    Kernel kernel;
    hipEvent_t start;
    hipEvent_t stop;
};

std::atomic<bool> g_enabled{false};
std::mutex g_mutex;
std::vector<hipEvent_t> g_pool;
std::vector<Pending> g_pending;
double g_total_ms[kKernelCount] = {};
uint64_t g_launches[kKernelCount] = {};

hipEvent_t take_event()
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Waits for every recorded pair and folds it into the totals.
int resolve()
{
    std::lock_guard<std::mutex> lock(g_mutex);
    int rc = DSDX_OK;
    for (const Pending& p : g_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.stop) != hipSuccess ||
            hipEventElapsedTime(&ms, p.start, p.stop) != hipSuccess) {
            rc = DSDX_EDEVICE;
        } else {
            g_total_ms[p.kernel] += ms;
            g_launches[p.kernel] += 1;
        }
        g_pool.push_back(p.start);
        g_pool.push_back(p.stop);
    }
    g_pending.clear();
    return rc;
}

}  // namespace

bool enabled() { return g_enabled.load(std::memory_order_relaxed); }

namespace {
thread_local Scope* t_open = nullptr;

void give_back(hipEvent_t e)
{
    if (!e) return;
    std::lock_guard<std::mutex> lock(g_mutex);
    g_pool.push_back(e);
}
}  // namespace

Scope::Scope(Kernel kernel, hipStream_t stream) : kernel_(kernel)
{
    if (!enabled()) return;
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &capture) != hipSuccess || capture != hipStreamCaptureStatusNone)
        return;
    start_ = take_event();
    stop_ = take_event();
    if (!start_ || !stop_) {
        give_back(start_);
        give_back(stop_);
        start_ = stop_ = nullptr;
        return;
    }
    outer_ = t_open;
    t_open = this;
}

Scope::~Scope()
{
    if (!start_) return;
    t_open = outer_;
    if (launched_) {
        std::lock_guard<std::mutex> lock(g_mutex);
        g_pending.push_back({kernel_, start_, stop_});
    } else {
        give_back(start_);
        give_back(stop_);
    }
}

void claim(hipEvent_t* start, hipEvent_t* stop)
{
    Scope* s = t_open;
    if (!s || s->launched_) {
        *start = *stop = nullptr;
        return;
    }
    s->launched_ = true;
    *start = s->start_;
    *stop = s->stop_;
}

}  // namespace dsdx::timing

extern "C" {

int dsdx_timing_enable(int32_t enable)
{
    dsdx::timing::g_enabled.store(enable != 0);
    return DSDX_OK;
}

int dsdx_timing_reset(void)
{
    const int rc = dsdx::timing::resolve();
    std::lock_guard<std::mutex> lock(dsdx::timing::g_mutex);
    std::memset(dsdx::timing::g_total_ms, 0, sizeof(dsdx::timing::g_total_ms));
    std::memset(dsdx::timing::g_launches, 0, sizeof(dsdx::timing::g_launches));
    return rc;
}

int dsdx_timing_read(const char* kernel, double* total_ms, uint64_t* launches)
{
    if (!kernel) return DSDX_EINVAL;
    const int rc = dsdx::timing::resolve();
    std::lock_guard<std::mutex> lock(dsdx::timing::g_mutex);
    for (int k = 0; k < dsdx::timing::kKernelCount; ++k) {
        if (std::strcmp(kernel, dsdx::timing::kNames[k]) == 0) {
            if (total_ms) *total_ms = dsdx::timing::g_total_ms[k];
            if (launches) *launches = dsdx::timing::g_launches[k];
            return rc;
        }
    }
    return DSDX_EINVAL;
}

}  // extern "C"
