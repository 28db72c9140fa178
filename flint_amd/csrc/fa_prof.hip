// fa_prof.hip -- per-kernel launch timing with HIP events (bench support)
// and a plain copy kernel used to calibrate hardware byte counters.
#include <string.h>

#include <vector>

#include "fa_internal.h"

namespace {
struct pending_t {
    const char* name;
    hipEvent_t start, stop;
};
struct total_t {
    const char* name;
    double ms;
    long count;
};
thread_local int prof_enabled = 0;
thread_local std::vector<pending_t> pending;
thread_local std::vector<total_t> totals;
thread_local std::vector<hipEvent_t> event_pool;

hipEvent_t take_event()
{
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return NULL;
    return e;
}
}  // namespace

int fa_prof_on() { return prof_enabled; }

void fa_prof_begin(const char* name, hipStream_t s)
{
    pending_t p;
    p.name = name;
    p.start = take_event();
    p.stop = take_event();
    if (!p.start || !p.stop) return;
    (void)hipEventRecord(p.start, s);
    pending.push_back(p);
}

void fa_prof_end(hipStream_t s)
{
    if (pending.empty()) return;
    (void)hipEventRecord(pending.back().stop, s);
}

// accumulate finished launches into the per-name totals
static void collect()
{
    for (auto& p : pending) {
        (void)hipEventSynchronize(p.stop);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.start, p.stop);
        bool found = false;
        for (auto& t : totals)
            if (strcmp(t.name, p.name) == 0) {
                t.ms += ms;
                t.count++;
                found = true;
                break;
            }
        if (!found) totals.push_back({p.name, (double)ms, 1});
        event_pool.push_back(p.start);
        event_pool.push_back(p.stop);
    }
    pending.clear();
}

extern "C" void fa_prof_enable(int on)
{
    if (!on) collect();
    prof_enabled = on;
}

extern "C" void fa_prof_reset(void)
{
    collect();
    totals.clear();
}

// copies up to max entries: names (64 chars each), total ms, launch counts
extern "C" int fa_prof_read(char* names, double* ms, int64_t* counts, int max)
{
    collect();
    int n = 0;
    for (auto& t : totals) {
        if (n >= max) break;
        strncpy(names + 64 * n, t.name, 63);
        names[64 * n + 63] = 0;
        ms[n] = t.ms;
        counts[n] = t.count;
        n++;
    }
    return n;
}

__global__ void fa_k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// plain device copy of nbytes (a multiple of 16) for counter calibration
extern "C" int fa_copy_device(void* dst, const void* src, uint64_t nbytes, void* stream)
{
    if (nbytes % 16) return fa_set_error(FA_E_ARG, "fa_copy_device: size not a multiple of 16");
    uint64_t n = nbytes / 16;
    unsigned grid = (unsigned)((n + 255) / 256);
    if (grid > 65536) grid = 65536;
    FA_LAUNCH("fa_k_copy", fa_k_copy, dim3(grid), dim3(256), 0, (hipStream_t)stream,
              (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError() == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "copy launch failed");
}
