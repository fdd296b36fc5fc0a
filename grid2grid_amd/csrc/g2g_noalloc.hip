// g2g_transform_plans_noalloc (include/grid2grid_amd.h): several plans in one
// persistent gather launch, like g2g_batch_create + g2g_batch_execute, with
// every device buffer the launch needs beyond the callers' blocks -- the
// work descriptors, the per-work-item records and the work counters -- in
// one file-scope device array instead of allocations, for callers that
// cannot allocate (code under hipGraph capture, serving loops that must not
// call hipMalloc). Every call on a device shares that device's array: calls
// that overlap on one GPU race.
#include "grid2grid_amd.h"
#include "g2g_gather.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace g2g {
namespace {

// Layout of the array: work counters, then descriptors, then work records.
constexpr size_t kNoallocCounterBytes = 256;
constexpr int kNoallocMaxPlans = 4096;
constexpr size_t kNoallocDescBytes = sizeof(work_desc) * kNoallocMaxPlans;
constexpr size_t kScratchBytes = 4u << 20;  // 4 MiB per device, held until the process ends
constexpr size_t kNoallocMaxWork = (kScratchBytes - kNoallocCounterBytes - kNoallocDescBytes) / sizeof(work_rec);
static_assert(kNoallocDescBytes % 256 == 0, "work records stay 256-byte aligned");
static_assert(kNoallocMaxWork >= 32768, "room for the work records of a sizeable launch");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the address of that device's copy of `scratch`.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

int scratch_address(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return internal_fail(G2G_ERR_HIP, "noalloc: no current device (or device index above %d)", kMaxDevices - 1);
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return internal_fail(G2G_ERR_HIP, "noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                                 hipGetErrorString(e));
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return G2G_OK;
}

// Descriptors travel to the device as kernel arguments, kDescChunk at a time:
// a launch only, no host-to-device copy.
constexpr int kDescChunk = 16;
struct desc_chunk {
    work_desc d[kDescChunk];
    uint32_t first;  // index of d[0] in the table
    uint32_t n;
};
static_assert(sizeof(desc_chunk) <= 4000, "kernel argument size");

__global__ void k_put_descs(work_desc *table, desc_chunk c) {
    if (threadIdx.x < c.n) table[c.first + threadIdx.x] = c.d[threadIdx.x];
}

}  // namespace
}  // namespace g2g

using namespace g2g;

extern "C" int g2g_transform_plans_noalloc(int n, const g2g_plan *plans, const void *const *alphas,
                                           const void *const *betas, void *stream) {
    if (n < 0 || (n > 0 && !plans)) return internal_fail(G2G_ERR_INVALID, "noalloc: bad arguments");
    if (n > kNoallocMaxPlans)
        return internal_fail(G2G_ERR_INVALID, "noalloc: %d plans exceed the maximum of %d per launch", n,
                             kNoallocMaxPlans);
    if (n == 0) return G2G_OK;
    hipStream_t s = (hipStream_t)stream;
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc != G2G_OK) return rc;
    uint32_t *counters = (uint32_t *)base;
    work_desc *table = (work_desc *)((unsigned char *)base + kNoallocCounterBytes);
    work_rec *work = (work_rec *)((unsigned char *)base + kNoallocCounterBytes + kNoallocDescBytes);
    // descriptors (checked for size first, so that nothing is launched for a
    // request the array cannot hold)
    uint64_t tiles = 0;
    launch_cfg cfg;
    cfg.features = FEAT_COPY;
    for (int k = 0; k < n; ++k) {
        work_desc d;
        launch_cfg c;
        rc = internal_desc(plans[k], alphas ? alphas[k] : nullptr, betas ? betas[k] : nullptr, (uint32_t)tiles, d, c);
        if (rc != G2G_OK) return rc;
        if (k == 0) {
            cfg.threads = c.threads;
            cfg.unroll_bytes = c.unroll_bytes;
            cfg.nontemporal = c.nontemporal;
        }
        cfg.features = std::max(cfg.features, c.features);
        tiles += d.n_tiles;
    }
    if (tiles > kNoallocMaxWork)
        return internal_fail(G2G_ERR_INVALID,
                             "noalloc: %llu work items exceed the maximum of %llu per launch (4 MiB of scratch); "
                             "split the plans over several calls or retune them to larger work items",
                             (unsigned long long)tiles, (unsigned long long)kNoallocMaxWork);
    desc_chunk chunk;
    uint32_t base_tile = 0;
    for (int k0 = 0; k0 < n; k0 += kDescChunk) {
        chunk.first = (uint32_t)k0;
        chunk.n = (uint32_t)std::min(kDescChunk, n - k0);
        for (uint32_t j = 0; j < chunk.n; ++j) {
            launch_cfg c;
            const int k = k0 + (int)j;
            rc = internal_desc(plans[k], alphas ? alphas[k] : nullptr, betas ? betas[k] : nullptr, base_tile,
                               chunk.d[j], c);
            if (rc != G2G_OK) return rc;
            base_tile += chunk.d[j].n_tiles;
        }
        hipLaunchKernelGGL(k_put_descs, dim3(1), dim3(64), 0, s, table, chunk);
        if (hipGetLastError() != hipSuccess) return internal_fail(G2G_ERR_HIP, "noalloc: descriptor launch failed");
    }
    if (launch_work_records(table, n, work, s) != hipSuccess)
        return internal_fail(G2G_ERR_HIP, "noalloc: work record launch failed");
    gather_args a;
    std::memset(&a, 0, sizeof a);
    a.descs = table;
    a.work = work;
    a.total_tiles = (uint32_t)tiles;
    a.queue = counters;  // zeroed by the launch's own reset kernel
    cfg.persistent = true;
    cfg.xcd_ranges = false;
    const hipError_t e = launch_gather(a, cfg, internal_grid_limit(), s);
    if (e != hipSuccess) return internal_fail(G2G_ERR_HIP, "noalloc: gather launch failed: %s", hipGetErrorString(e));
    return G2G_OK;
}
