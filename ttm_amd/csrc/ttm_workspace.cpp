// Device workspace of one product call: per-device pool, or a fixed region
// bound by an allocation-free entry point.
#include "ttm_workspace.hpp"

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "ttm_plan.hpp"

namespace ttm_amd
{

namespace
{
thread_local void *t_base = nullptr;
thread_local size_t t_cap = 0;
thread_local size_t t_used = 0;
} // namespace

hipMemPool_t workspace_pool()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    static std::mutex mu;
    static std::vector<hipMemPool_t> pools;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)pools.size() <= dev) pools.resize(dev + 1, nullptr);
    if (!pools[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) != hipSuccess) return nullptr;
        uint64_t keep = UINT64_MAX;
        if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            return nullptr;
        }
        pools[dev] = pool;
    }
    return pools[dev];
}

ttm_status_t ws_acquire(size_t bytes, hipStream_t s, void **out, const char *what)
{
    *out = nullptr;
    size_t const need = (bytes + 255) / 256 * 256;
    if (t_base) {
        if (need > t_cap - t_used) {
            char msg[256];
            std::snprintf(msg, sizeof msg, "%s: %zu bytes of scratch needed, %zu of the fixed %zu-byte region left",
                          what, need, t_cap - t_used, t_cap);
            set_error(TTM_ERR_ALLOC, msg);
            return TTM_ERR_ALLOC;
        }
        *out = static_cast<char *>(t_base) + t_used;
        t_used += need;
        return TTM_SUCCESS;
    }
    hipMemPool_t const pool = workspace_pool();
    if (!pool || hipMallocFromPoolAsync(out, need ? need : 256, pool, s) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        set_error(TTM_ERR_ALLOC, what);
        return TTM_ERR_ALLOC;
    }
    return TTM_SUCCESS;
}

void ws_release(void *p, hipStream_t s)
{
    if (!p || t_base) return;
    (void)hipFreeAsync(p, s);
}

ScratchScope::ScratchScope(void *base, size_t cap) : prev_base_(t_base), prev_cap_(t_cap), prev_used_(t_used)
{
    t_base = base;
    t_cap = cap;
    t_used = 0;
}

ScratchScope::~ScratchScope()
{
    t_base = prev_base_;
    t_cap = prev_cap_;
    t_used = prev_used_;
}

bool scratch_bound() { return t_base != nullptr; }

size_t scratch_available() { return t_base ? t_cap - t_used : SIZE_MAX; }

// Queried once per device; concurrent first calls store the same value.
int device_cus()
{
    static std::atomic<int> cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 256;
    }
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            n = 256;
        }
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

} // namespace ttm_amd
