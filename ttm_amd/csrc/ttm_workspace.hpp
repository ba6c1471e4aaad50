// Device workspace of one product call (split-contraction partials, converted
// or quantized operands, scale exponents).
//
// By default a call takes its workspace from a per-device stream-ordered
// memory pool that keeps its memory between calls (hipMallocFromPoolAsync /
// hipFreeAsync), so a steady stream of products allocates nothing new from
// the driver after the first.  An allocation-free entry point instead binds a
// fixed device region for the duration of its call (ScratchScope); requests
// are then carved from that region and fail with TTM_ERR_ALLOC when it is too
// small.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "ttm_amd.h"

namespace ttm_amd
{

// The per-device pool (nullptr when it cannot be created).
hipMemPool_t workspace_pool();

// Workspace of `bytes` bytes ordered on stream s; *out is 256-byte aligned.
// On failure the thread's last error names `what`.
ttm_status_t ws_acquire(size_t bytes, hipStream_t s, void **out, const char *what);
// Return a workspace taken with ws_acquire (stream ordered; no-op for a
// fixed region).
void ws_release(void *p, hipStream_t s);

// While alive, ws_acquire on this thread hands out consecutive pieces of
// [base, base + cap) instead of pool memory.
class ScratchScope
{
  public:
    ScratchScope(void *base, size_t cap);
    ~ScratchScope();
    ScratchScope(const ScratchScope &) = delete;
    ScratchScope &operator=(const ScratchScope &) = delete;

  private:
    void *prev_base_;
    size_t prev_cap_, prev_used_;
};

// True while a ScratchScope is bound on this thread.
bool scratch_bound();
// Bytes a ws_acquire can still hand out: the rest of the bound region, or
// SIZE_MAX for the pool.
size_t scratch_available();

// Compute units of the current device (256 when it cannot be queried).
int device_cus();

} // namespace ttm_amd
