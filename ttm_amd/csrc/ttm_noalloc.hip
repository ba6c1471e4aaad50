// Allocation-free variant of the float32 device entry (ttm_sttm_ex_noalloc of
// include/ttm_amd.h): every device scratch buffer of the call -- the
// partial products of a contraction split, the converted or quantized
// operands and the scale exponents of the reduced-precision kernels -- is
// carved from `scratch`, a fixed array of this library in each GPU's memory,
// instead of the stream-ordered pool.  The call allocates nothing.
//
// Every byte of scratch that a kernel reads was written earlier in the same
// call (partials by their split, operand planes and padding by the
// conversion passes), so the array's contents from an earlier call never
// reach a result.  Calls that overlap on one GPU (from other threads or on
// other streams) share the array and race: the caller serialises them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ttm_amd.h"
#include "ttm_plan.hpp"
#include "ttm_workspace.hpp"

namespace ttm_amd
{
ttm_status_t sttm_device(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                         const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                         const size_t *nc, const size_t *wc, ttm_precision_t prec, hipStream_t s);
} // namespace ttm_amd

namespace
{
constexpr size_t kScratchBytes = TTM_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");

// In host code `scratch` names a host-side stand-in, not the device array:
// kernels and copies only ever get the address that hipGetSymbolAddress
// returns for the current device.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device
// with that device current.
ttm_status_t scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        ttm_amd::set_error(TTM_ERR_HIP, e != hipSuccess ? hipGetErrorString(e) : "device index out of range");
        return TTM_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            ttm_amd::set_error(TTM_ERR_HIP, e != hipSuccess ? hipGetErrorString(e) : "scratch address is null");
            return TTM_ERR_HIP;
        }
        g_ptr[dev] = ptr;
    }
    *out = g_ptr[dev];
    return TTM_SUCCESS;
}
} // namespace

extern "C" ttm_status_t ttm_sttm_ex_noalloc(unsigned q, unsigned p, const float *a, const size_t *na,
                                            const size_t *wa, const size_t *pia, const float *b, const size_t *nb,
                                            const size_t *pib, float *c, const size_t *nc, const size_t *wc,
                                            ttm_precision_t precision, ttm_stream_t stream)
{
    void *base = nullptr;
    if (ttm_status_t const st = scratch_address(&base); st != TTM_SUCCESS) return st;
    ttm_amd::ScratchScope scope(base, kScratchBytes);
    return ttm_amd::sttm_device(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, precision,
                                static_cast<hipStream_t>(stream));
}
