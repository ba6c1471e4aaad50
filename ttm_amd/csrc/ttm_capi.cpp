// extern "C" entry points declared in include/ttm_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "ttm_amd.h"
#include "ttm_kernels.hpp"
#include "ttm_plan.hpp"
#include "ttm_tune.hpp"
#include "ttm_workspace.hpp"

using namespace ttm_amd;

namespace
{

Kernel g_force = Kernel::none; // test hook, see ttm_force_kernel

bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof attr);
    hipError_t const e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError(); // unregistered host memory reports an error; clear it
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

ttm_status_t hip_fail(hipError_t e)
{
    set_error(TTM_ERR_HIP, hipGetErrorString(e));
    return TTM_ERR_HIP;
}

// ---- staging of pageable host buffers ------------------------------------
// A copy between pageable host memory and the device runs as a pipeline of
// 64 MiB chunks through two pinned buffers: host threads copy chunk i + 1
// into one pinned buffer while the DMA engine moves chunk i from (or to) the
// other.  Measured on MI355X for 2 GiB: 55 GB/s host to device (the pinned
// DMA alone reaches 57 GB/s) against 29 GB/s for a plain copy from pageable
// memory.  The pinned buffers are allocated once per process and reused
// (the host entries are synchronous; a mutex serialises their use).
constexpr size_t kStageChunk = 64ull << 20;
constexpr size_t kStageMin = 16ull << 20; // smaller copies go straight through hipMemcpy

struct Stager
{
    std::mutex mu;
    char *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
    int device = -1;
    bool ready = false;
};

Stager &stager()
{
    static Stager st;
    return st;
}

unsigned copy_threads()
{
    unsigned const hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(8u, hc ? hc : 1u));
}

// Copy n bytes with several host threads (pieces page-aligned).  A piece
// whose thread cannot be started is copied by the calling thread (nothing
// may throw out of the C entry points).
void parallel_copy(char *dst, const char *src, size_t n)
{
    unsigned const nt = copy_threads();
    if (nt <= 1 || n < (4ull << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    size_t per = (n + nt - 1) / nt;
    per = (per + 4095) & ~size_t(4095);
    std::thread threads[8];
    unsigned started = 0;
    for (unsigned t = 1; t < nt; ++t) {
        size_t const off = (size_t)t * per;
        if (off >= n) break;
        size_t const len = std::min(per, n - off);
        try {
            threads[started] = std::thread([=] { std::memcpy(dst + off, src + off, len); });
            ++started;
        } catch (...) {
            std::memcpy(dst + off, src + off, len);
        }
    }
    std::memcpy(dst, src, std::min(per, n));
    for (unsigned t = 0; t < started; ++t) threads[t].join();
}

// Set up the pinned buffers, events and stream on the current device (with
// the mutex held).  Returns false when any of them cannot be had; the caller
// then copies directly.
bool stager_ready(Stager &st)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (st.ready && st.device == dev) return true;
    if (st.ready) { // another device: the pinned buffers stay, the stream and events follow the device
        (void)hipStreamDestroy(st.stream);
        (void)hipEventDestroy(st.ev[0]);
        (void)hipEventDestroy(st.ev[1]);
        st.ready = false;
    }
    for (int b = 0; b < 2; ++b)
        if (!st.pin[b] && hipHostMalloc(reinterpret_cast<void **>(&st.pin[b]), kStageChunk, 0) != hipSuccess) {
            (void)hipGetLastError();
            st.pin[b] = nullptr;
            return false;
        }
    st.stream = nullptr;
    st.ev[0] = st.ev[1] = nullptr;
    if (hipStreamCreate(&st.stream) != hipSuccess ||
        hipEventCreateWithFlags(&st.ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&st.ev[1], hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (st.ev[1]) (void)hipEventDestroy(st.ev[1]);
        if (st.ev[0]) (void)hipEventDestroy(st.ev[0]);
        if (st.stream) (void)hipStreamDestroy(st.stream);
        st.stream = nullptr;
        st.ev[0] = st.ev[1] = nullptr;
        return false;
    }
    st.device = dev;
    st.ready = true;
    return true;
}

// Pageable host -> device.
hipError_t copy_h2d(void *dst, const void *src, size_t bytes)
{
    Stager &st = stager();
    std::lock_guard<std::mutex> lock(st.mu);
    if (bytes < kStageMin || !stager_ready(st)) return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    char *d = static_cast<char *>(dst);
    const char *h = static_cast<const char *>(src);
    bool used[2] = {false, false};
    hipError_t e = hipSuccess;
    for (size_t off = 0, i = 0; off < bytes && e == hipSuccess; off += kStageChunk, ++i) {
        size_t const len = std::min(kStageChunk, bytes - off);
        int const b = (int)(i & 1);
        if (used[b] && (e = hipEventSynchronize(st.ev[b])) != hipSuccess) break; // buffer b is free again
        parallel_copy(st.pin[b], h + off, len);
        e = hipMemcpyAsync(d + off, st.pin[b], len, hipMemcpyHostToDevice, st.stream);
        if (e == hipSuccess) e = hipEventRecord(st.ev[b], st.stream);
        used[b] = true;
    }
    hipError_t const e2 = hipStreamSynchronize(st.stream);
    return e != hipSuccess ? e : e2;
}

// Device -> pageable host.
hipError_t copy_d2h(void *dst, const void *src, size_t bytes)
{
    Stager &st = stager();
    std::lock_guard<std::mutex> lock(st.mu);
    if (bytes < kStageMin || !stager_ready(st)) return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    char *h = static_cast<char *>(dst);
    const char *d = static_cast<const char *>(src);
    size_t const nchunks = (bytes + kStageChunk - 1) / kStageChunk;
    auto issue = [&](size_t i) -> hipError_t {
        size_t const off = i * kStageChunk;
        size_t const len = std::min(kStageChunk, bytes - off);
        int const b = (int)(i & 1);
        hipError_t const e = hipMemcpyAsync(st.pin[b], d + off, len, hipMemcpyDeviceToHost, st.stream);
        return e == hipSuccess ? hipEventRecord(st.ev[b], st.stream) : e;
    };
    hipError_t e = issue(0);
    for (size_t i = 0; i < nchunks && e == hipSuccess; ++i) {
        int const b = (int)(i & 1);
        if ((e = hipEventSynchronize(st.ev[b])) != hipSuccess) break;
        if (i + 1 < nchunks && (e = issue(i + 1)) != hipSuccess) break; // the next chunk moves while this one is copied out
        size_t const off = i * kStageChunk;
        parallel_copy(h + off, st.pin[b], std::min(kStageChunk, bytes - off));
    }
    hipError_t const e2 = hipStreamSynchronize(st.stream);
    return e != hipSuccess ? e : e2;
}

// The last product this thread launched (ttm_last_kernel / ttm_last_splits).
struct LastRun
{
    LaunchInfo info;
    Fold fold;
    bool is_double = false;
    ttm_precision_t prec = TTM_PREC_DEFAULT;
    bool valid = false;
};
thread_local LastRun t_last;
thread_local std::string t_last_symbol;

void note_run(const LaunchInfo &info, const Fold &f, bool is_double, ttm_precision_t prec)
{
    t_last.info = info;
    t_last.fold = f;
    t_last.is_double = is_double;
    t_last.prec = prec;
    t_last.valid = true;
}

template <class T>
ttm_status_t launch_typed(const Fold &f, const T *a, const T *b, T *c, ttm_precision_t prec, hipStream_t s,
                          unsigned want_splits)
{
    LaunchInfo info;
    ttm_status_t st;
    if constexpr (std::is_same_v<T, float>) {
        st = launch_f32(f, a, b, c, prec, s, g_force, &info, want_splits);
    } else {
        if (prec != TTM_PREC_DEFAULT) {
            set_error(TTM_ERR_PRECISION);
            return TTM_ERR_PRECISION;
        }
        st = launch_f64(f, a, b, c, s, g_force, &info, want_splits);
    }
    if (st == TTM_SUCCESS) note_run(info, f, std::is_same_v<T, double>, prec);
    return st;
}

// Device pointer path shared by the typed entries.
template <class T>
ttm_status_t run_device(unsigned q, unsigned p, const T *a, const size_t *na, const size_t *wa,
                        const size_t *pia, const T *b, const size_t *nb, const size_t *pib, T *c,
                        const size_t *nc, const size_t *wc, ttm_precision_t prec, hipStream_t s)
{
    Fold f;
    ttm_status_t st = plan(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, f);
    if (st != TTM_SUCCESS) return st;
    st = launch_typed<T>(f, a, b, c, prec, s, 0);
    if (st == TTM_SUCCESS) clear_error();
    return st;
}

// Time the split candidates of a product on its own buffers (C is
// overwritten) and cache the fastest for its shape class.
template <class T>
ttm_status_t run_tune(unsigned q, unsigned p, const T *a, const size_t *na, const size_t *wa, const size_t *pia,
                      const T *b, const size_t *nb, const size_t *pib, T *c, const size_t *nc, const size_t *wc,
                      hipStream_t s, unsigned *splits_out)
{
    Fold f;
    ttm_status_t st = plan(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, f);
    if (st != TTM_SUCCESS) return st;
    bool const dbl = std::is_same_v<T, double>;
    // the family the planner picks for these buffers (no split yet)
    LaunchInfo info;
    st = launch_typed<T>(f, a, b, c, TTM_PREC_DEFAULT, s, 1);
    if (st != TTM_SUCCESS) return st;
    info = t_last.info;
    if (info.kernel != Kernel::gemm && info.kernel != Kernel::skinny) {
        if (splits_out) *splits_out = 1;
        clear_error();
        return TTM_SUCCESS;
    }
    // candidates: 1 and powers of two up to the family's limit, deduplicated
    // by the split count they actually give
    std::vector<unsigned> cand;
    for (unsigned want = 1; want <= 65535; want *= 2) {
        unsigned const got = launch_splits(info.kernel, f, dbl, want);
        if (std::find(cand.begin(), cand.end(), got) == cand.end()) cand.push_back(got);
        if (got < want) break;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return hip_fail(hipGetLastError());
    }
    unsigned best = 1;
    float best_ms = 0;
    for (unsigned want : cand) {
        float ms_min = 0;
        for (int rep = 0; rep < 4 && st == TTM_SUCCESS; ++rep) {
            (void)hipEventRecord(e0, s);
            st = launch_typed<T>(f, a, b, c, TTM_PREC_DEFAULT, s, want);
            (void)hipEventRecord(e1, s);
            if (st != TTM_SUCCESS) break;
            if (hipEventSynchronize(e1) != hipSuccess) {
                st = hip_fail(hipGetLastError());
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && (rep == 1 || ms < ms_min)) ms_min = ms; // rep 0 warms up
        }
        if (st != TTM_SUCCESS) break;
        if (want == cand.front() || ms_min < best_ms) {
            best = want;
            best_ms = ms_min;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != TTM_SUCCESS) return st;
    store_tuned_splits(info.kernel, f, dbl, best);
    if (splits_out) *splits_out = best;
    clear_error();
    return TTM_SUCCESS;
}

// Host (or mixed) pointer path: validates first, stages every buffer that is
// not device memory, runs on the null stream and waits for C.
template <class T>
ttm_status_t run_host(unsigned q, unsigned p, const T *a, const size_t *na, const size_t *wa,
                      const size_t *pia, const T *b, const size_t *nb, const size_t *pib, T *c,
                      const size_t *nc, const size_t *wc, ttm_precision_t prec)
{
    Fold f;
    ttm_status_t st = plan(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, f);
    if (st != TTM_SUCCESS) return st;
    if constexpr (std::is_same_v<T, double>) {
        if (prec != TTM_PREC_DEFAULT) {
            set_error(TTM_ERR_PRECISION);
            return TTM_ERR_PRECISION;
        }
    }

    size_t const bytes_a = f.numel_a() * sizeof(T);
    size_t const bytes_b = f.numel_b() * sizeof(T);
    size_t const bytes_c = f.numel_c() * sizeof(T);
    bool const a_dev = is_device_ptr(a), b_dev = is_device_ptr(b), c_dev = is_device_ptr(c);

    // Device copies of the host buffers come from the per-device pool, which
    // keeps its memory: repeated calls of one size allocate nothing new.
    T *da = nullptr, *db = nullptr, *dc = nullptr;
    auto release = [&]() {
        if (!a_dev && da) ws_release(da, nullptr);
        if (!b_dev && db) ws_release(db, nullptr);
        if (!c_dev && dc) ws_release(dc, nullptr);
    };
    auto alloc = [&](T **dst, size_t bytes) -> bool {
        void *ptr = nullptr;
        if (ws_acquire(bytes, nullptr, &ptr, "device copy of a host buffer") != TTM_SUCCESS) return false;
        *dst = static_cast<T *>(ptr);
        return true;
    };
    if ((!a_dev && !alloc(&da, bytes_a)) || (!b_dev && !alloc(&db, bytes_b)) ||
        (!c_dev && !alloc(&dc, bytes_c))) {
        release();
        return TTM_ERR_ALLOC;
    }
    if (a_dev) da = const_cast<T *>(a);
    if (b_dev) db = const_cast<T *>(b);
    if (c_dev) dc = c;

    // the pool hands out memory in null-stream order; the staging copies run
    // on their own stream
    hipError_t e = hipStreamSynchronize(nullptr);
    if (e == hipSuccess && !a_dev) e = copy_h2d(da, a, bytes_a);
    if (e == hipSuccess && !b_dev) e = copy_h2d(db, b, bytes_b);
    if (e != hipSuccess) {
        release();
        return hip_fail(e);
    }
    st = launch_typed<T>(f, da, db, dc, prec, nullptr, 0);
    if (st == TTM_SUCCESS) {
        e = hipDeviceSynchronize(); // the product ran on the null stream; the staging stream is separate
        if (e == hipSuccess && !c_dev) e = copy_d2h(c, dc, bytes_c);
        if (e != hipSuccess) st = hip_fail(e);
    }
    release();
    if (st == TTM_SUCCESS) clear_error();
    return st;
}

} // namespace

namespace ttm_amd
{
// The float32 device entry, for the allocation-free variant (ttm_noalloc.hip).
ttm_status_t sttm_device(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                         const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                         const size_t *nc, const size_t *wc, ttm_precision_t prec, hipStream_t s)
{
    return run_device<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, prec, s);
}
} // namespace ttm_amd

extern "C" {

ttm_status_t ttm_sttm(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                      const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                      const size_t *nc, const size_t *wc, ttm_stream_t stream)
{
    return run_device<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, TTM_PREC_DEFAULT,
                             static_cast<hipStream_t>(stream));
}

ttm_status_t ttm_dttm(unsigned q, unsigned p, const double *a, const size_t *na, const size_t *wa,
                      const size_t *pia, const double *b, const size_t *nb, const size_t *pib, double *c,
                      const size_t *nc, const size_t *wc, ttm_stream_t stream)
{
    return run_device<double>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, TTM_PREC_DEFAULT,
                              static_cast<hipStream_t>(stream));
}

ttm_status_t ttm_sttm_ex(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                         const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                         const size_t *nc, const size_t *wc, ttm_precision_t precision, ttm_stream_t stream)
{
    return run_device<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, precision,
                             static_cast<hipStream_t>(stream));
}

ttm_status_t ttm_sttm_host(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                           const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                           const size_t *nc, const size_t *wc)
{
    return run_host<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, TTM_PREC_DEFAULT);
}

ttm_status_t ttm_dttm_host(unsigned q, unsigned p, const double *a, const size_t *na, const size_t *wa,
                           const size_t *pia, const double *b, const size_t *nb, const size_t *pib,
                           double *c, const size_t *nc, const size_t *wc)
{
    return run_host<double>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, TTM_PREC_DEFAULT);
}

ttm_status_t ttm_sttm_host_ex(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                              const size_t *pia, const float *b, const size_t *nb, const size_t *pib,
                              float *c, const size_t *nc, const size_t *wc, ttm_precision_t precision)
{
    return run_host<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, precision);
}

const char *ttm_last_error(void) { return last_error(); }

const char *ttm_status_message(ttm_status_t status) { return status_text(status); }

int ttm_is_device_pointer(const void *ptr) { return is_device_ptr(ptr) ? 1 : 0; }

namespace
{
// Fold of a dense problem described by extents and layouts only.
bool fold_dense(unsigned q, unsigned p, const size_t *na, const size_t *pia, const size_t *nb, const size_t *pib,
                Fold &f)
{
    if (!na || !pia || !nb || !pib || p == 0 || p > 64 || q == 0 || q > p) return false;
    size_t nc[64] = {}, wa[64] = {}, wc[64] = {}; // modes a non-permutation pia skips stay 0 (rejected by plan)
    for (unsigned r = 0; r < p; ++r) nc[r] = na[r];
    nc[q - 1] = nb[0];
    size_t sa = 1, sc = 1;
    for (unsigned r = 0; r < p; ++r) {
        if (pia[r] == 0 || pia[r] > p) return false;
        wa[pia[r] - 1] = sa;
        wc[pia[r] - 1] = sc;
        sa *= na[pia[r] - 1];
        sc *= nc[pia[r] - 1];
    }
    static const char dummy = 0;
    return plan(q, p, &dummy, na, wa, pia, &dummy, nb, pib, &dummy, nc, wc, f) == TTM_SUCCESS;
}
thread_local std::string g_symbol;
} // namespace

unsigned ttm_plan_splits(unsigned q, unsigned p, int is_double, const size_t *na, const size_t *pia,
                         const size_t *nb, const size_t *pib, ttm_precision_t precision)
{
    Fold f;
    if (!fold_dense(q, p, na, pia, nb, pib, f)) return 0;
    Kernel const k = choose_kernel(f, is_double != 0, precision, true, g_force);
    return launch_splits(k, f, is_double != 0, 0);
}

const char *ttm_last_kernel(void)
{
    if (!t_last.valid) return "";
    t_last_symbol = kernel_symbol(t_last.info.kernel, t_last.fold, t_last.is_double, t_last.prec, t_last.info.aligned);
    return t_last_symbol.c_str();
}

unsigned ttm_last_splits(void) { return t_last.valid ? t_last.info.splits : 0; }

ttm_status_t ttm_stune(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                       const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                       const size_t *nc, const size_t *wc, ttm_stream_t stream, unsigned *splits)
{
    return run_tune<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, static_cast<hipStream_t>(stream), splits);
}

ttm_status_t ttm_dtune(unsigned q, unsigned p, const double *a, const size_t *na, const size_t *wa,
                       const size_t *pia, const double *b, const size_t *nb, const size_t *pib, double *c,
                       const size_t *nc, const size_t *wc, ttm_stream_t stream, unsigned *splits)
{
    return run_tune<double>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, static_cast<hipStream_t>(stream), splits);
}

void ttm_tune_clear(void) { clear_tuned_splits(); }

size_t ttm_workspace_bytes(void)
{
    hipMemPool_t const pool = workspace_pool();
    if (!pool) return 0;
    uint64_t v = 0;
    if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &v) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (size_t)v;
}

const char *ttm_plan_kernel(unsigned q, unsigned p, int is_double, const size_t *na, const size_t *pia,
                            const size_t *nb, const size_t *pib, ttm_precision_t precision)
{
    Fold f;
    if (!fold_dense(q, p, na, pia, nb, pib, f)) return "";
    return kernel_name(choose_kernel(f, is_double != 0, precision, true, g_force));
}

const char *ttm_plan_kernel_symbol(unsigned q, unsigned p, int is_double, const size_t *na, const size_t *pia,
                                   const size_t *nb, const size_t *pib, ttm_precision_t precision)
{
    Fold f;
    if (!fold_dense(q, p, na, pia, nb, pib, f)) return "";
    g_symbol = kernel_symbol(choose_kernel(f, is_double != 0, precision, true, g_force), f, is_double != 0, precision);
    return g_symbol.c_str();
}

const char *ttm_version(void) { return "ttm_amd 0.1.0 (gfx950)"; }

namespace
{
ttm_status_t hip_status(hipError_t e, ttm_status_t on_error)
{
    if (e == hipSuccess) {
        clear_error();
        return TTM_SUCCESS;
    }
    (void)hipGetLastError();
    set_error(on_error, hipGetErrorString(e));
    return on_error;
}
} // namespace

ttm_status_t ttm_device_alloc(size_t bytes, void **ptr)
{
    if (!ptr) {
        set_error(TTM_ERR_ARGUMENT, "output pointer is null");
        return TTM_ERR_ARGUMENT;
    }
    *ptr = nullptr;
    if (bytes == 0) {
        clear_error();
        return TTM_SUCCESS;
    }
    ttm_status_t const st = hip_status(hipMalloc(ptr, bytes), TTM_ERR_ALLOC);
    if (st != TTM_SUCCESS) *ptr = nullptr;
    return st;
}

ttm_status_t ttm_device_free(void *ptr)
{
    if (!ptr) {
        clear_error();
        return TTM_SUCCESS;
    }
    return hip_status(hipFree(ptr), TTM_ERR_HIP);
}

ttm_status_t ttm_device_copy(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) {
        clear_error();
        return TTM_SUCCESS;
    }
    if (!dst || !src) {
        set_error(TTM_ERR_ARGUMENT, "copy with a null pointer");
        return TTM_ERR_ARGUMENT;
    }
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDefault), TTM_ERR_HIP);
}

ttm_status_t ttm_device_zero(void *ptr, size_t bytes)
{
    if (bytes == 0) {
        clear_error();
        return TTM_SUCCESS;
    }
    if (!ptr) {
        set_error(TTM_ERR_ARGUMENT, "zero with a null pointer");
        return TTM_ERR_ARGUMENT;
    }
    return hip_status(hipMemset(ptr, 0, bytes), TTM_ERR_HIP);
}

ttm_status_t ttm_device_synchronize(void) { return hip_status(hipDeviceSynchronize(), TTM_ERR_HIP); }

// Testing hook (not part of the public header's contract): force a kernel
// family by name ("" restores the planner's choice).  Returns 0 when the
// name is known.
int ttm_force_kernel(const char *name)
{
    if (!name || !*name) {
        g_force = Kernel::none;
        return 0;
    }
    Kernel const k = kernel_from_name(name);
    if (k == Kernel::none) return -1;
    g_force = k;
    return 0;
}

} // extern "C"
