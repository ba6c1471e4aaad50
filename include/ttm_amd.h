/*
 * ttm_amd.h -- C ABI of the MI355X (gfx950) mode-q tensor-times-matrix product
 *
 *     C(i1..i(q-1), j, i(q+1)..ip) = sum_iq  A(i1..iq..ip) * B(j, iq)
 *
 * Every entry point takes plain pointers and sizes.  Extents, strides and
 * layouts are one-based-layout size_t arrays of length p exactly as in the
 * reference's low-level call; B is m x n_q (nb = {m, n_q}) and is
 * column-major when pib[0] == 1, row-major otherwise.  C has A's layout.
 *
 * The arrays A, B, C must be dense in their layout (the reference requires
 * contiguous tensors: README.md:40); the stride vectors are validated exactly
 * as the reference validates them.
 *
 * Two residencies:
 *   ttm_?ttm*       device pointers; the call is asynchronous and ordered on
 *                   `stream` (a hipStream_t, 0 = the null stream).
 *   ttm_?ttm_host*  host pointers; the call stages through device memory and
 *                   returns when C is complete (the reference's semantics).
 *
 * Errors: every entry returns a ttm_status_t.  Statuses 1..18 correspond
 * one-to-one to the reference's validation errors (include/tlib/ttm.h:65-90),
 * checked in the same order, and ttm_last_error() returns the reference's
 * message text for them (thread-local, valid until the next call on the
 * same thread).
 */
#ifndef TTM_AMD_H
#define TTM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t, passed opaquely so that the header needs no HIP include. */
typedef void *ttm_stream_t;

typedef enum ttm_status_t {
    TTM_SUCCESS = 0,
    /* include/tlib/ttm.h:65-90, same order and meaning */
    TTM_ERR_ORDER_ZERO = 1,     /* p == 0 */
    TTM_ERR_MODE = 2,           /* q == 0 || q > p */
    TTM_ERR_NULL_A = 3,
    TTM_ERR_NULL_B = 4,
    TTM_ERR_NULL_C = 5,
    TTM_ERR_NULL_NA = 6,
    TTM_ERR_NULL_NB = 7,
    TTM_ERR_NULL_NC = 8,
    TTM_ERR_NULL_WA = 9,
    TTM_ERR_NULL_WC = 10,
    TTM_ERR_NULL_PIA = 11,
    TTM_ERR_CONTRACTION_DIM = 12, /* na[q-1] != nb[1] */
    TTM_ERR_FREE_DIM = 13,        /* nc[q-1] != nb[0] */
    TTM_ERR_SHAPE_A = 14,
    TTM_ERR_SHAPE_C = 15,
    TTM_ERR_LAYOUT_A = 16,
    TTM_ERR_STRIDES_A = 17,
    TTM_ERR_STRIDES_C = 18,
    /* conditions the reference does not check */
    TTM_ERR_NULL_PIB = 19,        /* the reference dereferences pib unchecked */
    TTM_ERR_NOT_DENSE = 20,       /* valid but non-dense strides */
    TTM_ERR_SHAPE_C_MISMATCH = 21,/* nc differs from na outside mode q */
    TTM_ERR_PRECISION = 22,       /* precision not available for this element type */
    TTM_ERR_HIP = 23,             /* a HIP runtime call failed */
    TTM_ERR_ALLOC = 24,           /* device allocation failed */
    TTM_ERR_ARGUMENT = 25,        /* other invalid argument (batched descriptors) */
    TTM_ERR_POLICY = 26           /* a policy tag triple the reference does not define
                                     (include/tlib/detail/ttm.h:391-400) */
} ttm_status_t;

/* Arithmetic used for a float32 product.  TTM_PREC_DEFAULT is the plain
 * FP32/FP64 drop-in.  The others take float32 inputs and return float32
 * results, accumulating in FP32 on the matrix cores. */
typedef enum ttm_precision_t {
    TTM_PREC_DEFAULT = 0,   /* FP32 in, FP32 MFMA/FMA arithmetic (FP64 for double) */
    TTM_PREC_BF16X3 = 1,    /* three-term bfloat16 split, 6 products */
    TTM_PREC_BF16 = 2,      /* inputs rounded to bfloat16 */
    TTM_PREC_FP16 = 3,      /* inputs rounded to binary16 */
    TTM_PREC_FP8_E4M3 = 4,  /* OCP e4m3fn with per-slice power-of-two scaling */
    TTM_PREC_FP16X2 = 5     /* two-term binary16 split of A and B, each scaled by one power
                               of two (from its largest magnitude), 3 products */
} ttm_precision_t;

/* --- device pointers, stream ordered ---------------------------------- */
/* Replaces tlib::ttm::ttm(ep, sp, fp, q, p, a, na, wa, pia, b, nb, pib, c, nc, wc)
 * of include/tlib/ttm.h:54-93 for value_t = float / double (the policy tags
 * select CPU threading strategies and have no GPU counterpart). */
ttm_status_t ttm_sttm(unsigned q, unsigned p,
                      const float *a, const size_t *na, const size_t *wa, const size_t *pia,
                      const float *b, const size_t *nb, const size_t *pib,
                      float *c, const size_t *nc, const size_t *wc,
                      ttm_stream_t stream);

ttm_status_t ttm_dttm(unsigned q, unsigned p,
                      const double *a, const size_t *na, const size_t *wa, const size_t *pia,
                      const double *b, const size_t *nb, const size_t *pib,
                      double *c, const size_t *nc, const size_t *wc,
                      ttm_stream_t stream);

/* The float32 product with a selectable arithmetic (no reference
 * counterpart: the reference is float/double only, README.md:31,41). */
ttm_status_t ttm_sttm_ex(unsigned q, unsigned p,
                         const float *a, const size_t *na, const size_t *wa, const size_t *pia,
                         const float *b, const size_t *nb, const size_t *pib,
                         float *c, const size_t *nc, const size_t *wc,
                         ttm_precision_t precision, ttm_stream_t stream);

/* --- allocation-free variant -------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The same product as ttm_sttm_ex, on the same kernels,
 * with every device scratch buffer the call needs (the partial products of a
 * contraction split, the converted or quantized operands and the scales of
 * the reduced-precision kernels) taken from a fixed array of
 * TTM_NOALLOC_SCRATCH_BYTES bytes that the library holds in each GPU's
 * memory until the process ends, instead of from its memory pool; the call
 * allocates nothing.  A product whose scratch would exceed the array is
 * rejected with TTM_ERR_ALLOC before any launch (ttm_last_error() gives the
 * bytes needed); contraction splits are limited so that their partials fit.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other threads, other streams) race and must be serialised by the
 * caller. */
#define TTM_NOALLOC_SCRATCH_BYTES (64u << 20)
ttm_status_t ttm_sttm_ex_noalloc(unsigned q, unsigned p,
                                 const float *a, const size_t *na, const size_t *wa, const size_t *pia,
                                 const float *b, const size_t *nb, const size_t *pib,
                                 float *c, const size_t *nc, const size_t *wc,
                                 ttm_precision_t precision, ttm_stream_t stream);

/* --- host pointers, synchronous ---------------------------------------- */
/* Same contract as include/tlib/ttm.h:54-93 with caller-owned host buffers:
 * C is complete when the call returns. */
ttm_status_t ttm_sttm_host(unsigned q, unsigned p,
                           const float *a, const size_t *na, const size_t *wa, const size_t *pia,
                           const float *b, const size_t *nb, const size_t *pib,
                           float *c, const size_t *nc, const size_t *wc);

ttm_status_t ttm_dttm_host(unsigned q, unsigned p,
                           const double *a, const size_t *na, const size_t *wa, const size_t *pia,
                           const double *b, const size_t *nb, const size_t *pib,
                           double *c, const size_t *nc, const size_t *wc);

ttm_status_t ttm_sttm_host_ex(unsigned q, unsigned p,
                              const float *a, const size_t *na, const size_t *wa,
                              const size_t *pia, const float *b, const size_t *nb,
                              const size_t *pib, float *c, const size_t *nc, const size_t *wc,
                              ttm_precision_t precision);

/* --- contraction-split autotuner ------------------------------------------ */
/* No reference counterpart (the reference leaves blocking to its BLAS).
 * Products whose blocks cannot fill the GPU (few columns, or a long
 * contraction with a small C: the GEMV-like cases, p = 1 included) split the
 * contraction over the chip and add the partial products in a second,
 * deterministic pass.  ttm_?tune times the split counts for the problem on
 * its own device buffers (C is overwritten with the product) and keeps the
 * fastest for the problem's shape class -- kernel family, element type, view
 * (q fastest in the layout or not), B's orientation and the power-of-two
 * classes of m, n_q and the column count -- for every later call of the
 * process (and in the file named by TTM_AMD_SPLIT_CACHE when set).  *splits
 * receives the chosen count (1 for kernels that do not split). */
ttm_status_t ttm_stune(unsigned q, unsigned p,
                       const float *a, const size_t *na, const size_t *wa, const size_t *pia,
                       const float *b, const size_t *nb, const size_t *pib,
                       float *c, const size_t *nc, const size_t *wc,
                       ttm_stream_t stream, unsigned *splits);
ttm_status_t ttm_dtune(unsigned q, unsigned p,
                       const double *a, const size_t *na, const size_t *wa, const size_t *pia,
                       const double *b, const size_t *nb, const size_t *pib,
                       double *c, const size_t *nc, const size_t *wc,
                       ttm_stream_t stream, unsigned *splits);
/* Forget every tuned split of this process (the file is left alone). */
void ttm_tune_clear(void);

/* --- host executor (no GPU) ---------------------------------------------- */
/* The reference's policy tag triples (include/tlib/detail/tags.h:22-69), for
 * the host executor: how the work is cut (slices of the fastest mode's extent
 * or whole subtensors below q) and scheduled over threads (sequential, one
 * contiguous range per thread, chunks or single items on demand, inside each
 * product, both, one flat batch, or the reference's `combined` choice). */
typedef enum ttm_cpu_parallel_t {
    TTM_CPU_SEQUENTIAL = 0,          /* parallel_policy::sequential */
    TTM_CPU_PARALLEL_BLAS = 1,       /* parallel_policy::parallel_blas */
    TTM_CPU_PARALLEL_LOOP = 2,       /* parallel_policy::parallel_loop */
    TTM_CPU_PARALLEL_TASKLOOP = 3,   /* parallel_policy::parallel_taskloop */
    TTM_CPU_PARALLEL_TASK = 4,       /* parallel_policy::parallel_task */
    TTM_CPU_PARALLEL_LOOP_BLAS = 5,  /* parallel_policy::parallel_loop_blas */
    TTM_CPU_BATCHED_GEMM = 6,        /* parallel_policy::batched_gemm */
    TTM_CPU_COMBINED = 7             /* parallel_policy::combined */
} ttm_cpu_parallel_t;
typedef enum ttm_cpu_slicing_t {
    TTM_CPU_SLICE = 0,               /* slicing_policy::slice */
    TTM_CPU_SUBTENSOR = 1,           /* slicing_policy::subtensor */
    TTM_CPU_SLICING_COMBINED = 2     /* slicing_policy::combined */
} ttm_cpu_slicing_t;
typedef enum ttm_cpu_fusion_t {
    TTM_CPU_FUSION_NONE = 0,         /* fusion_policy::none */
    TTM_CPU_FUSION_OUTER = 1,        /* fusion_policy::outer */
    TTM_CPU_FUSION_ALL = 2           /* fusion_policy::all */
} ttm_cpu_fusion_t;

/* Replaces the reference's host executors, tlib::ttm::ttm(ep, sp, fp, ...)
 * for each of the 19 tag triples it defines (include/tlib/detail/ttm.h:405-1474),
 * for callers without a GPU: host pointers, synchronous, `threads` worker
 * threads (0 = the hardware's), `ratio` the share of the threads on the outer
 * loop for parallel_loop_blas (ignored otherwise; 0 = half).  Same validation
 * and statuses as the GPU entries; an undefined triple returns
 * TTM_ERR_POLICY.  Every triple and thread count gives the same bits: each
 * element of C is one in-order sum over the contraction. */
ttm_status_t ttm_sttm_cpu(unsigned q, unsigned p,
                          const float *a, const size_t *na, const size_t *wa, const size_t *pia,
                          const float *b, const size_t *nb, const size_t *pib,
                          float *c, const size_t *nc, const size_t *wc,
                          ttm_cpu_parallel_t parallel, ttm_cpu_slicing_t slicing, ttm_cpu_fusion_t fusion,
                          unsigned threads, double ratio);
ttm_status_t ttm_dttm_cpu(unsigned q, unsigned p,
                          const double *a, const size_t *na, const size_t *wa, const size_t *pia,
                          const double *b, const size_t *nb, const size_t *pib,
                          double *c, const size_t *nc, const size_t *wc,
                          ttm_cpu_parallel_t parallel, ttm_cpu_slicing_t slicing, ttm_cpu_fusion_t fusion,
                          unsigned threads, double ratio);

/* --- batched multi-TTM chains (Tucker-style) ------------------------------ */
/* No single reference counterpart: generalises ttmpy.ttms (the sequential
 * multi-TTM chain of ttmpy/src/wrapped_ttm.cpp:81-222) to many tensors of
 * uneven extents in one plan.  Tensor t of a plan computes
 *     C = A x_{k1} U_{k1} x_{k2} U_{k2} ...   over every mode k with u[k-1] != NULL
 * where U_k is the row-major r[k-1] x n[k-1] factor of mode k; modes with
 * u[k-1] == NULL keep their extent (their stage is a copy into the next
 * stage's layout, not a product).  A and C are dense in layout pi, in device
 * memory.  Modes are contracted in layout order; every stage reads rows that
 * are contiguous along the contracted mode and writes that mode slowest, so the
 * result lands back in layout pi without any transposition. */
#define TTM_CHAIN_MAX_ORDER 8
typedef struct ttm_chain_desc_t {
    unsigned p;                         /* order, 1..TTM_CHAIN_MAX_ORDER */
    size_t n[TTM_CHAIN_MAX_ORDER];      /* extents of A */
    size_t pi[TTM_CHAIN_MAX_ORDER];     /* one-based layout of A and C, fastest first */
    const void *a;                      /* A (float or double, as the plan) */
    const void *u[TTM_CHAIN_MAX_ORDER]; /* factor of mode k+1, row-major r[k] x n[k], or NULL */
    size_t r[TTM_CHAIN_MAX_ORDER];      /* rows of u[k] = extent of mode k+1 in C */
    void *c;                            /* C: extents r[k] where u[k] != NULL, else n[k] */
} ttm_chain_desc_t;

typedef struct ttm_chain_plan_s *ttm_chain_plan_t;

/* Validate `count` descriptors and prepare a plan (allocates device memory for
 * descriptors, tile maps and intermediates).  precision: TTM_PREC_DEFAULT
 * (float or double as is_double says) or, for float, TTM_PREC_BF16X3 or
 * TTM_PREC_FP16X2 (scaled per tile and slice of 32 contracted indices). */
ttm_status_t ttm_chain_plan_create(size_t count, const ttm_chain_desc_t *descs, int is_double,
                                   ttm_precision_t precision, ttm_chain_plan_t *plan);
/* Run every chain of the plan: one persistent grouped launch per stage,
 * ordered on `stream`; allocates nothing. */
ttm_status_t ttm_chain_run(ttm_chain_plan_t plan, ttm_stream_t stream);
/* The same as one hipGraph: captured on the first call, replayed afterwards. */
ttm_status_t ttm_chain_run_graph(ttm_chain_plan_t plan, ttm_stream_t stream);
/* Time each compiled tile shape per stage on the plan's own buffers, keep the
 * fastest (also for later plans of the same shape class in this process, and
 * in the file named by TTM_AMD_TUNE_CACHE when set).  Drops a captured graph.
 * Stages whose contractions are 4..256 long and whose ranks are at most 64
 * run the short-row kernels by default (float32, BF16X3, FP16X2, float64)
 * (TTM_AMD_CHAIN_ROWS=0 at plan creation keeps them on the tiles); the
 * plain-precision tuner times the short-row kernel next to the tile shapes.
 * Small stages run their rank classes up to 64 as one launch
 * (TTM_AMD_CHAIN_MERGE=0 at plan creation keeps one launch per class). */
ttm_status_t ttm_chain_tune(ttm_chain_plan_t plan, ttm_stream_t stream);
/* Launches of the plan (each stage of the chains split by row class of the
 * ranks, in execution order) and the tile shape (kernel instance) of launch `s`. */
unsigned ttm_chain_stages(ttm_chain_plan_t plan);
const char *ttm_chain_stage_kernel(ttm_chain_plan_t plan, unsigned s);
void ttm_chain_plan_destroy(ttm_chain_plan_t plan);

/* --- device memory for callers without a HIP toolchain ------------------ */
/* No reference counterpart (the reference's tensor is host memory only,
 * include/tlib/detail/tensor.h:59-121): these back the device-resident
 * tlib::ttm::device_tensor of include/tlib/device.h, so that C++ callers keep
 * tensors in HBM across calls without including HIP headers.  All are
 * synchronous on the current device. */
ttm_status_t ttm_device_alloc(size_t bytes, void **ptr);   /* *ptr = NULL when bytes == 0 */
ttm_status_t ttm_device_free(void *ptr);                   /* NULL is a no-op */
/* Copy `bytes` between any host / device pointers (direction from the
 * pointers themselves). */
ttm_status_t ttm_device_copy(void *dst, const void *src, size_t bytes);
ttm_status_t ttm_device_zero(void *ptr, size_t bytes);
/* Wait for all work submitted to the current device. */
ttm_status_t ttm_device_synchronize(void);

/* --- diagnostics -------------------------------------------------------- */
/* Message of the last failing call on this thread ("" after a success). */
const char *ttm_last_error(void);
/* Fixed message for a status (the reference's text for 1..18). */
const char *ttm_status_message(ttm_status_t status);
/* 1 if `ptr` is device (or managed) memory visible to the current device. */
int ttm_is_device_pointer(const void *ptr);
/* Name of the kernel the planner picks for a problem, without running it
 * (for tests and the bench); returns a static string or "" on invalid input. */
const char *ttm_plan_kernel(unsigned q, unsigned p, int is_double, const size_t *na,
                            const size_t *pia, const size_t *nb, const size_t *pib,
                            ttm_precision_t precision);
/* Template instance of that kernel as profilers name it, e.g.
 * "k_gemm<float, 4, 4, 2, 16, false, true>" (thread-local string). */
const char *ttm_plan_kernel_symbol(unsigned q, unsigned p, int is_double, const size_t *na,
                                   const size_t *pia, const size_t *nb, const size_t *pib,
                                   ttm_precision_t precision);
/* Contraction splits the planner uses for a dense problem with 16-byte
 * aligned buffers (1 = none; 0 on invalid input). */
unsigned ttm_plan_splits(unsigned q, unsigned p, int is_double, const size_t *na, const size_t *pia,
                         const size_t *nb, const size_t *pib, ttm_precision_t precision);
/* Template instance of the kernel that the last successful product call of
 * this thread ran ("" before any), and its contraction splits (0 before any). */
const char *ttm_last_kernel(void);
unsigned ttm_last_splits(void);
/* Device memory held by the current device's workspace pool (bytes). */
size_t ttm_workspace_bytes(void);
/* Library version string. */
const char *ttm_version(void);

/* --- testing hook --------------------------------------------------------- */
/* Force a kernel family by name for every later call of this process
 * ("generic", "nn_stream", "tn_stream", ...; "" or NULL restores the
 * planner's choice).  A forced family that cannot run a problem is ignored
 * for it.  Returns 0 for a known name, -1 otherwise.  Not thread-safe. */
int ttm_force_kernel(const char *name);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* TTM_AMD_H */
