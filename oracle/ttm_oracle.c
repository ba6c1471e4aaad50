/*
 * ttm_oracle.c -- CPU oracle for the mode-q tensor-times-matrix product.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only the test-suite, the smoke check and the
 * cpu_baseline leg of bench.py use it, and only as the reference result / the CPU
 * baseline, never as the thing being measured on the GPU.
 *
 * It is a plain-C restatement of the reference algorithm (bassoy/ttm,
 * TLIB(TTM)); every function names the reference lines it follows.  It has
 * three layers:
 *   1. metadata   : shape / layout / strides / inverse_mode / is_case / product
 *   2. dispatch   : validation + the `combined` executor with its three
 *                   case-8 strategies and the direct cases 1..7, each issuing
 *                   the same GEMM/GEMV calls (same M,N,K,lda,ldb,ldc) as the
 *                   reference.  The GEMM is either a naive loop (value-type
 *                   accumulation, deterministic) or a CBLAS found at run time
 *                   by dlopen (for the timed CPU baseline).
 *   3. fp64 brute force : the mathematical definition over arbitrary strides
 *                   with long-double accumulation -- the "FP64 oracle" used to
 *                   measure the error of every precision.
 *
 * How each part is pinned (see tests/test_oracle.py):
 *   metadata   -- known answers of test/src/gtest_tlib_{shape,layout,strides}.cpp
 *                 and the real reference headers compiled under oracle/_ref.
 *   dispatch   -- closed-form triangular-number answers of
 *                 test/src/gtest_tlib_ttm.cpp:27-134 (all p! layouts, both B
 *                 layouts, orders 2..4) extended with a non-trivial B;
 *                 numpy.einsum in float64; ttmpy/tests/test.py vectors.
 *   brute force -- numpy.einsum and the same closed forms.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* 1. metadata                                                               */
/* ------------------------------------------------------------------------ */

/* include/tlib/detail/shape.h:31-35 */
int or_is_valid_shape(const size_t *n, size_t p)
{
    if (p == 0) return 0;
    for (size_t r = 0; r < p; ++r)
        if (n[r] == 0) return 0;
    return 1;
}

/* include/tlib/detail/shape.h:39-47 */
int or_is_scalar(const size_t *n, size_t p)
{
    if (!or_is_valid_shape(n, p)) return 0;
    for (size_t r = 0; r < p; ++r)
        if (n[r] != 1) return 0;
    return 1;
}

/* include/tlib/detail/shape.h:50-67 */
int or_is_vector(const size_t *n, size_t p)
{
    if (!or_is_valid_shape(n, p)) return 0;
    if (p == 1) return n[0] > 1;
    int any_gt1 = (n[0] > 1) || (n[1] > 1);
    int any_eq1 = (n[0] == 1) || (n[1] == 1);
    for (size_t r = 2; r < p; ++r)
        if (n[r] != 1) return 0;
    return any_gt1 && any_eq1;
}

/* include/tlib/detail/shape.h:69-79 */
int or_is_matrix(const size_t *n, size_t p)
{
    if (!or_is_valid_shape(n, p)) return 0;
    if (p < 2) return 0;
    if (!(n[0] > 1 && n[1] > 1)) return 0;
    for (size_t r = 2; r < p; ++r)
        if (n[r] != 1) return 0;
    return 1;
}

/* include/tlib/detail/shape.h:82-90 */
int or_is_tensor(const size_t *n, size_t p)
{
    if (!or_is_valid_shape(n, p)) return 0;
    if (p < 3) return 0;
    for (size_t r = 2; r < p; ++r)
        if (n[r] > 1) return 1;
    return 0;
}

/* include/tlib/detail/layout.h:29-55: a permutation of 1..p */
int or_is_valid_layout(const size_t *pi, size_t p)
{
    if (p == 0) return 0;
    for (size_t r = 0; r < p; ++r) {
        if (pi[r] == 0 || pi[r] > p) return 0;
        for (size_t s = r + 1; s < p; ++s)
            if (pi[s] == pi[r]) return 0;
    }
    return 1;
}

/* include/tlib/detail/layout.h:57-78: {k,k-1,..,1,k+1,..,p}; k==0 -> last order */
int or_compute_k_order_layout(size_t *pi, size_t p, size_t k)
{
    if (p == 0) return -1;
    if (k == 0) k = p;
    size_t m = k > p ? p : k;
    size_t r = 0;
    for (size_t v = m; r < m; ++r, --v) pi[r] = v;
    for (size_t v = m + 1; r < p; ++r, ++v) pi[r] = v;
    return 0;
}

/* include/tlib/detail/layout.h:86-102 */
int or_is_first_order(const size_t *pi, size_t p)
{
    for (size_t r = 0; r < p; ++r)
        if (pi[r] != r + 1) return 0;
    return 1;
}

int or_is_last_order(const size_t *pi, size_t p)
{
    for (size_t r = 0; r < p; ++r)
        if (pi[r] != p - r) return 0;
    return 1;
}

/* include/tlib/detail/layout.h:120-144: one-based position of `mode` in pi,
 * 0 when the reference would throw. */
size_t or_inverse_mode(const size_t *pi, size_t p, size_t mode)
{
    if (!or_is_valid_layout(pi, p)) return 0;
    if (mode == 0 || mode > p) return 0;
    for (size_t r = 0; r < p; ++r)
        if (pi[r] == mode) return r + 1;
    return 0;
}

/* include/tlib/detail/strides.h:31-57.  Returns -1 where the reference
 * throws.  Vectors and scalars get all-ones strides (strides.h:44-48). */
int or_compute_strides(const size_t *n, const size_t *pi, size_t p, size_t *w)
{
    if (!or_is_valid_shape(n, p)) return -1;
    if (!or_is_valid_layout(pi, p)) return -1;
    for (size_t r = 0; r < p; ++r) w[r] = 1;
    if (or_is_vector(n, p) || or_is_scalar(n, p)) return 0;
    if (!or_is_matrix(n, p) && !or_is_tensor(n, p)) return 0;
    for (size_t r = 1; r < p; ++r) {
        size_t cur = pi[r] - 1, prev = pi[r - 1] - 1;
        w[cur] = w[prev] * n[prev];
    }
    return 0;
}

/* include/tlib/detail/strides.h:76-101: strides non-decreasing along pi.
 * Returns -1 where the reference throws (invalid layout). */
int or_is_valid_strides(const size_t *pi, size_t p, const size_t *w)
{
    if (!or_is_valid_layout(pi, p)) return -1;
    if (p == 1 && w[0] == 1) return 1;
    for (size_t r = 0; r + 1 < p; ++r)
        if (w[pi[r] - 1] > w[pi[r + 1] - 1]) return 0;
    return 1;
}

/* include/tlib/detail/cases.h:25-37 */
int or_is_case(unsigned case_nr, unsigned p, size_t q, const size_t *pi)
{
    switch (case_nr) {
    case 1: return p == 1;
    case 2: return p == 2 && q == 1 && pi[0] == 1;
    case 3: return p == 2 && q == 2 && pi[0] == 1;
    case 4: return p == 2 && q == 1 && pi[0] == 2;
    case 5: return p == 2 && q == 2 && pi[0] == 2;
    case 6: return p >= 3 && pi[0] == q;
    case 7: return p >= 3 && pi[p - 1] == q;
    case 8: return p >= 3 && !(pi[0] == q || pi[p - 1] == q);
    default: return 0;
    }
}

/* include/tlib/detail/ttm.h:181-191: product of n[pi[r]] for r in [start,finish) (one-based) */
size_t or_product(const size_t *n, const size_t *pi, unsigned start, unsigned finish)
{
    size_t nn = 1;
    for (unsigned r = start - 1; r + 1 < finish; ++r) nn *= n[pi[r] - 1];
    return nn;
}

/* include/tlib/detail/ttm.h:60-92: sockets x cores-per-socket from lscpu,
 * capped by the logical CPU count. */
static unsigned read_lscpu_field(const char *cmd)
{
    unsigned v = 0;
    char out[16];
    FILE *fp = popen(cmd, "r");
    if (fp) {
        if (fgets(out, sizeof out, fp)) v = (unsigned)atoi(out);
        pclose(fp);
    }
    return v;
}

static unsigned g_cores; /* computed once, like the reference's static `cores` (ttm.h:96) */

unsigned or_cores(void)
{
    if (g_cores) return g_cores;
    unsigned sockets = read_lscpu_field("lscpu | grep 'Socket' | cut -d':' -f2 | tr -d ' '");
    unsigned per_socket = read_lscpu_field("lscpu | grep 'socket' | cut -d':' -f2 | tr -d ' '");
    unsigned logical = (unsigned)sysconf(_SC_NPROCESSORS_ONLN);
    unsigned c = sockets * per_socket;
    if (c == 0 || c > logical) c = logical;
    g_cores = c;
    return c;
}

/* include/tlib/ttm.h:65-90: validation in the reference's order.  Returns 0
 * or the one-based index of the failing check (same numbering as
 * ttm_status_t in the product header). */
int or_validate(unsigned q, unsigned p, const void *a, const size_t *na, const size_t *wa,
                const size_t *pia, const void *b, const size_t *nb, const void *c,
                const size_t *nc, const size_t *wc)
{
    if (p == 0) return 1;
    if (q == 0 || q > p) return 2;
    if (!a) return 3;
    if (!b) return 4;
    if (!c) return 5;
    if (!na) return 6;
    if (!nb) return 7;
    if (!nc) return 8;
    if (!wa) return 9;
    if (!wc) return 10;
    if (!pia) return 11;
    if (na[q - 1] != nb[1]) return 12;
    if (nc[q - 1] != nb[0]) return 13;
    if (!or_is_valid_shape(na, p)) return 14;
    if (!or_is_valid_shape(nc, p)) return 15;
    if (!or_is_valid_layout(pia, p)) return 16;
    if (or_is_valid_strides(pia, p, wa) != 1) return 17;
    if (or_is_valid_strides(pia, p, wc) != 1) return 18;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 2. GEMM back ends + the reference dispatch                                */
/* ------------------------------------------------------------------------ */

/* Standard CBLAS enumerator values (the public CBLAS interface). */
enum { ROW = 101, COL = 102, NOTR = 111, TR = 112 };

typedef void (*sgemm_lp64_t)(int, int, int, int, int, int, float, const float *, int,
                             const float *, int, float, float *, int);
typedef void (*dgemm_lp64_t)(int, int, int, int, int, int, double, const double *, int,
                             const double *, int, double, double *, int);
typedef void (*sgemm_ilp64_t)(int, int, int, int64_t, int64_t, int64_t, float, const float *,
                              int64_t, const float *, int64_t, float, float *, int64_t);
typedef void (*dgemm_ilp64_t)(int, int, int, int64_t, int64_t, int64_t, double, const double *,
                              int64_t, const double *, int64_t, double, double *, int64_t);
typedef void (*set_threads_int_t)(int);
typedef void (*set_threads_i64_t)(int64_t);

static struct {
    void *handle;
    int ilp64;
    void *sgemm, *dgemm, *set_threads;
    char name[512];
} g_blas;

/* Bind a CBLAS by path.  kind: 0 = plain LP64 names (cblas_sgemm,
 * openblas_set_num_threads / MKL_Set_Num_Threads), 1 = the ILP64
 * scipy-openblas naming (scipy_cblas_sgemm64_, scipy_openblas_set_num_threads64_).
 * Returns 0 on success. */
int or_blas_load(const char *path, int kind)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    void *sg, *dg, *st;
    if (kind == 1) {
        sg = dlsym(h, "scipy_cblas_sgemm64_");
        dg = dlsym(h, "scipy_cblas_dgemm64_");
        st = dlsym(h, "scipy_openblas_set_num_threads64_");
    } else {
        sg = dlsym(h, "cblas_sgemm");
        dg = dlsym(h, "cblas_dgemm");
        st = dlsym(h, "openblas_set_num_threads");
        if (!st) st = dlsym(h, "MKL_Set_Num_Threads");
    }
    if (!sg || !dg) { dlclose(h); return -2; }
    g_blas.handle = h;
    g_blas.ilp64 = kind == 1;
    g_blas.sgemm = sg;
    g_blas.dgemm = dg;
    g_blas.set_threads = st;
    snprintf(g_blas.name, sizeof g_blas.name, "%s", path);
    return 0;
}

const char *or_blas_name(void) { return g_blas.handle ? g_blas.name : "naive"; }

static void blas_threads(int n)
{
    if (!g_blas.handle || !g_blas.set_threads) return;
    if (g_blas.ilp64) ((set_threads_i64_t)g_blas.set_threads)((int64_t)n);
    else ((set_threads_int_t)g_blas.set_threads)(n);
}

/* Naive GEMM with CBLAS argument semantics, alpha=1, beta=0, accumulating in
 * the value type in index order (what a straightforward BLAS does). */
#define DEFINE_NAIVE_GEMM(NAME, T)                                                        \
    static void NAME(int layout, int ta, int tb, size_t M, size_t N, size_t K,            \
                     const T *A, size_t lda, const T *B, size_t ldb, T *C, size_t ldc)   \
    {                                                                                     \
        for (size_t i = 0; i < M; ++i)                                                    \
            for (size_t j = 0; j < N; ++j) {                                              \
                T s = 0;                                                                  \
                for (size_t l = 0; l < K; ++l) {                                          \
                    T av, bv;                                                             \
                    if (layout == ROW) {                                                  \
                        av = ta == NOTR ? A[i * lda + l] : A[l * lda + i];                \
                        bv = tb == NOTR ? B[l * ldb + j] : B[j * ldb + l];                \
                    } else {                                                              \
                        av = ta == NOTR ? A[i + l * lda] : A[l + i * lda];                \
                        bv = tb == NOTR ? B[l + j * ldb] : B[j + l * ldb];                \
                    }                                                                     \
                    s += av * bv;                                                         \
                }                                                                         \
                if (layout == ROW) C[i * ldc + j] = s;                                    \
                else C[i + j * ldc] = s;                                                  \
            }                                                                             \
    }
DEFINE_NAIVE_GEMM(naive_sgemm, float)
DEFINE_NAIVE_GEMM(naive_dgemm, double)

static int g_use_blas = 0; /* set per call */

static void gemm_f(int layout, int ta, int tb, size_t M, size_t N, size_t K, const float *A,
                   size_t lda, const float *B, size_t ldb, float *C, size_t ldc)
{
    if (g_use_blas && g_blas.handle) {
        if (g_blas.ilp64)
            ((sgemm_ilp64_t)g_blas.sgemm)(layout, ta, tb, (int64_t)M, (int64_t)N, (int64_t)K, 1.0f,
                                          A, (int64_t)lda, B, (int64_t)ldb, 0.0f, C, (int64_t)ldc);
        else
            ((sgemm_lp64_t)g_blas.sgemm)(layout, ta, tb, (int)M, (int)N, (int)K, 1.0f, A, (int)lda,
                                         B, (int)ldb, 0.0f, C, (int)ldc);
        return;
    }
    naive_sgemm(layout, ta, tb, M, N, K, A, lda, B, ldb, C, ldc);
}

static void gemm_d(int layout, int ta, int tb, size_t M, size_t N, size_t K, const double *A,
                   size_t lda, const double *B, size_t ldb, double *C, size_t ldc)
{
    if (g_use_blas && g_blas.handle) {
        if (g_blas.ilp64)
            ((dgemm_ilp64_t)g_blas.dgemm)(layout, ta, tb, (int64_t)M, (int64_t)N, (int64_t)K, 1.0,
                                          A, (int64_t)lda, B, (int64_t)ldb, 0.0, C, (int64_t)ldc);
        else
            ((dgemm_lp64_t)g_blas.dgemm)(layout, ta, tb, (int)M, (int)N, (int)K, 1.0, A, (int)lda,
                                         B, (int)ldb, 0.0, C, (int)ldc);
        return;
    }
    naive_dgemm(layout, ta, tb, M, N, K, A, lda, B, ldb, C, ldc);
}

/* A GEMV y = A x restated as a one-column GEMM: include/tlib/detail/mtm.h:117-141
 * (NoTrans, incx = incy = 1). */
#define GEMV(GEMM, layout, A, x, y, m, n, lda)                                             \
    do {                                                                                  \
        if ((layout) == ROW) GEMM(ROW, NOTR, NOTR, m, 1, n, A, lda, x, 1, y, 1);          \
        else GEMM(COL, NOTR, NOTR, m, 1, n, A, lda, x, n, y, m);                          \
    } while (0)

/* The typed body is written once as a macro over (T, GEMM). */
#define DEFINE_TTM(SUFFIX, T, GEMM)                                                              \
/* include/tlib/detail/mtm.h:145-182 (B row-major) */                                           \
static void mtm_rm_##SUFFIX(unsigned q, unsigned p, const T *a, const size_t *na,               \
                            const size_t *pia, const T *b, const size_t *nb, T *c)               \
{                                                                                                \
    size_t m = na[0], n = p > 1 ? na[1] : 1, nq = na[q - 1], u = nb[0];                         \
    size_t nn = 1;                                                                               \
    for (unsigned r = 0; r < p; ++r) nn *= na[r];                                                \
    nn /= nq;                                                                                    \
    if (or_is_case(1, p, q, pia)) GEMV(GEMM, ROW, b, a, c, u, m, m);                             \
    else if (or_is_case(2, p, q, pia)) GEMM(ROW, NOTR, TR, n, u, m, a, m, b, m, c, u);           \
    else if (or_is_case(3, p, q, pia)) GEMM(ROW, NOTR, NOTR, u, m, n, b, n, a, m, c, m);         \
    else if (or_is_case(4, p, q, pia)) GEMM(ROW, NOTR, NOTR, u, n, m, b, m, a, n, c, n);         \
    else if (or_is_case(5, p, q, pia)) GEMM(ROW, NOTR, TR, m, u, n, a, n, b, n, c, u);           \
    else if (or_is_case(6, p, q, pia)) GEMM(ROW, NOTR, TR, nn, u, nq, a, nq, b, nq, c, u);       \
    else if (or_is_case(7, p, q, pia)) GEMM(ROW, NOTR, NOTR, u, nn, nq, b, nq, a, nn, c, nn);    \
}                                                                                                \
/* include/tlib/detail/mtm.h:186-223 (B column-major) */                                        \
static void mtm_cm_##SUFFIX(unsigned q, unsigned p, const T *a, const size_t *na,               \
                            const size_t *pia, const T *b, const size_t *nb, T *c)               \
{                                                                                                \
    size_t m = na[0], n = p > 1 ? na[1] : 1, nq = na[q - 1], u = nb[0];                         \
    size_t nn = 1;                                                                               \
    for (unsigned r = 0; r < p; ++r) nn *= na[r];                                                \
    nn /= nq;                                                                                    \
    if (or_is_case(1, p, q, pia)) GEMV(GEMM, COL, b, a, c, u, m, u);                             \
    else if (or_is_case(2, p, q, pia)) GEMM(COL, NOTR, NOTR, u, n, m, b, u, a, m, c, u);         \
    else if (or_is_case(3, p, q, pia)) GEMM(COL, NOTR, TR, m, u, n, a, m, b, u, c, m);           \
    else if (or_is_case(4, p, q, pia)) GEMM(COL, NOTR, TR, n, u, m, a, n, b, u, c, n);           \
    else if (or_is_case(5, p, q, pia)) GEMM(COL, NOTR, NOTR, u, m, n, b, u, a, n, c, u);         \
    else if (or_is_case(6, p, q, pia)) GEMM(COL, NOTR, NOTR, u, nn, nq, b, u, a, nq, c, u);      \
    else if (or_is_case(7, p, q, pia)) GEMM(COL, NOTR, TR, nn, u, nq, a, nn, b, u, c, nn);       \
}                                                                                                \
/* include/tlib/detail/ttm.h:295-318: recursion over the modes above qh */                      \
static void subtensor_loops_##SUFFIX(int is_cm, unsigned r, unsigned qh, const T *a,            \
                                     const size_t *na, const size_t *wa, const size_t *pia,     \
                                     const T *b, T *c, const size_t *wc, size_t nnq, size_t m,  \
                                     size_t nq)                                                  \
{                                                                                                \
    if (r > 1) {                                                                                 \
        if (r <= qh) {                                                                           \
            subtensor_loops_##SUFFIX(is_cm, r - 1, qh, a, na, wa, pia, b, c, wc, nnq, m, nq);   \
        } else {                                                                                 \
            size_t pr = pia[r - 1] - 1;                                                          \
            for (size_t i = 0; i < na[pr]; ++i, a += wa[pr], c += wc[pr])                        \
                subtensor_loops_##SUFFIX(is_cm, r - 1, qh, a, na, wa, pia, b, c, wc, nnq, m, nq);\
        }                                                                                        \
    } else if (is_cm) {                                                                          \
        GEMM(COL, NOTR, TR, nnq, m, nq, a, nnq, b, m, c, nnq);                                  \
    } else {                                                                                     \
        GEMM(ROW, NOTR, NOTR, m, nnq, nq, b, nq, a, nnq, c, nnq);                               \
    }                                                                                            \
}                                                                                                \
/* include/tlib/detail/ttm.h:1408-1474 (combined) and the three executors it                   \
 * picks: :1135-1194, :700-767, :950-985.  threads = cores as in the reference. */              \
static void ttm_combined_##SUFFIX(unsigned q, unsigned p, const T *a, const size_t *na,         \
                                  const size_t *wa, const size_t *pia, const T *b,              \
                                  const size_t *nb, const size_t *pib, T *c, const size_t *nc,  \
                                  const size_t *wc, unsigned cores, int *branch)                \
{                                                                                                \
    int is_cm = pib[0] == 1;                                                                     \
    if (!or_is_case(8, p, q, pia)) {                                                             \
        *branch = 0;                                                                             \
        blas_threads((int)cores);                                                                \
        if (is_cm) mtm_cm_##SUFFIX(q, p, a, na, pia, b, nb, c);                                 \
        else mtm_rm_##SUFFIX(q, p, a, na, pia, b, nb, c);                                       \
        return;                                                                                  \
    }                                                                                            \
    unsigned qh = (unsigned)or_inverse_mode(pia, p, q);                                          \
    size_t inner = or_product(na, pia, 2, qh);                                                   \
    size_t outer = or_product(na, pia, qh + 1, p + 1);                                           \
    size_t m = nc[q - 1], nq = na[q - 1];                                                        \
    if (outer >= cores) {                                                                        \
        /* par-loop / subtensor / all: one GEMM w x m x nq per outer slab */                    \
        *branch = 1;                                                                             \
        size_t nnq = or_product(na, pia, 1, qh);                                                 \
        size_t waq = wa[pia[qh] - 1], wcq = wc[pia[qh] - 1];                                     \
        _Pragma("omp parallel for schedule(static) num_threads(cores)")                         \
        for (size_t k = 0; k < outer; ++k) {                                                     \
            const T *aa = a + k * waq;                                                           \
            T *cc = c + k * wcq;                                                                 \
            blas_threads(1);                                                                     \
            if (is_cm) GEMM(COL, NOTR, TR, nnq, m, nq, aa, nnq, b, m, cc, nnq);                 \
            else GEMM(ROW, NOTR, NOTR, m, nnq, nq, b, nq, aa, nnq, cc, nnq);                    \
        }                                                                                        \
    } else if (inner * outer >= cores) {                                                         \
        /* par-loop / slice / all: GEMMs n1 x m x nq over outer x inner slices */               \
        *branch = 2;                                                                             \
        size_t wao = wa[pia[qh] - 1], wco = wc[pia[qh] - 1];                                     \
        size_t wai = na[pia[0] - 1], wci = nc[pia[0] - 1];                                       \
        size_t n1 = na[pia[0] - 1], wq = wa[q - 1];                                              \
        _Pragma("omp parallel for schedule(static) collapse(2) num_threads(cores)")             \
        for (size_t k = 0; k < outer; ++k)                                                       \
            for (size_t j = 0; j < inner; ++j) {                                                 \
                const T *aa = a + k * wao + j * wai;                                             \
                T *cc = c + k * wco + j * wci;                                                   \
                blas_threads(1);                                                                 \
                if (is_cm) GEMM(COL, NOTR, TR, n1, m, nq, aa, wq, b, m, cc, wq);                \
                else GEMM(ROW, NOTR, NOTR, m, n1, nq, b, nq, aa, wq, cc, wq);                   \
            }                                                                                    \
    } else {                                                                                     \
        /* par-blas / subtensor / none: serial recursion, threaded GEMMs */                     \
        *branch = 3;                                                                             \
        blas_threads((int)cores);                                                                \
        size_t nnq = or_product(na, pia, 1, qh);                                                 \
        subtensor_loops_##SUFFIX(is_cm, p, qh, a, na, wa, pia, b, c, wc, nnq, m, nq);            \
    }                                                                                            \
}

DEFINE_TTM(s, float, gemm_f)
DEFINE_TTM(d, double, gemm_d)

/* Public oracle entry points: validate (returns the failing check number, as
 * or_validate) and run the reference's `combined` algorithm.  use_blas=0
 * uses the naive GEMM; use_blas=1 the dlopen'ed CBLAS (when one is bound).
 * cores=0 means "compute it the reference's way".  *branch reports which
 * combined branch ran (0 direct case 1..7, 1 par-loop/subtensor,
 * 2 par-loop/slice, 3 par-blas/subtensor). */
int or_ttm_s(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
             const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
             const size_t *nc, const size_t *wc, int use_blas, unsigned cores, int *branch)
{
    int st = or_validate(q, p, a, na, wa, pia, b, nb, c, nc, wc);
    if (st) return st;
    int br = -1;
    g_use_blas = use_blas;
    ttm_combined_s(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, cores ? cores : or_cores(), &br);
    if (branch) *branch = br;
    return 0;
}

int or_ttm_d(unsigned q, unsigned p, const double *a, const size_t *na, const size_t *wa,
             const size_t *pia, const double *b, const size_t *nb, const size_t *pib, double *c,
             const size_t *nc, const size_t *wc, int use_blas, unsigned cores, int *branch)
{
    int st = or_validate(q, p, a, na, wa, pia, b, nb, c, nc, wc);
    if (st) return st;
    int br = -1;
    g_use_blas = use_blas;
    ttm_combined_d(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, cores ? cores : or_cores(), &br);
    if (branch) *branch = br;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3. FP64 brute force over arbitrary strides (the mathematical definition) */
/*    C(i1..j..ip) = sum_l A(i1..l..ip) * B(j,l)            README.md:13-18  */
/* ------------------------------------------------------------------------ */

/* a_is_double / b_is_double select the element type of A, B (float or
 * double); C is always written as double.  B(j,l) is at j + l*m when
 * b_colmajor, else j*nq + l.  Strides are element strides per mode (any
 * values, not only dense ones), so the definition is independent of the
 * layout-folding logic that is under test. */
void or_ttm_brute_f64(unsigned q, unsigned p, const void *a, int a_is_double, const size_t *na,
                      const size_t *wa, const void *b, int b_is_double, size_t m, int b_colmajor,
                      double *c, const size_t *wc)
{
    size_t nq = na[q - 1];
    size_t total = 1; /* number of index tuples over the non-q modes */
    for (unsigned r = 0; r < p; ++r)
        if (r != q - 1) total *= na[r];
    size_t idx[64];
    for (size_t t = 0; t < total; ++t) {
        size_t rem = t, offa = 0, offc = 0;
        for (unsigned r = 0; r < p; ++r) {
            if (r == q - 1) { idx[r] = 0; continue; }
            idx[r] = rem % na[r];
            rem /= na[r];
            offa += idx[r] * wa[r];
            offc += idx[r] * wc[r];
        }
        for (size_t j = 0; j < m; ++j) {
            long double s = 0.0L;
            for (size_t l = 0; l < nq; ++l) {
                size_t ia = offa + l * wa[q - 1];
                size_t ib = b_colmajor ? j + l * m : j * nq + l;
                long double av = a_is_double ? ((const double *)a)[ia] : ((const float *)a)[ia];
                long double bv = b_is_double ? ((const double *)b)[ib] : ((const float *)b)[ib];
                s += av * bv;
            }
            c[offc + j * wc[q - 1]] = (double)s;
        }
    }
}

/* Seeded generators shared with the test-suite and bench (SplitMix64). */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform[-1,1] as 2*(u53 * 2^-53) - 1, computed in double */
void or_fill_uniform_d(double *x, size_t n, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i) x[i] = 2.0 * ((double)(splitmix64(&s) >> 11) * 0x1.0p-53) - 1.0;
}

void or_fill_uniform_s(float *x, size_t n, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i)
        x[i] = (float)(2.0 * ((double)(splitmix64(&s) >> 11) * 0x1.0p-53) - 1.0);
}

/* sign-random log-uniform magnitude in [1e-3, 1e3] */
void or_fill_loguniform_s(float *x, size_t n, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i) {
        double u = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
        double sg = (splitmix64(&s) >> 63) ? -1.0 : 1.0;
        x[i] = (float)(sg * pow(10.0, -3.0 + 6.0 * u));
    }
}
