// Host executor of the folded product for callers without a GPU
// (ttm_sttm_cpu / ttm_dttm_cpu of include/ttm_amd.h).
//
//     C[k][j][i] = sum_{l < nq} B(j, l) * A[k][l][i]
//
// The reference runs this product through OpenMP and a multithreaded BLAS,
// one strategy per policy tag triple (include/tlib/detail/ttm.h:405-1403):
// loops over slices (GEMMs of n_{pi_1} columns) or subtensors (GEMMs of the
// whole block w below q), parallelised by an omp for, taskloops, recursive
// tasks, the BLAS's own threads, both, or one batched GEMM, and `combined`
// choosing among them by case and core count (:1408-1474).  Here the triple
// chooses the same decomposition of the work into tiles and the same kind of
// scheduling over a pool of std::threads:
//
//   slicing   subtensor: a work item is a slab k (all of its w columns);
//             slice: a slab's columns in pieces of n_{pi_1} (the GEMMs of
//             the reference's slice loops; for q first in the layout, one
//             slab row at a time)
//   parallel  sequential: one thread, items in order
//             parallel_loop: items split into one contiguous range per thread
//             parallel_taskloop: items handed out in chunks on demand
//             parallel_task: items handed out one at a time on demand
//             parallel_blas: items in order, each item's product split over
//             the threads (rows j x column blocks)
//             parallel_loop_blas: threads split between the item loop (ratio)
//             and each item's product
//             batched_gemm: every (item, row block, column block) tile of all
//             items as one flat, evenly split batch
//             combined: the reference's choice (cases 1-7: parallel_blas;
//             case 8: parallel_loop over subtensors when outer >= threads,
//             over slices when the slice count reaches the threads, else
//             parallel_blas over subtensors)
//   fusion    none / outer / all: how the reference nests or fuses the loops
//             over the non-contracted modes; the fold has fused them all, so
//             the tag only selects which triples exist.
//
// Every tile computes each of its C elements as one sum over l = 0 .. nq-1
// in order, so all policies and thread counts give bit-identical results.
// The file is compiled without floating-point contraction, so the result does
// not depend on whether the host has FMA instructions either.
#include "ttm_amd.h"
#include "ttm_plan.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace ttm_amd
{
namespace
{

struct Tile
{
    uint64_t k;      // slab
    uint64_t i0, i1; // columns of the slab
    uint64_t j0, j1; // rows of C
};

constexpr uint64_t kColBlock = 512; // columns of a tile's C rows kept in cache
constexpr uint64_t kRowBlock = 16;  // rows j of a tile

// One tile of the NN form (w > 1): C[k][j][i0:i1] for j in [j0, j1), by rank-1
// updates over l (the C rows of the tile stay in cache, A's row piece in L1).
template <class T>
void tile_nn(const Fold &f, const T *a, const T *b, T *c, const Tile &t)
{
    uint64_t const w = f.w, nq = f.nq, m = f.m;
    uint64_t const bsj = f.b_stride_j(), bsl = f.b_stride_l();
    T const *ak = a + t.k * nq * w;
    T *ck = c + t.k * m * w;
    for (uint64_t i0 = t.i0; i0 < t.i1; i0 += kColBlock) {
        uint64_t const i1 = std::min(t.i1, i0 + kColBlock);
        for (uint64_t j = t.j0; j < t.j1; ++j) std::fill(ck + j * w + i0, ck + j * w + i1, T(0));
        for (uint64_t l = 0; l < nq; ++l) {
            T const *ar = ak + l * w;
            for (uint64_t j = t.j0; j < t.j1; ++j) {
                T const bj = b[j * bsj + l * bsl];
                T *cr = ck + j * w;
                for (uint64_t i = i0; i < i1; ++i) cr[i] += bj * ar[i];
            }
        }
    }
}

// One tile of the TN form (w == 1): C[k][j] for slabs k in [i0, i1) (the
// columns are the slabs) and rows j in [j0, j1).
template <class T>
void tile_tn(const Fold &f, const T *a, const T *b, T *c, const Tile &t)
{
    uint64_t const nq = f.nq, m = f.m;
    if (f.b_colmajor) {
        // B(j, l) at b[j + l*m]: rows j contiguous for each l
        for (uint64_t k = t.i0; k < t.i1; ++k) {
            T *ck = c + k * m;
            std::fill(ck + t.j0, ck + t.j1, T(0));
            T const *ak = a + k * nq;
            for (uint64_t l = 0; l < nq; ++l) {
                T const al = ak[l];
                T const *bl = b + l * m;
                for (uint64_t j = t.j0; j < t.j1; ++j) ck[j] += bl[j] * al;
            }
        }
    } else {
        // B(j, l) at b[j*nq + l]: four slabs x four rows of dot products
        // at a time, each an in-order sum over l
        for (uint64_t k0 = t.i0; k0 < t.i1; k0 += 4) {
            uint64_t const kn = std::min<uint64_t>(4, t.i1 - k0);
            for (uint64_t j0 = t.j0; j0 < t.j1; j0 += 4) {
                uint64_t const jn = std::min<uint64_t>(4, t.j1 - j0);
                T acc[4][4] = {};
                T const *ap[4], *bp[4];
                for (uint64_t x = 0; x < 4; ++x) {
                    ap[x] = a + (k0 + std::min(x, kn - 1)) * nq;
                    bp[x] = b + (j0 + std::min(x, jn - 1)) * nq;
                }
                for (uint64_t l = 0; l < nq; ++l)
                    for (int x = 0; x < 4; ++x)
                        for (int y = 0; y < 4; ++y) acc[x][y] += bp[y][l] * ap[x][l];
                for (uint64_t x = 0; x < kn; ++x)
                    for (uint64_t y = 0; y < jn; ++y) c[(k0 + x) * m + j0 + y] = acc[x][y];
            }
        }
    }
}

template <class T>
void run_tile(const Fold &f, const T *a, const T *b, T *c, const Tile &t)
{
    if (f.w == 1) tile_tn(f, a, b, c, t);
    else tile_nn(f, a, b, c, t);
}

// Work items of a slicing policy (each covers all rows j).  In the TN form
// the columns are the slabs, so an item is a range of slabs.
std::vector<Tile> items(const Fold &f, bool subtensor, uint64_t slice_cols)
{
    std::vector<Tile> out;
    if (f.w == 1) {
        uint64_t const per = subtensor ? 64 : 1;
        for (uint64_t k = 0; k < f.outer; k += per) out.push_back({0, k, std::min(f.outer, k + per), 0, f.m});
        return out;
    }
    uint64_t const step = subtensor ? f.w : std::max<uint64_t>(1, std::min(slice_cols, f.w));
    for (uint64_t k = 0; k < f.outer; ++k)
        for (uint64_t i = 0; i < f.w; i += step) out.push_back({k, i, std::min(f.w, i + step), 0, f.m});
    return out;
}

// An item cut into tiles of kRowBlock rows and (NN form) kColBlock columns.
void split_item(const Fold &f, const Tile &it, std::vector<Tile> &out)
{
    uint64_t const cstep = f.w == 1 ? 16 : kColBlock;
    for (uint64_t j = it.j0; j < it.j1; j += kRowBlock)
        for (uint64_t i = it.i0; i < it.i1; i += cstep)
            out.push_back({it.k, i, std::min(it.i1, i + cstep), j, std::min(it.j1, j + kRowBlock)});
}

// Run fn(worker) on `n` threads (the caller is worker 0); a thread that
// cannot be started has its share run by the caller.
template <class F>
void on_threads(unsigned n, F fn)
{
    if (n <= 1) {
        fn(0u);
        return;
    }
    std::vector<std::thread> pool;
    std::vector<unsigned> inline_ids;
    for (unsigned t = 1; t < n; ++t) {
        try {
            pool.emplace_back(fn, t);
        } catch (...) {
            inline_ids.push_back(t);
        }
    }
    fn(0u);
    for (unsigned t : inline_ids) fn(t);
    for (auto &th : pool) th.join();
}

template <class T>
void run_static(const Fold &f, const T *a, const T *b, T *c, const std::vector<Tile> &tiles, unsigned nt)
{
    nt = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nt, tiles.size()));
    on_threads(nt, [&](unsigned t) {
        size_t const lo = tiles.size() * t / nt, hi = tiles.size() * (t + 1) / nt;
        for (size_t x = lo; x < hi; ++x) run_tile(f, a, b, c, tiles[x]);
    });
}

template <class T>
void run_dynamic(const Fold &f, const T *a, const T *b, T *c, const std::vector<Tile> &tiles, unsigned nt,
                 size_t chunk)
{
    nt = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nt, tiles.size()));
    std::atomic<size_t> next{0};
    on_threads(nt, [&](unsigned) {
        for (;;) {
            size_t const lo = next.fetch_add(chunk);
            if (lo >= tiles.size()) break;
            size_t const hi = std::min(tiles.size(), lo + chunk);
            for (size_t x = lo; x < hi; ++x) run_tile(f, a, b, c, tiles[x]);
        }
    });
}

template <class T>
ttm_status_t execute(const Fold &f, const T *a, const T *b, T *c, int par, int slicing, int fusion, unsigned nt,
                     unsigned na_first, double ratio)
{
    bool const defined =
        (par == TTM_CPU_COMBINED && slicing == TTM_CPU_SLICING_COMBINED && fusion == TTM_CPU_FUSION_ALL) ||
        (slicing == TTM_CPU_SLICE &&
         (((par == TTM_CPU_SEQUENTIAL || par == TTM_CPU_PARALLEL_BLAS || par == TTM_CPU_PARALLEL_TASKLOOP ||
            par == TTM_CPU_PARALLEL_TASK) && fusion == TTM_CPU_FUSION_NONE) ||
          ((par == TTM_CPU_PARALLEL_BLAS || par == TTM_CPU_PARALLEL_LOOP || par == TTM_CPU_PARALLEL_LOOP_BLAS) &&
           fusion == TTM_CPU_FUSION_ALL) ||
          (par == TTM_CPU_PARALLEL_LOOP && fusion == TTM_CPU_FUSION_OUTER))) ||
        (slicing == TTM_CPU_SUBTENSOR &&
         (((par == TTM_CPU_SEQUENTIAL || par == TTM_CPU_PARALLEL_BLAS || par == TTM_CPU_PARALLEL_TASKLOOP ||
            par == TTM_CPU_PARALLEL_TASK) && fusion == TTM_CPU_FUSION_NONE) ||
          ((par == TTM_CPU_PARALLEL_BLAS || par == TTM_CPU_PARALLEL_LOOP || par == TTM_CPU_PARALLEL_LOOP_BLAS ||
            par == TTM_CPU_BATCHED_GEMM) && fusion == TTM_CPU_FUSION_ALL)));
    if (!defined) {
        set_error(TTM_ERR_POLICY);
        return TTM_ERR_POLICY;
    }
    if (nt == 0) nt = std::max(1u, std::thread::hardware_concurrency());
    bool subtensor = slicing == TTM_CPU_SUBTENSOR;
    if (par == TTM_CPU_COMBINED) {
        // the reference's dispatch: the direct cases (w == 1 or outer == 1)
        // run one multithreaded product; case 8 parallelises the slab loop
        // when there are enough slabs, else the slice loop, else the product
        if (f.w == 1 || f.outer == 1) {
            par = TTM_CPU_PARALLEL_BLAS;
            subtensor = true;
        } else if (f.outer >= nt) {
            par = TTM_CPU_PARALLEL_LOOP;
            subtensor = true;
        } else if (f.outer * (f.w / std::max<uint64_t>(1, na_first)) >= nt) {
            par = TTM_CPU_PARALLEL_LOOP;
            subtensor = false;
        } else {
            par = TTM_CPU_PARALLEL_BLAS;
            subtensor = true;
        }
    }
    std::vector<Tile> const its = items(f, subtensor, na_first);
    switch (par) {
    case TTM_CPU_SEQUENTIAL: run_static(f, a, b, c, its, 1); break;
    case TTM_CPU_PARALLEL_LOOP: run_static(f, a, b, c, its, nt); break;
    case TTM_CPU_PARALLEL_TASKLOOP:
        run_dynamic(f, a, b, c, its, nt, std::max<size_t>(1, its.size() / (4 * (size_t)nt)));
        break;
    case TTM_CPU_PARALLEL_TASK: run_dynamic(f, a, b, c, its, nt, 1); break;
    case TTM_CPU_PARALLEL_BLAS: {
        // items in order, each one's tiles shared out over all threads
        // (one pool for the whole loop: thread t takes tiles t, t + nt, ...
        // of every item)
        std::vector<std::vector<Tile>> tiles(its.size());
        for (size_t x = 0; x < its.size(); ++x) split_item(f, its[x], tiles[x]);
        on_threads(nt, [&](unsigned t) {
            for (auto const &tl : tiles)
                for (size_t y = t; y < tl.size(); y += nt) run_tile(f, a, b, c, tl[y]);
        });
        break;
    }
    case TTM_CPU_PARALLEL_LOOP_BLAS: {
        // a `ratio` share of the threads on the item loop (the reference's
        // default splits them evenly), the rest inside each item: thread
        // (g, r) takes group g's contiguous range of items and, in each, the
        // tiles r, r + inner, ...
        double const r = ratio > 0 && ratio <= 1 ? ratio : 0.5;
        unsigned const outer_t = (unsigned)std::min<uint64_t>(std::max(1u, (unsigned)(nt * r)), its.size());
        unsigned const inner_t = std::max(1u, nt / outer_t);
        on_threads(outer_t * inner_t, [&](unsigned t) {
            unsigned const g = t / inner_t, rk = t % inner_t;
            size_t const lo = its.size() * g / outer_t, hi = its.size() * (g + 1) / outer_t;
            std::vector<Tile> tl;
            for (size_t x = lo; x < hi; ++x) {
                tl.clear();
                split_item(f, its[x], tl);
                for (size_t y = rk; y < tl.size(); y += inner_t) run_tile(f, a, b, c, tl[y]);
            }
        });
        break;
    }
    case TTM_CPU_BATCHED_GEMM: {
        std::vector<Tile> tiles;
        for (Tile const &it : its) split_item(f, it, tiles);
        run_static(f, a, b, c, tiles, nt);
        break;
    }
    default: set_error(TTM_ERR_POLICY); return TTM_ERR_POLICY;
    }
    return TTM_SUCCESS;
}

template <class T>
ttm_status_t run_cpu(unsigned q, unsigned p, const T *a, const size_t *na, const size_t *wa, const size_t *pia,
                     const T *b, const size_t *nb, const size_t *pib, T *c, const size_t *nc, const size_t *wc,
                     int par, int slicing, int fusion, unsigned threads, double ratio)
{
    Fold f;
    ttm_status_t st = plan(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, f);
    if (st != TTM_SUCCESS) return st;
    // extent of the fastest mode (the reference's slice width); q itself
    // when q is fastest
    unsigned const first = (unsigned)na[pia[0] - 1];
    st = execute<T>(f, a, b, c, par, slicing, fusion, threads, first, ratio);
    if (st == TTM_SUCCESS) clear_error();
    return st;
}

} // namespace
} // namespace ttm_amd

extern "C" {

ttm_status_t ttm_sttm_cpu(unsigned q, unsigned p, const float *a, const size_t *na, const size_t *wa,
                          const size_t *pia, const float *b, const size_t *nb, const size_t *pib, float *c,
                          const size_t *nc, const size_t *wc, ttm_cpu_parallel_t parallel,
                          ttm_cpu_slicing_t slicing, ttm_cpu_fusion_t fusion, unsigned threads, double ratio)
{
    return ttm_amd::run_cpu<float>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, parallel, slicing, fusion, threads,
                                   ratio);
}

ttm_status_t ttm_dttm_cpu(unsigned q, unsigned p, const double *a, const size_t *na, const size_t *wa,
                          const size_t *pia, const double *b, const size_t *nb, const size_t *pib, double *c,
                          const size_t *nc, const size_t *wc, ttm_cpu_parallel_t parallel,
                          ttm_cpu_slicing_t slicing, ttm_cpu_fusion_t fusion, unsigned threads, double ratio)
{
    return ttm_amd::run_cpu<double>(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, parallel, slicing, fusion,
                                    threads, ratio);
}

} // extern "C"
