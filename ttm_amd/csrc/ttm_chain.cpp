// Batched multi-TTM chains: plan, persistent grouped launches, hipGraph
// replay and the tile-shape autotuner (declared in include/ttm_amd.h).
//
// A chain over the modes of one tensor is run in layout order pi_1, pi_2, ...:
// stage s reads the current tensor as rows that are contiguous along mode
// pi_s (it is the fastest mode), multiplies them by U_{pi_s} and writes the
// result with that mode slowest (the TT view of ttm_gemm.hip).  The layout
// therefore rotates by one position per stage and after p stages is pi again.
// A mode without a factor is carried through by that rotation alone: its
// stage copies the current tensor into the rotated layout (a batched
// transpose, k_transpose_grouped) instead of multiplying by an identity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "ttm_amd.h"
#include "ttm_gemm.hpp"
#include "ttm_lowp.hpp"
#include "ttm_plan.hpp"
#include "ttm_rows.hpp"

using namespace ttm_amd;

namespace
{

struct Stage
{
    std::vector<GroupProblem> probs;   // tile_begin / nj filled for `cfg`
    std::vector<unsigned> tile_prob;
    GroupConfig cfg{4, 4, 1, 16};
    bool transpose = false;            // modes without a factor: a copy, 64 x 64 tiles
    // float32 stages whose contractions are 4..256 long and whose ranks are
    // at most 64 may run the short-row kernel (rows_height rows of B)
    int rows_height = 0;               // 0: not eligible
    bool rows = false;                 // the short-row kernel is chosen
    std::vector<RowsProblem> rprobs;
    RowsLayout rlayout;
    unsigned total = 0;
    unsigned grid = 0;
    GroupProblem *d_probs = nullptr;   // inside the plan's tile allocation
    unsigned *d_tile_prob = nullptr;
    RowsProblem *d_rprobs = nullptr;
    RowsSched rsched{};                // device schedule (inside the plan's tile allocation)
    std::string key;                   // shape class for the tuning cache
};

// Index of the short-row kernel in the tuning cache (after the tile shapes).
int rows_choice() { return group_config_count(); }

// ---- tuning cache ---------------------------------------------------------
std::mutex g_tune_mutex;
std::map<std::string, int> g_tune_cache;
bool g_tune_loaded = false;

const char *tune_file() { return std::getenv("TTM_AMD_TUNE_CACHE"); }

void tune_load_locked()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = tune_file();
    if (!path) return;
    std::ifstream in(path);
    std::string key;
    int idx;
    while (in >> key >> idx)
        if (idx >= 0 && idx <= rows_choice()) g_tune_cache[key] = idx;
}

void tune_store_locked()
{
    const char *path = tune_file();
    if (!path) return;
    std::ofstream out(path, std::ios::trunc);
    for (auto const &kv : g_tune_cache) out << kv.first << ' ' << kv.second << '\n';
}

// TTM_AMD_CHAIN_ROWS=0 keeps every stage on the grouped tiles.
bool rows_enabled()
{
    const char *v = std::getenv("TTM_AMD_CHAIN_ROWS");
    return !(v && v[0] == '0');
}

// TTM_AMD_CHAIN_MERGE=0 keeps every rank class of a stage in a launch of its
// own (no class is merged into a taller one).
bool merge_enabled()
{
    const char *v = std::getenv("TTM_AMD_CHAIN_MERGE");
    return !(v && v[0] == '0');
}

int log2_class(uint64_t x)
{
    int c = 0;
    while ((1ull << c) < x && c < 63) ++c;
    return c;
}

} // namespace

struct ttm_chain_plan_s
{
    bool is_double = false;
    ttm_precision_t prec = TTM_PREC_DEFAULT;
    std::vector<Stage> stages;
    void *d_static = nullptr;   // intermediates
    void *d_tiles = nullptr;    // descriptors + tile maps of all stages
    size_t tiles_bytes = 0;
    int num_cu = 0;
    hipGraphExec_t graph = nullptr;
    std::string symbol;
};

namespace
{

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

RowsKind rows_kind(ttm_chain_plan_t plan)
{
    return plan->is_double                    ? ROWS_F64
           : plan->prec == TTM_PREC_BF16X3 ? ROWS_BF16X3
           : plan->prec == TTM_PREC_FP16X2 ? ROWS_FP16X2
                                           : ROWS_F32;
}

ttm_status_t fail(ttm_status_t st, const char *detail)
{
    set_error(st, detail);
    return st;
}

ttm_status_t hip_check(hipError_t e)
{
    if (e == hipSuccess) return TTM_SUCCESS;
    (void)hipGetLastError();
    set_error(TTM_ERR_HIP, hipGetErrorString(e));
    return TTM_ERR_HIP;
}

// Heuristic tile shape for a stage: rows just above the largest rank, wide
// column blocks (the stages stream their input once).
GroupConfig default_config(uint64_t max_m, bool reduced)
{
    // the reduced-precision kernels have two shapes: 64 or 128 rows by 128 columns
    if (reduced) return max_m <= 64 ? GroupConfig{2, 4, 2, 32} : GroupConfig{4, 4, 2, 32};
    if (max_m <= 16) return group_config(0);
    if (max_m <= 32) return group_config(1);
    if (max_m <= 64) return group_config(2);
    return group_config(3);
}

bool config_allowed(const GroupConfig &cfg, bool is_double)
{
    if (is_double && cfg.bk > 16) return false; // slices of 32 are compiled for float32 only
    return cfg.nt * 16 * (4 / cfg.waves_m) >= 64;
}

// (Re)build the tile list of a stage for its configuration and upload it.
ttm_status_t layout_tiles(ttm_chain_plan_t plan)
{
    size_t bytes = 0;
    for (Stage &st : plan->stages) {
        if (st.rows) {
            st.rprobs.clear();
            for (GroupProblem const &P : st.probs)
                st.rprobs.push_back(RowsProblem{P.a, P.b, P.c, P.ncols, (unsigned)P.nq, (unsigned)P.m, 0, 0});
            st.grid = (unsigned)std::max(plan->num_cu, 1) * rows_wg_per_cu(st.rows_height);
            uint64_t units = 0;
            for (RowsProblem const &P : st.rprobs) units += (P.ncols + 15) / 16;
            if (units >= (1ull << 31)) return fail(TTM_ERR_ARGUMENT, "too many columns in one stage");
            rows_layout(st.rprobs, st.rows_height, st.grid, st.rlayout, rows_kind(plan));
            st.total = st.rlayout.units;
            bytes = align_up(bytes, 256) + st.rprobs.size() * sizeof(RowsProblem);
            bytes = align_up(bytes, 256) + (st.rlayout.wg_units.size() + st.rlayout.wg_prob.size()) * sizeof(unsigned);
            continue;
        }
        unsigned const bm = st.transpose ? 64u : (unsigned)st.cfg.bm();
        unsigned const bn = st.transpose ? 64u : (unsigned)st.cfg.bn();
        st.tile_prob.clear();
        unsigned begin = 0;
        // Workgroup b takes tiles b, b + grid, b + 2 grid, ...; a tile's cost
        // grows with its contraction length (16..256 indices in a Tucker
        // batch).  Listing the problems longest contraction first gives every
        // round of grid tiles nearly equal costs and leaves the cheapest tiles
        // for the last, partial round, so no workgroup finishes far behind.
        std::vector<size_t> order(st.probs.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return st.probs[x].nq > st.probs[y].nq; });
        for (size_t i : order) {
            GroupProblem &P = st.probs[i];
            unsigned const nj = (unsigned)((P.m + bm - 1) / bm);
            uint64_t const nn = (P.ncols + bn - 1) / bn;
            uint64_t const tiles = (uint64_t)nj * nn;
            if (begin + tiles >= (1ull << 31)) return fail(TTM_ERR_ARGUMENT, "too many tiles in one stage");
            P.tile_begin = begin;
            P.nj = nj;
            P.flags = (P.flags & ~TILE_STREAM) | (nj == 1 ? TILE_STREAM : 0u); // A read once
            st.tile_prob.insert(st.tile_prob.end(), (size_t)tiles, (unsigned)i);
            begin += (unsigned)tiles;
        }
        st.total = begin;
        int occ = st.transpose                     ? 8
                  : lowp_group_precision(plan->prec) ? lowp_group_occupancy(plan->prec, st.cfg.bm())
                                                     : group_occupancy(st.cfg, plan->is_double);
        if (occ <= 0) occ = 1;
        uint64_t const resident = (uint64_t)occ * (uint64_t)std::max(plan->num_cu, 1);
        st.grid = (unsigned)std::min<uint64_t>(st.total, resident);
        bytes = align_up(bytes, 256) + st.probs.size() * sizeof(GroupProblem);
        bytes = align_up(bytes, 256) + st.tile_prob.size() * sizeof(unsigned);
    }
    if (plan->d_tiles && bytes > plan->tiles_bytes) {
        (void)hipFree(plan->d_tiles);
        plan->d_tiles = nullptr;
    }
    if (!plan->d_tiles) {
        if (hipMalloc(&plan->d_tiles, std::max<size_t>(bytes, 256)) != hipSuccess) {
            (void)hipGetLastError();
            plan->d_tiles = nullptr;
            return fail(TTM_ERR_ALLOC, "chain tile maps");
        }
        plan->tiles_bytes = std::max<size_t>(bytes, 256);
    }
    size_t off = 0;
    char *base = static_cast<char *>(plan->d_tiles);
    for (Stage &st : plan->stages) {
        if (st.rows) {
            off = align_up(off, 256);
            st.d_rprobs = reinterpret_cast<RowsProblem *>(base + off);
            ttm_status_t s = hip_check(hipMemcpy(st.d_rprobs, st.rprobs.data(), st.rprobs.size() * sizeof(RowsProblem),
                                                 hipMemcpyHostToDevice));
            if (s) return s;
            off += st.rprobs.size() * sizeof(RowsProblem);
            off = align_up(off, 256);
            // the ranges' first units and their problems
            RowsLayout const &L = st.rlayout;
            std::vector<unsigned> words(L.wg_units);
            words.insert(words.end(), L.wg_prob.begin(), L.wg_prob.end());
            unsigned *w = reinterpret_cast<unsigned *>(base + off);
            s = hip_check(hipMemcpy(w, words.data(), words.size() * sizeof(unsigned), hipMemcpyHostToDevice));
            if (s) return s;
            st.rsched.wg_units = w;
            st.rsched.wg_prob = w + L.wg_units.size();
            off += words.size() * sizeof(unsigned);
            continue;
        }
        off = align_up(off, 256);
        st.d_probs = reinterpret_cast<GroupProblem *>(base + off);
        ttm_status_t s = hip_check(
            hipMemcpy(st.d_probs, st.probs.data(), st.probs.size() * sizeof(GroupProblem), hipMemcpyHostToDevice));
        if (s) return s;
        off += st.probs.size() * sizeof(GroupProblem);
        off = align_up(off, 256);
        st.d_tile_prob = reinterpret_cast<unsigned *>(base + off);
        s = hip_check(hipMemcpy(st.d_tile_prob, st.tile_prob.data(), st.tile_prob.size() * sizeof(unsigned),
                                hipMemcpyHostToDevice));
        if (s) return s;
        off += st.tile_prob.size() * sizeof(unsigned);
    }
    return TTM_SUCCESS;
}

ttm_status_t launch_stage(ttm_chain_plan_t plan, const Stage &st, hipStream_t s)
{
    if (st.total == 0) return TTM_SUCCESS;
    if (st.rows)
        return launch_rows_grouped(st.rows_height, st.d_rprobs, st.rsched, st.grid, s, rows_kind(plan));
    if (st.transpose)
        return plan->is_double ? launch_transpose_grouped<double>(st.d_probs, st.d_tile_prob, st.total, st.grid, s)
                               : launch_transpose_grouped<float>(st.d_probs, st.d_tile_prob, st.total, st.grid, s);
    if (lowp_group_precision(plan->prec))
        return launch_lowp_grouped(plan->prec, st.cfg.bm(), st.d_probs, st.d_tile_prob, st.total, st.grid, s);
    if (plan->is_double)
        return launch_gemm_grouped<double>(st.cfg, st.d_probs, st.d_tile_prob, st.total, st.grid, s);
    return launch_gemm_grouped<float>(st.cfg, st.d_probs, st.d_tile_prob, st.total, st.grid, s);
}

void drop_graph(ttm_chain_plan_t plan)
{
    if (plan->graph) {
        (void)hipGraphExecDestroy(plan->graph);
        plan->graph = nullptr;
    }
}

} // namespace

extern "C" {

ttm_status_t ttm_chain_plan_create(size_t count, const ttm_chain_desc_t *descs, int is_double,
                                   ttm_precision_t precision, ttm_chain_plan_t *plan_out)
{
    if (!plan_out) return fail(TTM_ERR_ARGUMENT, "plan pointer is null");
    *plan_out = nullptr;
    if (count > 0 && !descs) return fail(TTM_ERR_ARGUMENT, "descriptor array is null");
    if (precision != TTM_PREC_DEFAULT && (is_double || !lowp_group_precision(precision)))
        return fail(TTM_ERR_PRECISION, "chains run in the plain precision or, for float, BF16X3 or FP16X2");
    size_t const esz = is_double ? sizeof(double) : sizeof(float);
    size_t const vec = 16 / esz;

    // ---- validation and sizes ------------------------------------------------
    size_t max_p = 0, ws_total = 0;
    std::vector<size_t> ws_of(count, 0);
    for (size_t t = 0; t < count; ++t) {
        ttm_chain_desc_t const &d = descs[t];
        char msg[96];
        std::snprintf(msg, sizeof msg, "chain descriptor %zu", t);
        if (d.p == 0 || d.p > TTM_CHAIN_MAX_ORDER) return fail(TTM_ERR_ARGUMENT, msg);
        if (!d.a || !d.c) return fail(TTM_ERR_ARGUMENT, msg);
        bool seen[TTM_CHAIN_MAX_ORDER + 1] = {false};
        for (unsigned r = 0; r < d.p; ++r) {
            if (d.pi[r] == 0 || d.pi[r] > d.p || seen[d.pi[r]]) return fail(TTM_ERR_LAYOUT_A, msg);
            seen[d.pi[r]] = true;
            if (d.n[r] == 0) return fail(TTM_ERR_SHAPE_A, msg);
            if (d.u[r] && d.r[r] == 0) return fail(TTM_ERR_SHAPE_C, msg);
        }
        max_p = std::max<size_t>(max_p, d.p);
        // intermediate sizes after each stage but the last
        uint64_t cur[TTM_CHAIN_MAX_ORDER];
        for (unsigned r = 0; r < d.p; ++r) cur[r] = d.n[r];
        uint64_t biggest = 0;
        for (unsigned s = 0; s < d.p; ++s) {
            unsigned const mode = (unsigned)d.pi[s] - 1;
            cur[mode] = d.u[mode] ? d.r[mode] : d.n[mode];
            if (s + 1 < d.p) {
                uint64_t sz = 1;
                for (unsigned r = 0; r < d.p; ++r) sz *= cur[r];
                biggest = std::max(biggest, sz);
            }
        }
        ws_of[t] = align_up((size_t)biggest * esz, 256);
        ws_total += 2 * ws_of[t];
    }

    auto *plan = new ttm_chain_plan_s;
    plan->is_double = is_double != 0;
    plan->prec = precision;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        delete plan;
        return fail(TTM_ERR_HIP, "device query failed");
    }
    plan->num_cu = prop.multiProcessorCount;

    // ---- static allocation: ping-pong intermediates ----------------------------
    size_t const static_bytes = std::max<size_t>(ws_total, 256);
    if (hipMalloc(&plan->d_static, static_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete plan;
        return fail(TTM_ERR_ALLOC, "chain intermediates");
    }
    char *sbase = static_cast<char *>(plan->d_static);

    // ---- stages ---------------------------------------------------------------
    // Stage s of every chain, split by row class of m (<= 16, <= 32, <= 48,
    // <= 64, larger) so that each launch gets a tile height that fits its
    // ranks, and within a class into the problems the short-row kernels can
    // take (4 <= n_q <= 256) and the rest, plus one copy launch for the modes
    // without a factor; launches of one stage are independent, stages run in
    // order.  (with the short-row kernels switched off, the reduced precisions
    // keep the two tile heights of their grouped kernel)
    int const classes = lowp_group_precision(precision) && !rows_enabled() ? 2 : 5;
    // per stage: the rank classes of the problems the short-row kernels can
    // take, the same classes of the rest, the copy launch
    int const groups_per_stage = 2 * classes + 1;
    auto m_class = [&](uint64_t m) {
        if (classes == 2) return m <= 64 ? 0 : 1; // the two reduced-precision tile heights
        return m <= 16 ? 0 : m <= 32 ? 1 : m <= 48 ? 2 : m <= 64 ? 3 : 4;
    };
    std::vector<Stage> groups(max_p * groups_per_stage);
    std::vector<uint64_t> stage_max_m(groups.size(), 0), stage_cols(groups.size(), 0), stage_nq(groups.size(), 0);
    size_t ws_off = 0;
    for (size_t t = 0; t < count; ++t) {
        ttm_chain_desc_t const &d = descs[t];
        char *buf[2] = {sbase + ws_off, sbase + ws_off + ws_of[t]};
        ws_off += 2 * ws_of[t];
        uint64_t cur[TTM_CHAIN_MAX_ORDER];
        for (unsigned r = 0; r < d.p; ++r) cur[r] = d.n[r];
        const void *in = d.a;
        for (unsigned s = 0; s < d.p; ++s) {
            unsigned const mode = (unsigned)d.pi[s] - 1;
            GroupProblem P{};
            P.a = in;
            P.nq = cur[mode];
            P.m = d.u[mode] ? d.r[mode] : d.n[mode];
            P.b = d.u[mode]; // null: the copy stage
            uint64_t ncols = 1;
            for (unsigned r = 0; r < d.p; ++r)
                if (r != mode) ncols *= cur[r];
            P.ncols = ncols;
            void *out = (s + 1 == d.p) ? d.c : static_cast<void *>(buf[s & 1]);
            P.c = out;
            bool const a_al = (reinterpret_cast<uintptr_t>(P.a) & 15u) == 0;
            bool const c_al = (reinterpret_cast<uintptr_t>(P.c) & 15u) == 0;
            bool const b_al = (reinterpret_cast<uintptr_t>(P.b) & 15u) == 0;
            P.flags = ((a_al && P.nq % vec == 0) ? TILE_A_VEC : 0u) | ((c_al && ncols % vec == 0) ? TILE_C_VEC : 0u) |
                      (b_al ? TILE_B_VEC : 0u);
            bool const short_rows = rows_enabled() && rows_eligible(P.nq, P.m, P.ncols);
            size_t const g = (size_t)s * groups_per_stage +
                             (d.u[mode] ? (size_t)m_class(P.m) + (short_rows ? 0 : classes) : 2 * (size_t)classes);
            groups[g].transpose = !d.u[mode];
            groups[g].probs.push_back(P);
            stage_max_m[g] = std::max(stage_max_m[g], P.m);
            stage_cols[g] += ncols;
            stage_nq[g] = std::max(stage_nq[g], P.nq);
            cur[mode] = P.m;
            in = out;
        }
    }
    // Ranks 33..48 keep a launch of their own (the 48-row short-row kernels:
    // a quarter less matrix work than on 64 rows) only where the stage gives
    // the two launches enough work, at least 64 units of 16 columns per
    // compute unit together; in smaller stages the second launch's start and
    // tail cost more than the rows save.  The tiles have one 64-row height
    // for both classes, so there the two are always one launch.  A stage
    // whose short-row problems of ranks up to 64 have fewer than 128 units
    // per compute unit in all runs them as one launch on the tallest height
    // it needs: each launch of so little work is mostly its start and its
    // tail.
    if (classes == 5 && merge_enabled()) {
        uint64_t const cu = (uint64_t)std::max(plan->num_cu, 1);
        auto units_of = [&](size_t g) {
            uint64_t u = 0;
            for (GroupProblem const &P : groups[g].probs) u += (P.ncols + 15) / 16;
            return u;
        };
        auto merge = [&](size_t from, size_t into) {
            Stage &lo = groups[from], &hi = groups[into];
            hi.probs.insert(hi.probs.end(), lo.probs.begin(), lo.probs.end());
            lo.probs.clear();
            stage_max_m[into] = std::max(stage_max_m[into], stage_max_m[from]);
            stage_cols[into] += stage_cols[from];
            stage_nq[into] = std::max(stage_nq[into], stage_nq[from]);
            stage_max_m[from] = stage_cols[from] = stage_nq[from] = 0;
        };
        for (size_t s = 0; s < max_p; ++s) {
            size_t const base = s * groups_per_stage;
            uint64_t all = 0;
            size_t top = base; // the tallest non-empty short-row class up to 64 rows
            for (size_t c = 0; c < 4; ++c) {
                all += units_of(base + c);
                if (!groups[base + c].probs.empty()) top = base + c;
            }
            if (all < 128 * cu) {
                for (size_t c = 0; c < 4; ++c)
                    if (base + c != top && !groups[base + c].probs.empty()) merge(base + c, top);
            } else if (!groups[base + 2].probs.empty() && units_of(base + 2) + units_of(base + 3) < 64 * cu) {
                merge(base + 2, base + 3);
            }
            size_t const rest = base + (size_t)classes;
            if (!groups[rest + 2].probs.empty()) merge(rest + 2, rest + 3);
        }
    }
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        tune_load_locked();
        for (size_t g = 0; g < groups.size(); ++g) {
            Stage &st = groups[g];
            if (st.probs.empty()) continue;
            if (st.transpose) {
                plan->stages.push_back(std::move(st));
                continue;
            }
            std::ostringstream key;
            key << (is_double ? "f64" : "f32") << "/p" << (int)precision << "/m" << log2_class(stage_max_m[g])
                << "/nq" << log2_class(stage_nq[g]) << "/cols" << log2_class(stage_cols[g]) << "/n"
                << log2_class(st.probs.size()) << "/r"
                << (g % groups_per_stage < (size_t)classes ? rows_height(stage_max_m[g]) : 0);
            st.key = key.str();
            st.cfg = default_config(stage_max_m[g], lowp_group_precision(precision));
            if (rows_enabled()) {
                bool ok = true;
                for (GroupProblem const &P : st.probs) ok = ok && rows_eligible(P.nq, P.m, P.ncols);
                if (ok) st.rows_height = rows_height(stage_max_m[g]);
            }
            st.rows = st.rows_height != 0; // the default where it applies
            auto it = g_tune_cache.find(st.key);
            if (!lowp_group_precision(precision) && it != g_tune_cache.end()) {
                if (it->second == rows_choice()) {
                    st.rows = st.rows_height != 0;
                } else {
                    st.rows = false;
                    st.cfg = group_config(it->second);
                }
            }
            plan->stages.push_back(std::move(st));
        }
    }
    ttm_status_t const st = layout_tiles(plan);
    if (st != TTM_SUCCESS) {
        ttm_chain_plan_destroy(plan);
        return st;
    }
    *plan_out = plan;
    clear_error();
    return TTM_SUCCESS;
}

ttm_status_t ttm_chain_run(ttm_chain_plan_t plan, ttm_stream_t stream)
{
    if (!plan) return fail(TTM_ERR_ARGUMENT, "plan is null");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    for (Stage const &st : plan->stages) {
        ttm_status_t const r = launch_stage(plan, st, s);
        if (r != TTM_SUCCESS) return r;
    }
    clear_error();
    return TTM_SUCCESS;
}

ttm_status_t ttm_chain_run_graph(ttm_chain_plan_t plan, ttm_stream_t stream)
{
    if (!plan) return fail(TTM_ERR_ARGUMENT, "plan is null");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    if (!plan->graph) {
        // capture on a private stream so the caller's stream may be the null stream
        hipStream_t cap = nullptr;
        ttm_status_t r = hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        if (r) return r;
        hipGraph_t g = nullptr;
        r = hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        if (!r) {
            for (Stage const &st : plan->stages) {
                r = launch_stage(plan, st, cap);
                if (r) break;
            }
            ttm_status_t const e = hip_check(hipStreamEndCapture(cap, &g));
            if (!r) r = e;
        }
        if (!r) r = hip_check(hipGraphInstantiate(&plan->graph, g, nullptr, nullptr, 0));
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cap);
        if (r) {
            plan->graph = nullptr;
            return r;
        }
    }
    ttm_status_t const r = hip_check(hipGraphLaunch(plan->graph, s));
    if (!r) clear_error();
    return r;
}

ttm_status_t ttm_chain_tune(ttm_chain_plan_t plan, ttm_stream_t stream)
{
    if (!plan) return fail(TTM_ERR_ARGUMENT, "plan is null");
    if (lowp_group_precision(plan->prec)) return TTM_SUCCESS; // one compiled shape per tile height
    drop_graph(plan);
    hipStream_t const s = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    ttm_status_t r = hip_check(hipEventCreate(&e0));
    if (r) return r;
    r = hip_check(hipEventCreate(&e1));
    if (r) {
        (void)hipEventDestroy(e0);
        return r;
    }
    for (size_t si = 0; si < plan->stages.size() && !r; ++si) {
        float best_ms = 1e30f;
        int best = -1;
        if (plan->stages[si].transpose) continue; // one copy kernel
        for (int ci = 0; ci <= rows_choice() && !r; ++ci) {
            if (ci == rows_choice()) {
                if (plan->stages[si].rows_height == 0) continue;
                plan->stages[si].rows = true;
            } else {
                GroupConfig const cfg = group_config(ci);
                if (!config_allowed(cfg, plan->is_double)) continue;
                plan->stages[si].rows = false;
                plan->stages[si].cfg = cfg;
            }
            r = layout_tiles(plan);
            if (r) break;
            Stage const &st = plan->stages[si];
            r = launch_stage(plan, st, s); // warm-up
            if (r) break;
            r = hip_check(hipEventRecord(e0, s));
            for (int it = 0; it < 3 && !r; ++it) r = launch_stage(plan, st, s);
            if (!r) r = hip_check(hipEventRecord(e1, s));
            if (!r) r = hip_check(hipEventSynchronize(e1));
            float ms = 0.f;
            if (!r) r = hip_check(hipEventElapsedTime(&ms, e0, e1));
            if (!r && ms < best_ms) {
                best_ms = ms;
                best = ci;
            }
        }
        if (!r && best >= 0) {
            plan->stages[si].rows = best == rows_choice();
            if (best != rows_choice()) plan->stages[si].cfg = group_config(best);
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            tune_load_locked();
            g_tune_cache[plan->stages[si].key] = best;
            tune_store_locked();
        }
    }
    if (!r) r = layout_tiles(plan);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!r) clear_error();
    return r;
}

unsigned ttm_chain_stages(ttm_chain_plan_t plan) { return plan ? (unsigned)plan->stages.size() : 0u; }

const char *ttm_chain_stage_kernel(ttm_chain_plan_t plan, unsigned s)
{
    if (!plan || s >= plan->stages.size()) return "";
    plan->symbol = plan->stages[s].rows               ? rows_symbol(plan->stages[s].rows_height, rows_kind(plan))
                   : plan->stages[s].transpose        ? (plan->is_double ? "k_transpose_grouped<double>"
                                                                         : "k_transpose_grouped<float>")
                   : lowp_group_precision(plan->prec) ? lowp_group_symbol(plan->prec, plan->stages[s].cfg.bm())
                                                      : group_symbol(plan->stages[s].cfg, plan->is_double);
    return plan->symbol.c_str();
}

void ttm_chain_plan_destroy(ttm_chain_plan_t plan)
{
    if (!plan) return;
    drop_graph(plan);
    // launches of this plan may still read its descriptors and intermediates
    if (plan->d_tiles || plan->d_static) (void)hipDeviceSynchronize();
    if (plan->d_tiles) (void)hipFree(plan->d_tiles);
    if (plan->d_static) (void)hipFree(plan->d_static);
    delete plan;
}

} // extern "C"
