// The short-row kernels' schedule (ttm_amd/csrc/ttm_rows.hip, rows_layout) on
// random problem lists: every unit of every problem lies in exactly one
// workgroup range, the ranges are contiguous and in order, wg_prob names the
// problem of each range's first unit, and the ranges' estimated costs are
// even (no range above twice the mean plus one unit).  Host code only.
#include "ttm_rows.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

using namespace ttm_amd;

int main()
{
    std::mt19937 rg(3);
    int bad = 0;
    for (int trial = 0; trial < 400; ++trial) {
        int const np = 1 + (int)(rg() % 300);
        std::vector<RowsProblem> P(np);
        for (auto &p : P) {
            p = RowsProblem{};
            p.ncols = 1 + rg() % (rg() % 4 == 0 ? 70000 : 300);
            p.nq = 4 + rg() % 253;
            p.m = 1 + rg() % 64;
        }
        int const height = (int[]){16, 32, 48, 64}[rg() % 4];
        RowsKind const kind = (RowsKind[]){ROWS_F32, ROWS_BF16X3, ROWS_F64}[rg() % 3];
        unsigned const grid = (unsigned[]){1, 7, 256, 512}[rg() % 4];
        RowsLayout L;
        rows_layout(P, height, grid, L, kind);
        unsigned total = 0;
        for (int i = 0; i < np; ++i) {
            if (P[i].unit_begin != total || P[i].nunits != (P[i].ncols + 15) / 16) ++bad;
            total += P[i].nunits;
        }
        if (total != L.units || L.wg_units.size() != grid + 1 || L.wg_prob.size() != grid) {
            ++bad;
            continue;
        }
        if (L.wg_units[0] != 0 || L.wg_units[grid] != total) ++bad;
        for (unsigned b = 0; b < grid; ++b) {
            unsigned const u0 = L.wg_units[b], u1 = L.wg_units[b + 1];
            if (u0 > u1) ++bad;
            if (u0 < u1) {
                unsigned const q = L.wg_prob[b];
                if (q >= (unsigned)np || !(P[q].unit_begin <= u0 && u0 < P[q].unit_begin + P[q].nunits)) ++bad;
            }
        }
        // even ranges: count units per range against the mean
        unsigned longest = 0;
        for (unsigned b = 0; b < grid; ++b) longest = std::max(longest, L.wg_units[b + 1] - L.wg_units[b]);
        double const mean = (double)total / grid;
        // unit costs differ by at most the ratio of the largest to the smallest unit
        // (contraction 4..256, rows 1..64): bound loosely by 130x the mean
        if (longest > 130.0 * mean + 1) ++bad;
    }
    std::printf("rows schedule: %d violations\n", bad);
    return bad != 0;
}
