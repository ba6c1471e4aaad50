// Grouped TT products with short contractions (float32, BF16X3, FP16X2, float64):
// the kernels for multi-TTM chain stages whose contracted extent is at most 256.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ttm_amd.h"

namespace ttm_amd
{

// One product of a short-row launch, in the TT view:
//   C[j * ncols + n] = sum_{l < nq} B[j * nq + l] * A[n * nq + l]
// (elements of the launch's kind: float for ROWS_F32 and ROWS_BF16X3, double
// for ROWS_F64).  Its columns are cut into units of 16 (unit u covers
// columns 16 u .. 16 u + 15); unit_begin is the problem's first unit in the
// launch's list.
struct RowsProblem
{
    const void *a;
    const void *b;
    void *c;
    uint64_t ncols;
    unsigned nq;
    unsigned m;
    unsigned unit_begin;
    unsigned nunits;
};

// The arithmetic of a launch: plain float32 (k_rows_grouped), float32 with A
// and B split into three bfloat16 terms and six products on the 16x16x32 bf16
// MFMA (k_rows_bf16x3), float32 with A and B scaled by powers of two and split
// into two binary16 terms, three products on the 16x16x32 f16 MFMA
// (k_rows_fp16x2), plain float64 (k_rows_f64).
enum RowsKind : int
{
    ROWS_F32 = 0,
    ROWS_BF16X3 = 1,
    ROWS_F64 = 2,
    ROWS_FP16X2 = 3
};

constexpr unsigned kRowsMinNq = 4;
constexpr unsigned kRowsMaxNq = 256;

// Row heights the kernels are compiled for (16, 32, 48, 64); 0 when m > 64.
int rows_height(uint64_t m);

// A problem the kernels can run: 4 <= nq <= 256, 1 <= m <= 64.
bool rows_eligible(uint64_t nq, uint64_t m, uint64_t ncols);

// Workgroups per compute unit of the kernel for a row height.
unsigned rows_wg_per_cu(int height);

// Kernel instance name, as profilers print it.
std::string rows_symbol(int height, RowsKind kind = ROWS_F32);

// Waves per workgroup of the kernel.
unsigned rows_waves(int height, RowsKind kind = ROWS_F32);

// A launch's schedule: workgroup b runs the units [wg_units[b],
// wg_units[b + 1]), which start inside problem wg_prob[b].
struct RowsLayout
{
    std::vector<unsigned> wg_units, wg_prob;
    unsigned units = 0;
};

// The same in device memory, passed to the kernel by value.
struct RowsSched
{
    const unsigned *wg_units;
    const unsigned *wg_prob;
};

// Units of the problems (unit_begin / nunits filled, in order) and the
// launch's schedule: contiguous ranges of nearly equal estimated cost (HBM
// bytes or padded matrix work, whichever is larger, plus a fixed cost per
// unit).
void rows_layout(std::vector<RowsProblem> &probs, int height, unsigned grid, RowsLayout &out,
                 RowsKind kind = ROWS_F32);

ttm_status_t launch_rows_grouped(int height, const RowsProblem *d_probs, const RowsSched &sc, unsigned grid,
                                 hipStream_t s, RowsKind kind = ROWS_F32);

} // namespace ttm_amd
