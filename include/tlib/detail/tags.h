// Policy tags of the drop-in API (same namespaces, types and constants as
// include/tlib/detail/tags.h:22-69 of the reference).  The reference uses
// them to pick a CPU threading strategy (OpenMP loops, tasks, threaded BLAS).
// On MI355X every triple runs the same planner, so all of them are accepted.
#pragma once

namespace tlib::ttm::parallel_policy
{
struct sequential_t {};
struct parallel_blas_t {};
struct parallel_loop_t {};
struct parallel_taskloop_t {};
struct parallel_task_t {};
struct parallel_loop_blas_t {};
struct batched_gemm_t {};
struct combined_t {};

inline constexpr sequential_t sequential;
inline constexpr parallel_blas_t parallel_blas;
inline constexpr parallel_loop_t parallel_loop;
inline constexpr parallel_taskloop_t parallel_taskloop;
inline constexpr parallel_task_t parallel_task;
inline constexpr parallel_loop_blas_t parallel_loop_blas;
inline constexpr batched_gemm_t batched_gemm;
inline constexpr combined_t combined;
} // namespace tlib::ttm::parallel_policy

namespace tlib::ttm::slicing_policy
{
struct slice_t {};
struct subtensor_t {};
struct combined_t {};

inline constexpr combined_t combined;
inline constexpr slice_t slice;
inline constexpr subtensor_t subtensor;
} // namespace tlib::ttm::slicing_policy

namespace tlib::ttm::fusion_policy
{
struct none_t {};
struct outer_t {};
struct all_t {};

inline constexpr none_t none;
inline constexpr outer_t outer;
inline constexpr all_t all;
} // namespace tlib::ttm::fusion_policy
