// batch_solver.hpp -- independent LPs solved together in one persistent launch.
#pragma once

#include <vector>

#include "lp_types.hpp"

namespace cuopt_amd {

// Solves every LP of `lps` with the settings `st` (Stable3, FP64 or FP32-mixed
// iterates; no initial solutions, no refinement).  Results are in input order;
// solve_time is the wall time of the whole batch.
std::vector<SolveResult> batch_solve(const std::vector<const LpHost*>& lps, const PdlpSettings& st);

}  // namespace cuopt_amd
