// lp_types.hpp -- host-side problem, settings and solution records of the C ABI.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cuopt_amd_ext.h"

namespace cuopt_amd {

// optimization_problem_t's LP fields (pdlp/optimization_problem.cu) as the
// C ABI hands them over: CSR matrix, objective, row sense/rhs or ranged rows,
// variable bounds.
struct LpHost {
  int m = 0, n = 0;
  std::vector<int> rowptr, colidx;
  std::vector<double> vals;
  std::vector<double> c;
  bool maximize = false;
  double offset = 0.0;
  bool ranged   = false;
  std::vector<char> sense;   // 'E','L','G' (row-type form)
  std::vector<double> rhs;   // row-type form
  std::vector<double> lc, uc;  // ranged form
  std::vector<double> lb, ub;
  std::vector<char> var_types;
  // Names read from an MPS file (empty for problems built from arrays).
  std::vector<std::string> row_names, col_names;
  std::string objective_name, problem_name;
};

// pdlp_solver_settings_t fields used by the LP path (pdlp/solver_settings.hpp:193-262).
struct PdlpSettings {
  double abs_primal = 1e-4, rel_primal = 1e-4;
  double abs_dual = 1e-4, rel_dual = 1e-4;
  double abs_gap = 1e-4, rel_gap = 1e-4;
  double primal_infeasible_tol = 1e-10, dual_infeasible_tol = 1e-10;
  int iteration_limit = std::numeric_limits<int>::max();
  double time_limit   = std::numeric_limits<double>::infinity();
  int pdlp_solver_mode = CUOPT_PDLP_SOLVER_MODE_STABLE3;
  int method           = CUOPT_METHOD_CONCURRENT;
  int presolve         = CUOPT_PRESOLVE_DEFAULT;
  int precision        = CUOPT_AMD_PRECISION_FP64;
  int refinement       = 1;  // FP64 iterative refinement of FP32-mixed solves
  bool infeasibility_detection = false;
  bool strict_infeasibility    = false;
  bool per_constraint_residual = false;
  int per_constraint_absolute  = 0;  // 1: |r_i| - rel |w_i| in place of the signed form
  int spmv_autotune            = 0;  // 1: time tile / lanes candidates at setup
  int column_order             = -1; // -1 automatic, 0 as given, 1 first-touch relabelling
  int stop_on_nan              = 0;  // 1: a NaN residual ends a single-LP solve (NumericalError)
  int long_row_merge           = -1; // rows longer than a tile: 0 merge launch, 1 last chunk merges, -1 auto
  bool save_best_primal_so_far = false;
  bool first_primal_feasible   = false;
  bool log_to_console          = true;
  bool crossover               = false;
  bool dual_postsolve          = true;
  int num_cpu_threads          = -1;
  int num_gpus                 = 1;
  int random_seed              = -1;
  std::string log_file, solution_file, user_problem_file;
  std::vector<double> initial_primal, initial_dual;
  // Benchmark window (cuOptAmdBenchmarkSteps): time iterations [bench_warmup,
  // bench_warmup + bench_steps) and stop there.  Disabled when bench_steps <= 0.
  long long bench_warmup = 0, bench_steps = 0;
};

struct SolveResult {
  int status       = CUOPT_TERIMINATION_STATUS_NO_TERMINATION;
  int error        = CUOPT_SUCCESS;
  std::string error_message;
  std::vector<double> x, y, rc;
  double primal_objective = std::nan(""), dual_objective = std::nan("");
  double solve_time       = 0.0;
  double info[CUOPT_AMD_INFO_COUNT];
  std::vector<double> trace;  // rows of CUOPT_AMD_TRACE_WIDTH
  double bench[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // see cuOptAmdBenchmarkSteps
  SolveResult()
  {
    for (double& v : info) v = std::nan("");
  }
};

// Validation failure of the problem representation (problem_checking.cu).
struct validation_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// solve_lp for the PDLP path (solve.cu:1123-1342 with method PDLP).
SolveResult solve_lp(const LpHost& lp, const PdlpSettings& settings);

// [lc, uc] of every row: the ranged form as given, or the row types mapped as
// E: [b, b], L: [-inf, b], G: [b, +inf] (problem_helpers.cuh:26-51).
void constraint_bounds(const LpHost& lp, std::vector<double>& lc, std::vector<double>& uc);

// y = A x with the solver's tiled SpMV (test entry point).
void tiled_spmv_host(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                     const double* x, double* y, int precision);

// Device scratch of tiled_spmv_host_noalloc, a fixed array per GPU.
constexpr size_t kSpmvScratchBytes = size_t(64) << 20;

// The same SpMV with every device buffer in that fixed array: no device
// allocation.  Throws validation_error when the staged data exceed it.
void tiled_spmv_host_noalloc(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                             const double* x, double* y, int precision);

int device_count();

}  // namespace cuopt_amd
