// cuopt_c.cpp -- the cuOpt C ABI (LP subset) over the MI355X PDLP engine.
//
// Handle and error semantics follow cpp/src/pdlp/cuopt_c.cpp of the reference:
// null arguments give CUOPT_INVALID_ARGUMENT, unknown parameters give
// CUOPT_INVALID_ARGUMENT, cuOptSolve returns the solution's error status and
// solver exceptions never cross the ABI.  Create copies the caller's host
// arrays; getters fill caller-sized host buffers; destroy nulls the handle.
//
// Documented differences (see the solve entry): MIP problems, and the
// dual-simplex / barrier methods, are rejected with CUOPT_INVALID_ARGUMENT;
// every PDLP solver mode runs the Stable3 method; every presolve value but Off
// runs this engine's own presolve (presolve.hpp).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <new>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "../../include/cuopt_amd_ext.h"
#include "common.hpp"
#include "batch_solver.hpp"
#include "lp_types.hpp"
#include "mps_io.hpp"
#include "settings.hpp"
#include "batch_size.hpp"
#include "unrolled.hpp"

using namespace cuopt_amd;

namespace {

struct ProblemHandle {
  LpHost lp;
};

struct SettingsHandle {
  SolverSettingsRegistry reg;
  // MIP-only settings, kept as the reference keeps them; an LP solve ignores them.
  std::vector<std::vector<double>> mip_starts;
  cuOptMIPGetSolutionCallback mip_get_cb = nullptr;
  cuOptMIPSetSolutionCallback mip_set_cb = nullptr;
  void* mip_get_data = nullptr;
  void* mip_set_data = nullptr;
};

struct SolutionHandle {
  SolveResult res;
  int m = 0, n = 0;
};

ProblemHandle* as_problem(cuOptOptimizationProblem p) { return static_cast<ProblemHandle*>(p); }
SettingsHandle* as_settings(cuOptSolverSettings s) { return static_cast<SettingsHandle*>(s); }
SolutionHandle* as_solution(cuOptSolution s) { return static_cast<SolutionHandle*>(s); }

// Copies and validates the shared part of both create entry points.
cuopt_int_t fill_common(LpHost& lp,
                        cuopt_int_t m,
                        cuopt_int_t n,
                        cuopt_int_t sense,
                        cuopt_float_t offset,
                        const cuopt_float_t* c,
                        const cuopt_int_t* ptr,
                        const cuopt_int_t* idx,
                        const cuopt_float_t* val,
                        const cuopt_float_t* lb,
                        const cuopt_float_t* ub,
                        const char* vtypes)
{
  if (m < 0 || n < 0) return CUOPT_INVALID_ARGUMENT;
  if (ptr[0] != 0) return CUOPT_INVALID_ARGUMENT;
  for (int i = 0; i < m; ++i)
    if (ptr[i + 1] < ptr[i]) return CUOPT_INVALID_ARGUMENT;
  const int nnz = ptr[m];
  for (int k = 0; k < nnz; ++k)
    if (idx[k] < 0 || idx[k] >= n) return CUOPT_INVALID_ARGUMENT;
  lp.m        = m;
  lp.n        = n;
  lp.maximize = (sense == CUOPT_MAXIMIZE);
  lp.offset   = offset;
  lp.c.assign(c, c + n);
  lp.rowptr.assign(ptr, ptr + m + 1);
  lp.colidx.assign(idx, idx + nnz);
  lp.vals.assign(val, val + nnz);
  lp.lb.assign(lb, lb + n);
  lp.ub.assign(ub, ub + n);
  lp.var_types.resize(n);
  for (int j = 0; j < n; ++j) lp.var_types[j] = vtypes[j] == CUOPT_INTEGER ? CUOPT_INTEGER : CUOPT_CONTINUOUS;
  return CUOPT_SUCCESS;
}

bool is_mip(const LpHost& lp)
{
  for (char t : lp.var_types)
    if (t == CUOPT_INTEGER) return true;
  return false;
}

const char* status_string(int s)
{
  switch (s) {
    case CUOPT_TERIMINATION_STATUS_OPTIMAL: return "Optimal";
    case CUOPT_TERIMINATION_STATUS_INFEASIBLE: return "Primal Infeasible";
    case CUOPT_TERIMINATION_STATUS_UNBOUNDED: return "Dual Infeasible";
    case CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT: return "Iteration Limit";
    case CUOPT_TERIMINATION_STATUS_TIME_LIMIT: return "Time Limit";
    case CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR: return "A numerical error was encountered.";
    case CUOPT_TERIMINATION_STATUS_PRIMAL_FEASIBLE: return "Primal Feasible";
    case CUOPT_TERIMINATION_STATUS_CONCURRENT_LIMIT: return "Concurrent Limit";
    default: return "No Termination";
  }
}

template <typename T>
cuopt_int_t copy_out(const std::vector<T>& v, T* out)
{
  if (out == nullptr) return CUOPT_INVALID_ARGUMENT;
  if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(T));
  return CUOPT_SUCCESS;
}

}  // namespace

extern "C" {

int8_t cuOptGetFloatSize(void) { return sizeof(cuopt_float_t); }
int8_t cuOptGetIntSize(void) { return sizeof(cuopt_int_t); }

cuopt_int_t cuOptGetVersion(cuopt_int_t* major, cuopt_int_t* minor, cuopt_int_t* patch)
{
  if (!major || !minor || !patch) return CUOPT_INVALID_ARGUMENT;
  *major = 0;
  *minor = 1;
  *patch = 0;
  return CUOPT_SUCCESS;
}

// cuOptReadProblem (cpp/src/pdlp/cuopt_c.cpp:99-123): open failures give
// CUOPT_MPS_FILE_ERROR, malformed content CUOPT_MPS_PARSE_ERROR, and the handle
// is nulled on error.
cuopt_int_t cuOptReadProblem(const char* filename, cuOptOptimizationProblem* problem_ptr)
{
  if (!problem_ptr) return CUOPT_INVALID_ARGUMENT;
  *problem_ptr = nullptr;
  if (!filename) return CUOPT_INVALID_ARGUMENT;
  try {
    auto* h = new ProblemHandle();
    try {
      h->lp = read_mps(filename);
    } catch (...) {
      delete h;
      throw;
    }
    *problem_ptr = h;
  } catch (const mps_file_error&) {
    return CUOPT_MPS_FILE_ERROR;
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (const std::exception&) {
    return CUOPT_MPS_PARSE_ERROR;
  }
  return CUOPT_SUCCESS;
}

// cuOptWriteProblem (cpp/src/pdlp/cuopt_c.cpp:125-143).
cuopt_int_t cuOptWriteProblem(cuOptOptimizationProblem problem, const char* filename, cuopt_int_t format)
{
  if (!problem || !filename || filename[0] == '\0' || format != CUOPT_FILE_FORMAT_MPS)
    return CUOPT_INVALID_ARGUMENT;
  try {
    write_mps(as_problem(problem)->lp, filename);
  } catch (...) {
    return CUOPT_MPS_FILE_ERROR;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptCreateProblem(cuopt_int_t num_constraints,
                               cuopt_int_t num_variables,
                               cuopt_int_t objective_sense,
                               cuopt_float_t objective_offset,
                               const cuopt_float_t* objective_coefficients,
                               const cuopt_int_t* row_offsets,
                               const cuopt_int_t* column_indices,
                               const cuopt_float_t* values,
                               const char* constraint_sense,
                               const cuopt_float_t* rhs,
                               const cuopt_float_t* lower_bounds,
                               const cuopt_float_t* upper_bounds,
                               const char* variable_types,
                               cuOptOptimizationProblem* problem_ptr)
{
  if (!problem_ptr || !objective_coefficients || !row_offsets || !column_indices || !values || !constraint_sense ||
      !rhs || !lower_bounds || !upper_bounds || !variable_types)
    return CUOPT_INVALID_ARGUMENT;
  try {
    auto* h = new ProblemHandle();
    cuopt_int_t rc = fill_common(h->lp, num_constraints, num_variables, objective_sense, objective_offset,
                                 objective_coefficients, row_offsets, column_indices, values, lower_bounds,
                                 upper_bounds, variable_types);
    if (rc != CUOPT_SUCCESS) {
      delete h;
      return rc;
    }
    for (int i = 0; i < num_constraints; ++i) {
      const char s = constraint_sense[i];
      if (s != CUOPT_EQUAL && s != CUOPT_LESS_THAN && s != CUOPT_GREATER_THAN) {
        delete h;
        return CUOPT_INVALID_ARGUMENT;
      }
    }
    h->lp.ranged = false;
    h->lp.sense.assign(constraint_sense, constraint_sense + num_constraints);
    h->lp.rhs.assign(rhs, rhs + num_constraints);
    *problem_ptr = h;
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptCreateRangedProblem(cuopt_int_t num_constraints,
                                     cuopt_int_t num_variables,
                                     cuopt_int_t objective_sense,
                                     cuopt_float_t objective_offset,
                                     const cuopt_float_t* objective_coefficients,
                                     const cuopt_int_t* row_offsets,
                                     const cuopt_int_t* column_indices,
                                     const cuopt_float_t* values,
                                     const cuopt_float_t* constraint_lower_bounds,
                                     const cuopt_float_t* constraint_upper_bounds,
                                     const cuopt_float_t* variable_lower_bounds,
                                     const cuopt_float_t* variable_upper_bounds,
                                     const char* variable_types,
                                     cuOptOptimizationProblem* problem_ptr)
{
  if (!problem_ptr || !objective_coefficients || !row_offsets || !column_indices || !values ||
      !constraint_lower_bounds || !constraint_upper_bounds || !variable_lower_bounds || !variable_upper_bounds ||
      !variable_types)
    return CUOPT_INVALID_ARGUMENT;
  try {
    auto* h = new ProblemHandle();
    cuopt_int_t rc = fill_common(h->lp, num_constraints, num_variables, objective_sense, objective_offset,
                                 objective_coefficients, row_offsets, column_indices, values, variable_lower_bounds,
                                 variable_upper_bounds, variable_types);
    if (rc != CUOPT_SUCCESS) {
      delete h;
      return rc;
    }
    h->lp.ranged = true;
    h->lp.lc.assign(constraint_lower_bounds, constraint_lower_bounds + num_constraints);
    h->lp.uc.assign(constraint_upper_bounds, constraint_upper_bounds + num_constraints);
    *problem_ptr = h;
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

namespace {
// Nonzeros of a quadratic objective matrix given in CSR over n rows, or -1
// when the arguments cannot describe one.
long long quadratic_nnz(cuopt_int_t n, const cuopt_int_t* q_ptr)
{
  if (n < 0 || !q_ptr || q_ptr[0] != 0) return -1;
  return q_ptr[n] < 0 ? -1 : q_ptr[n];
}
}  // namespace

cuopt_int_t cuOptCreateQuadraticProblem(cuopt_int_t num_constraints,
                                        cuopt_int_t num_variables,
                                        cuopt_int_t objective_sense,
                                        cuopt_float_t objective_offset,
                                        const cuopt_float_t* objective_coefficients,
                                        const cuopt_int_t* q_ptr,
                                        const cuopt_int_t* q_idx,
                                        const cuopt_float_t* q_val,
                                        const cuopt_int_t* row_offsets,
                                        const cuopt_int_t* column_indices,
                                        const cuopt_float_t* values,
                                        const char* constraint_sense,
                                        const cuopt_float_t* rhs,
                                        const cuopt_float_t* lower_bounds,
                                        const cuopt_float_t* upper_bounds,
                                        cuOptOptimizationProblem* problem_ptr)
{
  (void)q_idx;
  (void)q_val;
  if (!problem_ptr) return CUOPT_INVALID_ARGUMENT;
  *problem_ptr = nullptr;
  if (quadratic_nnz(num_variables, q_ptr) != 0) return CUOPT_INVALID_ARGUMENT;  // a QP: barrier, not served here
  const std::vector<char> types(std::max(num_variables, 1), CUOPT_CONTINUOUS);
  return cuOptCreateProblem(num_constraints, num_variables, objective_sense, objective_offset, objective_coefficients,
                            row_offsets, column_indices, values, constraint_sense, rhs, lower_bounds, upper_bounds,
                            types.data(), problem_ptr);
}

cuopt_int_t cuOptCreateQuadraticRangedProblem(cuopt_int_t num_constraints,
                                              cuopt_int_t num_variables,
                                              cuopt_int_t objective_sense,
                                              cuopt_float_t objective_offset,
                                              const cuopt_float_t* objective_coefficients,
                                              const cuopt_int_t* q_ptr,
                                              const cuopt_int_t* q_idx,
                                              const cuopt_float_t* q_val,
                                              const cuopt_int_t* row_offsets,
                                              const cuopt_int_t* column_indices,
                                              const cuopt_float_t* values,
                                              const cuopt_float_t* constraint_lower_bounds,
                                              const cuopt_float_t* constraint_upper_bounds,
                                              const cuopt_float_t* variable_lower_bounds,
                                              const cuopt_float_t* variable_upper_bounds,
                                              cuOptOptimizationProblem* problem_ptr)
{
  (void)q_idx;
  (void)q_val;
  if (!problem_ptr) return CUOPT_INVALID_ARGUMENT;
  *problem_ptr = nullptr;
  if (quadratic_nnz(num_variables, q_ptr) != 0) return CUOPT_INVALID_ARGUMENT;
  const std::vector<char> types(std::max(num_variables, 1), CUOPT_CONTINUOUS);
  return cuOptCreateRangedProblem(num_constraints, num_variables, objective_sense, objective_offset,
                                  objective_coefficients, row_offsets, column_indices, values, constraint_lower_bounds,
                                  constraint_upper_bounds, variable_lower_bounds, variable_upper_bounds, types.data(),
                                  problem_ptr);
}

void cuOptDestroyProblem(cuOptOptimizationProblem* problem_ptr)
{
  if (!problem_ptr || !*problem_ptr) return;
  delete as_problem(*problem_ptr);
  *problem_ptr = nullptr;
}

#define CUOPT_PROBLEM_GETTER_PRELUDE(out) \
  if (!problem || !(out)) return CUOPT_INVALID_ARGUMENT; \
  const LpHost& lp = as_problem(problem)->lp;

cuopt_int_t cuOptGetNumConstraints(cuOptOptimizationProblem problem, cuopt_int_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  *out = lp.m;
  return CUOPT_SUCCESS;
}
cuopt_int_t cuOptGetNumVariables(cuOptOptimizationProblem problem, cuopt_int_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  *out = lp.n;
  return CUOPT_SUCCESS;
}
cuopt_int_t cuOptGetObjectiveSense(cuOptOptimizationProblem problem, cuopt_int_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  *out = lp.maximize ? CUOPT_MAXIMIZE : CUOPT_MINIMIZE;
  return CUOPT_SUCCESS;
}
cuopt_int_t cuOptGetObjectiveOffset(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  *out = lp.offset;
  return CUOPT_SUCCESS;
}
cuopt_int_t cuOptGetObjectiveCoefficients(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  return copy_out(lp.c, out);
}
cuopt_int_t cuOptGetNumNonZeros(cuOptOptimizationProblem problem, cuopt_int_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  *out = lp.rowptr[lp.m];
  return CUOPT_SUCCESS;
}
cuopt_int_t cuOptGetConstraintMatrix(cuOptOptimizationProblem problem,
                                     cuopt_int_t* offsets,
                                     cuopt_int_t* indices,
                                     cuopt_float_t* values)
{
  if (!problem || !offsets || !indices || !values) return CUOPT_INVALID_ARGUMENT;
  const LpHost& lp = as_problem(problem)->lp;
  copy_out(lp.rowptr, offsets);
  copy_out(lp.colidx, indices);
  copy_out(lp.vals, values);
  return CUOPT_SUCCESS;
}
// A problem holds the row-type form (created by cuOptCreateProblem), the ranged
// form (cuOptCreateRangedProblem), or both (read from MPS); a getter of a form
// the problem does not hold returns CUOPT_INVALID_ARGUMENT.
cuopt_int_t cuOptGetConstraintSense(cuOptOptimizationProblem problem, char* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  if (lp.m > 0 && lp.sense.empty()) return CUOPT_INVALID_ARGUMENT;
  return copy_out(lp.sense, out);
}
cuopt_int_t cuOptGetConstraintRightHandSide(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  if (lp.m > 0 && lp.rhs.empty()) return CUOPT_INVALID_ARGUMENT;
  return copy_out(lp.rhs, out);
}
cuopt_int_t cuOptGetConstraintLowerBounds(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  if (lp.m > 0 && lp.lc.empty()) return CUOPT_INVALID_ARGUMENT;
  return copy_out(lp.lc, out);
}
cuopt_int_t cuOptGetConstraintUpperBounds(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  if (lp.m > 0 && lp.uc.empty()) return CUOPT_INVALID_ARGUMENT;
  return copy_out(lp.uc, out);
}
cuopt_int_t cuOptGetVariableLowerBounds(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  return copy_out(lp.lb, out);
}
cuopt_int_t cuOptGetVariableUpperBounds(cuOptOptimizationProblem problem, cuopt_float_t* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  return copy_out(lp.ub, out);
}
cuopt_int_t cuOptGetVariableTypes(cuOptOptimizationProblem problem, char* out)
{
  CUOPT_PROBLEM_GETTER_PRELUDE(out)
  return copy_out(lp.var_types, out);
}

cuopt_int_t cuOptCreateSolverSettings(cuOptSolverSettings* settings_ptr)
{
  if (!settings_ptr) return CUOPT_INVALID_ARGUMENT;
  try {
    *settings_ptr = new SettingsHandle();
  } catch (...) {
    return CUOPT_OUT_OF_MEMORY;
  }
  return CUOPT_SUCCESS;
}

void cuOptDestroySolverSettings(cuOptSolverSettings* settings_ptr)
{
  if (!settings_ptr || !*settings_ptr) return;
  delete as_settings(*settings_ptr);
  *settings_ptr = nullptr;
}

cuopt_int_t cuOptSetParameter(cuOptSolverSettings settings, const char* name, const char* value)
{
  if (!settings || !name || !value) return CUOPT_INVALID_ARGUMENT;
  try {
    as_settings(settings)->reg.set_from_string(name, value);
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetParameter(cuOptSolverSettings settings, const char* name, cuopt_int_t size, char* value)
{
  if (!settings || !name || !value || size <= 0) return CUOPT_INVALID_ARGUMENT;
  try {
    const std::string s = as_settings(settings)->reg.get_as_string(name);
    std::snprintf(value, size, "%s", s.c_str());
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSetIntegerParameter(cuOptSolverSettings settings, const char* name, cuopt_int_t value)
{
  if (!settings || !name) return CUOPT_INVALID_ARGUMENT;
  auto& reg = as_settings(settings)->reg;
  try {
    reg.set_int(name, value);
  } catch (...) {
    try {  // an integer call may target a boolean parameter (cuopt_c.cpp:682-690)
      reg.set_bool(name, value != 0);
    } catch (...) {
      return CUOPT_INVALID_ARGUMENT;
    }
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetIntegerParameter(cuOptSolverSettings settings, const char* name, cuopt_int_t* value)
{
  if (!settings || !name || !value) return CUOPT_INVALID_ARGUMENT;
  auto& reg = as_settings(settings)->reg;
  try {
    *value = reg.get_int(name);
  } catch (...) {
    try {
      *value = reg.get_bool(name) ? 1 : 0;
    } catch (...) {
      return CUOPT_INVALID_ARGUMENT;
    }
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSetFloatParameter(cuOptSolverSettings settings, const char* name, cuopt_float_t value)
{
  if (!settings || !name) return CUOPT_INVALID_ARGUMENT;
  try {
    as_settings(settings)->reg.set_float(name, value);
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetFloatParameter(cuOptSolverSettings settings, const char* name, cuopt_float_t* value)
{
  if (!settings || !name || !value) return CUOPT_INVALID_ARGUMENT;
  try {
    *value = as_settings(settings)->reg.get_float(name);
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSetInitialPrimalSolution(cuOptSolverSettings settings, const cuopt_float_t* x, cuopt_int_t n)
{
  if (!settings || !x || n <= 0) return CUOPT_INVALID_ARGUMENT;
  as_settings(settings)->reg.pdlp.initial_primal.assign(x, x + n);
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSetInitialDualSolution(cuOptSolverSettings settings, const cuopt_float_t* y, cuopt_int_t m)
{
  if (!settings || !y || m <= 0) return CUOPT_INVALID_ARGUMENT;
  as_settings(settings)->reg.pdlp.initial_dual.assign(y, y + m);
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSetMIPGetSolutionCallback(cuOptSolverSettings settings, cuOptMIPGetSolutionCallback callback,
                                           void* user_data)
{
  if (!settings || !callback) return CUOPT_INVALID_ARGUMENT;
  as_settings(settings)->mip_get_cb   = callback;
  as_settings(settings)->mip_get_data = user_data;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSetMIPSetSolutionCallback(cuOptSolverSettings settings, cuOptMIPSetSolutionCallback callback,
                                           void* user_data)
{
  if (!settings || !callback) return CUOPT_INVALID_ARGUMENT;
  as_settings(settings)->mip_set_cb   = callback;
  as_settings(settings)->mip_set_data = user_data;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAddMIPStart(cuOptSolverSettings settings, const cuopt_float_t* solution, cuopt_int_t num_variables)
{
  if (!settings || !solution || num_variables <= 0) return CUOPT_INVALID_ARGUMENT;
  try {
    as_settings(settings)->mip_starts.emplace_back(solution, solution + num_variables);
  } catch (...) {
    return CUOPT_INVALID_ARGUMENT;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptIsMIP(cuOptOptimizationProblem problem, cuopt_int_t* is_mip_ptr)
{
  if (!problem || !is_mip_ptr) return CUOPT_INVALID_ARGUMENT;
  *is_mip_ptr = is_mip(as_problem(problem)->lp) ? 1 : 0;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptSolve(cuOptOptimizationProblem problem, cuOptSolverSettings settings, cuOptSolution* solution_ptr)
{
  if (!problem || !settings || !solution_ptr) return CUOPT_INVALID_ARGUMENT;
  const LpHost& lp         = as_problem(problem)->lp;
  const PdlpSettings& conf = as_settings(settings)->reg.pdlp;
  SolutionHandle* sol      = nullptr;
  try {
    sol = new SolutionHandle();
  } catch (...) {
    return CUOPT_OUT_OF_MEMORY;
  }
  sol->m = lp.m;
  sol->n = lp.n;
  SolveResult& R = sol->res;
  auto fail = [&](int code, const std::string& msg) {
    R.error         = code;
    R.error_message = msg;
    R.status        = CUOPT_TERIMINATION_STATUS_NO_TERMINATION;
  };
  if (is_mip(lp)) {
    fail(CUOPT_INVALID_ARGUMENT, "MIP problems are not handled by this LP engine");
  } else if (conf.method == CUOPT_METHOD_DUAL_SIMPLEX || conf.method == CUOPT_METHOD_BARRIER) {
    fail(CUOPT_INVALID_ARGUMENT, "only the PDLP method (or concurrent, which runs PDLP) is available");
  } else if (!conf.initial_primal.empty() && (int)conf.initial_primal.size() != lp.n) {
    fail(CUOPT_VALIDATION_ERROR, "initial primal solution size does not match the number of variables");
  } else if (!conf.initial_dual.empty() && (int)conf.initial_dual.size() != lp.m) {
    fail(CUOPT_VALIDATION_ERROR, "initial dual solution size does not match the number of constraints");
  } else {
    try {
      R = solve_lp(lp, conf);
      // a problem settled before the PDHG loop (presolve, closed form) took no steps
      const double steps = R.info[CUOPT_AMD_INFO_STEPS];
      const int iters    = std::isfinite(steps) ? (int)steps : 0;
      if (conf.log_to_console) {
        std::printf("Status: %s   Objective: %.8e  Iterations: %d  Time: %.3fs\n", status_string(R.status),
                    R.primal_objective, iters, R.solve_time);
        std::fflush(stdout);
      }
      // log_file: the same line, appended to the file (solve.cu:1136, logger.cpp:149-172)
      if (!conf.log_file.empty()) {
        if (FILE* f = std::fopen(conf.log_file.c_str(), "a")) {
          std::fprintf(f, "Status: %s   Objective: %.8e  Iterations: %d  Time: %.3fs\n", status_string(R.status),
                       R.primal_objective, iters, R.solve_time);
          std::fclose(f);
        }
      }
    } catch (const validation_error& e) {
      fail(CUOPT_VALIDATION_ERROR, e.what());
    } catch (const std::bad_alloc& e) {
      fail(CUOPT_OUT_OF_MEMORY, e.what());
    } catch (const hip_error& e) {
      const std::string msg = e.what();
      fail(msg.find("out of memory") != std::string::npos ? CUOPT_OUT_OF_MEMORY : CUOPT_RUNTIME_ERROR, msg);
    } catch (const std::exception& e) {
      fail(CUOPT_RUNTIME_ERROR, e.what());
    }
  }
  if (R.x.size() != (size_t)lp.n) R.x.assign(lp.n, 0.0);
  if (R.y.size() != (size_t)lp.m) R.y.assign(lp.m, 0.0);
  if (R.rc.size() != (size_t)lp.n) R.rc.assign(lp.n, 0.0);
  *solution_ptr = sol;
  return R.error;
}

cuopt_int_t cuOptAmdBatchSolve(cuopt_int_t num_problems,
                               const cuOptOptimizationProblem* problems,
                               cuOptSolverSettings settings,
                               cuOptSolution* solutions)
{
  const auto t_call = std::chrono::steady_clock::now();
  if (num_problems < 0 || (num_problems > 0 && (!problems || !solutions)) || !settings) return CUOPT_INVALID_ARGUMENT;
  for (int i = 0; i < num_problems; ++i) {
    solutions[i] = nullptr;
    if (!problems[i]) return CUOPT_INVALID_ARGUMENT;
  }
  const PdlpSettings& conf = as_settings(settings)->reg.pdlp;
  std::vector<SolutionHandle*> sols(num_problems, nullptr);
  try {
    for (auto& s : sols) s = new SolutionHandle();
  } catch (...) {
    for (auto* s : sols) delete s;
    return CUOPT_OUT_OF_MEMORY;
  }
  std::vector<const LpHost*> lps;
  std::vector<int> where;  // batch index -> problem index
  int first_error = CUOPT_SUCCESS;
  auto fail = [&](SolveResult& R, int code, const std::string& msg) {
    R.error         = code;
    R.error_message = msg;
    R.status        = CUOPT_TERIMINATION_STATUS_NO_TERMINATION;
    if (first_error == CUOPT_SUCCESS) first_error = code;
  };
  for (int i = 0; i < num_problems; ++i) {
    const LpHost& lp = as_problem(problems[i])->lp;
    sols[i]->m       = lp.m;
    sols[i]->n       = lp.n;
    if (is_mip(lp))
      fail(sols[i]->res, CUOPT_INVALID_ARGUMENT, "MIP problems are not handled by this LP engine");
    else if (conf.method == CUOPT_METHOD_DUAL_SIMPLEX || conf.method == CUOPT_METHOD_BARRIER)
      fail(sols[i]->res, CUOPT_INVALID_ARGUMENT, "only the PDLP method (or concurrent, which runs PDLP) is available");
    else if (conf.per_constraint_residual || conf.first_primal_feasible || conf.save_best_primal_so_far)
      // as the reference's batch mode (convergence_information.cu:393-395, pdlp.cu:931-933)
      fail(sols[i]->res, CUOPT_VALIDATION_ERROR,
           "per_constraint_residual, first_primal_feasible and save_best_primal_so_far are not supported in batch "
           "mode");
    else {
      lps.push_back(&lp);
      where.push_back(i);
    }
  }
  try {
    std::vector<SolveResult> res = batch_solve(lps, conf);
    for (size_t b = 0; b < res.size(); ++b) {
      if (res[b].error != CUOPT_SUCCESS && first_error == CUOPT_SUCCESS) first_error = res[b].error;
      sols[where[b]]->res = std::move(res[b]);
    }
  } catch (const std::exception& e) {
    const std::string msg = e.what();
    int code              = CUOPT_RUNTIME_ERROR;
    if (dynamic_cast<const validation_error*>(&e)) code = CUOPT_VALIDATION_ERROR;
    if (dynamic_cast<const std::bad_alloc*>(&e) || msg.find("out of memory") != std::string::npos)
      code = CUOPT_OUT_OF_MEMORY;
    for (int i : where) fail(sols[i]->res, code, msg);
  }
  // Every solution reports the whole call's wall time: checks, the device solve,
  // result assembly and the release of the working memory.
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_call).count();
  for (int i = 0; i < num_problems; ++i) {
    SolveResult& R = sols[i]->res;
    if (R.x.size() != (size_t)sols[i]->n) R.x.assign(sols[i]->n, 0.0);
    if (R.y.size() != (size_t)sols[i]->m) R.y.assign(sols[i]->m, 0.0);
    if (R.rc.size() != (size_t)sols[i]->n) R.rc.assign(sols[i]->n, 0.0);
    R.solve_time                     = secs;
    R.info[CUOPT_AMD_INFO_SOLVE_TIME] = secs;
    solutions[i] = sols[i];
  }
  return first_error;
}

void cuOptDestroySolution(cuOptSolution* solution_ptr)
{
  if (!solution_ptr || !*solution_ptr) return;
  delete as_solution(*solution_ptr);
  *solution_ptr = nullptr;
}

cuopt_int_t cuOptGetTerminationStatus(cuOptSolution solution, cuopt_int_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  *out = as_solution(solution)->res.status;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetErrorStatus(cuOptSolution solution, cuopt_int_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  *out = as_solution(solution)->res.error;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetErrorString(cuOptSolution solution, char* out, cuopt_int_t size)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  std::snprintf(out, size, "%s", as_solution(solution)->res.error_message.c_str());
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetPrimalSolution(cuOptSolution solution, cuopt_float_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  return copy_out(as_solution(solution)->res.x, out);
}

cuopt_int_t cuOptGetObjectiveValue(cuOptSolution solution, cuopt_float_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  *out = as_solution(solution)->res.primal_objective;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetSolveTime(cuOptSolution solution, cuopt_float_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  *out = as_solution(solution)->res.solve_time;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetDualSolution(cuOptSolution solution, cuopt_float_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  return copy_out(as_solution(solution)->res.y, out);
}

namespace {
cuopt_int_t copy_name(const std::string& s, char* buf, cuopt_int_t size)
{
  if (!buf || size < 1) return CUOPT_INVALID_ARGUMENT;
  const size_t n = std::min<size_t>(s.size(), (size_t)size - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = '\0';
  return CUOPT_SUCCESS;
}
}  // namespace

cuopt_int_t cuOptAmdGetVariableName(cuOptOptimizationProblem problem, cuopt_int_t index, char* buf, cuopt_int_t size)
{
  if (!problem) return CUOPT_INVALID_ARGUMENT;
  const LpHost& lp = as_problem(problem)->lp;
  if (index < 0 || index >= lp.n) return CUOPT_INVALID_ARGUMENT;
  return copy_name((size_t)index < lp.col_names.size() ? lp.col_names[index] : std::string(), buf, size);
}

cuopt_int_t cuOptAmdGetConstraintName(cuOptOptimizationProblem problem, cuopt_int_t index, char* buf, cuopt_int_t size)
{
  if (!problem) return CUOPT_INVALID_ARGUMENT;
  const LpHost& lp = as_problem(problem)->lp;
  if (index < 0 || index >= lp.m) return CUOPT_INVALID_ARGUMENT;
  return copy_name((size_t)index < lp.row_names.size() ? lp.row_names[index] : std::string(), buf, size);
}

cuopt_int_t cuOptAmdGetObjectiveName(cuOptOptimizationProblem problem, char* buf, cuopt_int_t size)
{
  if (!problem) return CUOPT_INVALID_ARGUMENT;
  return copy_name(as_problem(problem)->lp.objective_name, buf, size);
}

cuopt_int_t cuOptAmdGetProblemName(cuOptOptimizationProblem problem, char* buf, cuopt_int_t size)
{
  if (!problem) return CUOPT_INVALID_ARGUMENT;
  return copy_name(as_problem(problem)->lp.problem_name, buf, size);
}

namespace {
cuopt_int_t set_names(const char* const* names, cuopt_int_t count, int expected, std::vector<std::string>& out)
{
  if (!names || count != expected) return CUOPT_INVALID_ARGUMENT;
  std::vector<std::string> v((size_t)count);
  for (cuopt_int_t i = 0; i < count; ++i) {
    if (!names[i]) return CUOPT_INVALID_ARGUMENT;
    v[(size_t)i] = names[i];
  }
  out = std::move(v);
  return CUOPT_SUCCESS;
}
}  // namespace

cuopt_int_t cuOptAmdSetVariableNames(cuOptOptimizationProblem problem, const char* const* names, cuopt_int_t count)
{
  if (!problem) return CUOPT_INVALID_ARGUMENT;
  LpHost& lp = as_problem(problem)->lp;
  return set_names(names, count, lp.n, lp.col_names);
}

cuopt_int_t cuOptAmdSetConstraintNames(cuOptOptimizationProblem problem, const char* const* names, cuopt_int_t count)
{
  if (!problem) return CUOPT_INVALID_ARGUMENT;
  LpHost& lp = as_problem(problem)->lp;
  return set_names(names, count, lp.m, lp.row_names);
}

cuopt_int_t cuOptAmdSetObjectiveName(cuOptOptimizationProblem problem, const char* name)
{
  if (!problem || !name) return CUOPT_INVALID_ARGUMENT;
  as_problem(problem)->lp.objective_name = name;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdSetProblemName(cuOptOptimizationProblem problem, const char* name)
{
  if (!problem || !name) return CUOPT_INVALID_ARGUMENT;
  as_problem(problem)->lp.problem_name = name;
  return CUOPT_SUCCESS;
}

// Solutions of this engine are LP solutions: no MIP gap, no solution bound.
cuopt_int_t cuOptGetMIPGap(cuOptSolution solution, cuopt_float_t* mip_gap_ptr)
{
  (void)solution;
  (void)mip_gap_ptr;
  return CUOPT_INVALID_ARGUMENT;
}

cuopt_int_t cuOptGetSolutionBound(cuOptSolution solution, cuopt_float_t* solution_bound_ptr)
{
  (void)solution;
  (void)solution_bound_ptr;
  return CUOPT_INVALID_ARGUMENT;
}

cuopt_int_t cuOptGetDualObjectiveValue(cuOptSolution solution, cuopt_float_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  *out = as_solution(solution)->res.dual_objective;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptGetReducedCosts(cuOptSolution solution, cuopt_float_t* out)
{
  if (!solution || !out) return CUOPT_INVALID_ARGUMENT;
  return copy_out(as_solution(solution)->res.rc, out);
}

// ---- extensions -------------------------------------------------------------

cuopt_int_t cuOptAmdGetTerminationInfo(cuOptSolution solution, cuopt_float_t* info, cuopt_int_t n)
{
  if (!solution || !info || n < 0) return CUOPT_INVALID_ARGUMENT;
  const SolveResult& R = as_solution(solution)->res;
  for (int i = 0; i < n && i < CUOPT_AMD_INFO_COUNT; ++i) info[i] = R.info[i];
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdGetTrace(cuOptSolution solution, cuopt_float_t* rows, cuopt_int_t cap, cuopt_int_t* len)
{
  if (!solution || !len || cap < 0 || (cap > 0 && !rows)) return CUOPT_INVALID_ARGUMENT;
  const SolveResult& R = as_solution(solution)->res;
  const int nrec       = (int)(R.trace.size() / CUOPT_AMD_TRACE_WIDTH);
  *len                 = nrec;
  const int ncopy      = nrec < cap ? nrec : cap;
  if (ncopy > 0) std::memcpy(rows, R.trace.data(), (size_t)ncopy * CUOPT_AMD_TRACE_WIDTH * sizeof(double));
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdGetSolutions(cuopt_int_t num_solutions,
                                 const cuOptSolution* solutions,
                                 cuopt_float_t* store,
                                 const int64_t* offsets,
                                 cuopt_int_t* status,
                                 cuopt_int_t* error_status,
                                 cuopt_float_t* objectives)
{
  if (num_solutions < 0) return CUOPT_INVALID_ARGUMENT;
  if (num_solutions == 0) return CUOPT_SUCCESS;
  if (!solutions || !store || !offsets || !status || !error_status || !objectives) return CUOPT_INVALID_ARGUMENT;
  for (int i = 0; i < num_solutions; ++i) {
    if (!solutions[i] || offsets[i] < 0) return CUOPT_INVALID_ARGUMENT;
    const SolutionHandle* h = as_solution(solutions[i]);
    if (offsets[i + 1] - offsets[i] != 2 * (int64_t)h->n + h->m + CUOPT_AMD_INFO_COUNT) return CUOPT_INVALID_ARGUMENT;
  }
  auto one = [&](int i) {
    const SolutionHandle* h = as_solution(solutions[i]);
    const SolveResult& R    = h->res;
    double* o               = store + offsets[i];
    auto put = [&](const std::vector<double>& v, size_t len) {
      if (v.size() == len) {
        if (len) std::memcpy(o, v.data(), len * sizeof(double));
      } else {
        std::fill(o, o + len, 0.0);
      }
      o += len;
    };
    put(R.x, (size_t)h->n);
    put(R.y, (size_t)h->m);
    put(R.rc, (size_t)h->n);
    std::memcpy(o, R.info, sizeof(R.info));
    status[i]             = R.status;
    error_status[i]       = R.error;
    objectives[3 * i]     = R.primal_objective;
    objectives[3 * i + 1] = R.dual_objective;
    objectives[3 * i + 2] = R.solve_time;
  };
  // Threads only when there is enough to copy; a failed thread start copies serially.
  const int64_t bytes = offsets[num_solutions] * (int64_t)sizeof(double);
  const int nt        = bytes < (4 << 20) ? 1
                                          : std::max(1, std::min({16, num_solutions,
                                                                  (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> th;
  auto work = [&](int t) {
    for (int i = t; i < num_solutions; i += nt) one(i);
  };
  int started = 1;
  try {
    for (int t = 1; t < nt; ++t, ++started) th.emplace_back(work, t);
  } catch (...) {
  }
  work(0);
  for (auto& x : th) x.join();
  for (int t = started; t < nt; ++t) work(t);
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdSpmv(cuopt_int_t m,
                         cuopt_int_t n,
                         const cuopt_int_t* ptr,
                         const cuopt_int_t* idx,
                         const cuopt_float_t* val,
                         const cuopt_float_t* x,
                         cuopt_float_t* y,
                         cuopt_int_t precision)
{
  if (m < 0 || n < 0 || !ptr || !idx || !val || !x || !y) return CUOPT_INVALID_ARGUMENT;
  if (precision != CUOPT_AMD_PRECISION_FP64 && precision != CUOPT_AMD_PRECISION_FP32_MIXED)
    return CUOPT_INVALID_ARGUMENT;
  if (ptr[0] != 0) return CUOPT_INVALID_ARGUMENT;
  for (int i = 0; i < m; ++i)
    if (ptr[i + 1] < ptr[i]) return CUOPT_INVALID_ARGUMENT;
  for (int k = 0; k < ptr[m]; ++k)
    if (idx[k] < 0 || idx[k] >= n) return CUOPT_INVALID_ARGUMENT;
  try {
    tiled_spmv_host(m, n, ptr, idx, val, x, y, precision);
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (...) {
    return CUOPT_RUNTIME_ERROR;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdSpmvNoAlloc(cuopt_int_t m,
                                cuopt_int_t n,
                                const cuopt_int_t* ptr,
                                const cuopt_int_t* idx,
                                const cuopt_float_t* val,
                                const cuopt_float_t* x,
                                cuopt_float_t* y,
                                cuopt_int_t precision)
{
  if (m < 0 || n < 0 || !ptr || !idx || !val || !x || !y) return CUOPT_INVALID_ARGUMENT;
  if (precision != CUOPT_AMD_PRECISION_FP64 && precision != CUOPT_AMD_PRECISION_FP32_MIXED)
    return CUOPT_INVALID_ARGUMENT;
  if (ptr[0] != 0) return CUOPT_INVALID_ARGUMENT;
  for (int i = 0; i < m; ++i)
    if (ptr[i + 1] < ptr[i]) return CUOPT_INVALID_ARGUMENT;
  for (int k = 0; k < ptr[m]; ++k)
    if (idx[k] < 0 || idx[k] >= n) return CUOPT_INVALID_ARGUMENT;
  try {
    tiled_spmv_host_noalloc(m, n, ptr, idx, val, x, y, precision);
  } catch (const validation_error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return CUOPT_INVALID_ARGUMENT;
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (...) {
    return CUOPT_RUNTIME_ERROR;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdSpmvScratchBytes(size_t* bytes)
{
  if (!bytes) return CUOPT_INVALID_ARGUMENT;
  *bytes = kSpmvScratchBytes;
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdBenchmarkSteps(cuOptOptimizationProblem problem,
                                   cuOptSolverSettings settings,
                                   cuopt_int_t warmup,
                                   cuopt_int_t steps,
                                   cuopt_float_t* out,
                                   cuopt_int_t n)
{
  if (!problem || !settings || !out || warmup < 0 || steps <= 0 || n < 0) return CUOPT_INVALID_ARGUMENT;
  const LpHost& lp = as_problem(problem)->lp;
  if (is_mip(lp) || lp.m == 0) return CUOPT_INVALID_ARGUMENT;
  PdlpSettings conf  = as_settings(settings)->reg.pdlp;
  conf.bench_warmup  = warmup;
  conf.bench_steps   = steps;
  conf.iteration_limit = std::numeric_limits<int>::max();
  conf.refinement      = 0;  // the window times the plain PDHG loop
  try {
    SolveResult R = solve_lp(lp, conf);
    for (int i = 0; i < n && i < CUOPT_AMD_BENCH_COUNT; ++i) out[i] = R.bench[i];
    if (n > 6 && R.bench[3] == 0) out[6] = R.status;  // terminated before the window closed
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (const std::exception&) {
    return CUOPT_RUNTIME_ERROR;
  }
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdDeviceCount(void) { return device_count(); }

}  // extern "C"

// ---- differentiable unrolled PDHG (no reference counterpart) ----------------------

namespace {
template <class F>
cuopt_int_t guarded(F f)
{
  try {
    f();
  } catch (const validation_error&) {
    return CUOPT_VALIDATION_ERROR;
  } catch (const std::bad_alloc&) {
    return CUOPT_OUT_OF_MEMORY;
  } catch (const hip_error& e) {
    return std::string(e.what()).find("out of memory") != std::string::npos ? CUOPT_OUT_OF_MEMORY
                                                                            : CUOPT_RUNTIME_ERROR;
  } catch (...) {
    return CUOPT_RUNTIME_ERROR;
  }
  return CUOPT_SUCCESS;
}
}  // namespace

cuopt_int_t cuOptAmdUnrolledCreate(cuopt_int_t m,
                                   cuopt_int_t n,
                                   const cuopt_int_t* row_offsets,
                                   const cuopt_int_t* column_indices,
                                   const cuopt_float_t* values,
                                   const cuopt_float_t* lower_bounds,
                                   const cuopt_float_t* upper_bounds,
                                   cuopt_float_t ridge,
                                   cuopt_int_t precision,
                                   cuOptAmdUnrolled* handle_ptr)
{
  if (!handle_ptr) return CUOPT_INVALID_ARGUMENT;
  *handle_ptr = nullptr;
  if (m < 0 || n < 0 || !row_offsets || !column_indices || !values || !lower_bounds || !upper_bounds || !(ridge >= 0))
    return CUOPT_INVALID_ARGUMENT;
  if (precision != CUOPT_AMD_PRECISION_FP64 && precision != CUOPT_AMD_PRECISION_FP32_MIXED)
    return CUOPT_INVALID_ARGUMENT;
  if (row_offsets[0] != 0) return CUOPT_INVALID_ARGUMENT;
  for (int i = 0; i < m; ++i)
    if (row_offsets[i + 1] < row_offsets[i]) return CUOPT_INVALID_ARGUMENT;
  for (int k = 0; k < row_offsets[m]; ++k)
    if (column_indices[k] < 0 || column_indices[k] >= n) return CUOPT_INVALID_ARGUMENT;
  UnrolledHandle* h = nullptr;
  const cuopt_int_t rc = guarded([&] {
    h = unrolled_create(m, n, row_offsets, column_indices, values, lower_bounds, upper_bounds, ridge, precision);
  });
  *handle_ptr = h;
  return rc;
}

void cuOptAmdUnrolledDestroy(cuOptAmdUnrolled* handle_ptr)
{
  if (!handle_ptr || !*handle_ptr) return;
  unrolled_destroy(static_cast<UnrolledHandle*>(*handle_ptr));
  *handle_ptr = nullptr;
}

cuopt_int_t cuOptAmdUnrolledStepSize(cuOptAmdUnrolled handle, cuopt_float_t* eta)
{
  if (!handle || !eta) return CUOPT_INVALID_ARGUMENT;
  *eta = unrolled_step_size(static_cast<UnrolledHandle*>(handle));
  return CUOPT_SUCCESS;
}

cuopt_int_t cuOptAmdUnrolledOptimalBatchSize(cuOptAmdUnrolled handle, cuopt_int_t max_batch_size,
                                             cuopt_int_t* batch_size)
{
  if (!handle || !batch_size || max_batch_size < 1) return CUOPT_INVALID_ARGUMENT;
  return guarded(
    [&] { *batch_size = unrolled_optimal_batch_size(static_cast<UnrolledHandle*>(handle), max_batch_size); });
}

cuopt_int_t cuOptAmdBatchSizeSearch(cuopt_int_t max_batch_size,
                                    cuopt_float_t (*cost)(cuopt_int_t, void*),
                                    void* context,
                                    cuopt_int_t* batch_size)
{
  if (!cost || !batch_size || max_batch_size < 1) return CUOPT_INVALID_ARGUMENT;
  return guarded([&] { *batch_size = search_batch_size(max_batch_size, [&](int b) { return cost(b, context); }); });
}

cuopt_int_t cuOptAmdUnrolledForward(cuOptAmdUnrolled handle,
                                    cuopt_int_t batch,
                                    cuopt_int_t steps,
                                    cuopt_int_t checkpoint_every,
                                    const cuopt_float_t* c,
                                    const cuopt_float_t* lc,
                                    const cuopt_float_t* uc,
                                    cuopt_int_t bounds_per_batch,
                                    cuopt_float_t* x_out,
                                    cuopt_float_t* y_out,
                                    void* stream)
{
  if (!handle || !c || !lc || !uc || !x_out || batch <= 0 || steps < 0 || checkpoint_every <= 0)
    return CUOPT_INVALID_ARGUMENT;
  return guarded([&] {
    unrolled_forward(static_cast<UnrolledHandle*>(handle), batch, steps, checkpoint_every, c, lc, uc,
                     bounds_per_batch, x_out, y_out, stream);
  });
}

cuopt_int_t cuOptAmdUnrolledBackward(cuOptAmdUnrolled handle,
                                     const cuopt_float_t* gx,
                                     const cuopt_float_t* gy,
                                     cuopt_float_t* gc,
                                     cuopt_float_t* glc,
                                     cuopt_float_t* guc,
                                     void* stream)
{
  if (!handle || !gx || !gc || !glc || !guc) return CUOPT_INVALID_ARGUMENT;
  return guarded([&] { unrolled_backward(static_cast<UnrolledHandle*>(handle), gx, gy, gc, glc, guc, stream); });
}
