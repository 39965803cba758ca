// settings.cpp -- see settings.hpp.
#include "settings.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <stdexcept>

namespace cuopt_amd {

namespace {

bool parse_int(const std::string& s, int& out)
{
  if (s.empty()) return false;
  char* end = nullptr;
  errno     = 0;
  long v    = std::strtol(s.c_str(), &end, 10);
  if (errno != 0 || *end != '\0' || v < std::numeric_limits<int>::min() || v > std::numeric_limits<int>::max())
    return false;
  out = (int)v;
  return true;
}

bool parse_float(const std::string& s, double& out)
{
  if (s.empty()) return false;
  char* end = nullptr;
  errno     = 0;
  double v  = std::strtod(s.c_str(), &end);
  if (errno == ERANGE && !std::isinf(v)) return false;
  if (*end != '\0') return false;
  out = v;
  return true;
}

bool parse_bool(std::string s, bool& out)
{
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
  if (s == "true" || s == "1") {
    out = true;
    return true;
  }
  if (s == "false" || s == "0") {
    out = false;
    return true;
  }
  return false;
}

}  // namespace

SolverSettingsRegistry::SolverSettingsRegistry()
{
  const double inf  = std::numeric_limits<double>::infinity();
  const int imax    = std::numeric_limits<int>::max();
  PdlpSettings& p   = pdlp;
  // Storage for accepted-but-unused parameters (sized up front: pointers stay valid).
  int_store_.resize(32);
  float_store_.resize(32);
  bool_store_.resize(32);
  str_store_.resize(8);
  int ni = 0, nf = 0, nb = 0, ns = 0;
  auto si = [&](int def) { int_store_[ni] = def; return &int_store_[ni++]; };
  auto sf = [&](double def) { float_store_[nf] = def; return &float_store_[nf++]; };
  auto sb = [&](bool def) { bool_store_[nb] = def; return reinterpret_cast<bool*>(&bool_store_[nb++]); };
  auto ss = [&]() { return &str_store_[ns++]; };

  // Float parameters (solver_settings.cu:61-79).
  floats_ = {
    {CUOPT_TIME_LIMIT, &p.time_limit, 0.0, inf, inf},
    {CUOPT_WORK_LIMIT, sf(inf), 0.0, inf, inf},
    {CUOPT_ABSOLUTE_DUAL_TOLERANCE, &p.abs_dual, 0.0, 1e-1, 1e-4},
    {CUOPT_RELATIVE_DUAL_TOLERANCE, &p.rel_dual, 0.0, 1e-1, 1e-4},
    {CUOPT_ABSOLUTE_PRIMAL_TOLERANCE, &p.abs_primal, 0.0, 1e-1, 1e-4},
    {CUOPT_RELATIVE_PRIMAL_TOLERANCE, &p.rel_primal, 0.0, 1e-1, 1e-4},
    {CUOPT_ABSOLUTE_GAP_TOLERANCE, &p.abs_gap, 0.0, 1e-1, 1e-4},
    {CUOPT_RELATIVE_GAP_TOLERANCE, &p.rel_gap, 0.0, 1e-1, 1e-4},
    {CUOPT_MIP_ABSOLUTE_TOLERANCE, sf(1e-4), 0.0, 1e-1, 1e-4},
    {CUOPT_MIP_RELATIVE_TOLERANCE, sf(1e-4), 0.0, 1e-1, 1e-4},
    {CUOPT_MIP_INTEGRALITY_TOLERANCE, sf(1e-5), 0.0, 1e-1, 1e-5},
    {CUOPT_MIP_ABSOLUTE_GAP, sf(1e-10), 0.0, inf, 1e-10},
    {CUOPT_MIP_RELATIVE_GAP, sf(1e-4), 0.0, 1e-1, 1e-4},
    {CUOPT_PRIMAL_INFEASIBLE_TOLERANCE, &p.primal_infeasible_tol, 0.0, 1e-1, 1e-10},
    {CUOPT_DUAL_INFEASIBLE_TOLERANCE, &p.dual_infeasible_tol, 0.0, 1e-1, 1e-10},
    {CUOPT_MIP_CUT_CHANGE_THRESHOLD, sf(1e-3), 0.0, inf, 1e-3},
    {CUOPT_MIP_CUT_MIN_ORTHOGONALITY, sf(0.5), 0.0, 1.0, 0.5},
  };
  // Integer parameters (solver_settings.cu:81-107), plus the precision switch.
  ints_ = {
    {CUOPT_ITERATION_LIMIT, &p.iteration_limit, 0, imax, imax},
    {CUOPT_PDLP_SOLVER_MODE, &p.pdlp_solver_mode, CUOPT_PDLP_SOLVER_MODE_STABLE1, CUOPT_PDLP_SOLVER_MODE_STABLE3,
     CUOPT_PDLP_SOLVER_MODE_STABLE3},
    {CUOPT_METHOD, &p.method, CUOPT_METHOD_CONCURRENT, CUOPT_METHOD_BARRIER, CUOPT_METHOD_CONCURRENT},
    {CUOPT_NUM_CPU_THREADS, &p.num_cpu_threads, -1, imax, -1},
    {CUOPT_AUGMENTED, si(-1), -1, 1, -1},
    {CUOPT_FOLDING, si(-1), -1, 1, -1},
    {CUOPT_DUALIZE, si(-1), -1, 1, -1},
    {CUOPT_ORDERING, si(-1), -1, 1, -1},
    {CUOPT_BARRIER_DUAL_INITIAL_POINT, si(-1), -1, 1, -1},
    {CUOPT_MIP_CUT_PASSES, si(10), -1, imax, 10},
    {CUOPT_MIP_MIXED_INTEGER_ROUNDING_CUTS, si(-1), -1, 1, -1},
    {CUOPT_MIP_MIXED_INTEGER_GOMORY_CUTS, si(-1), -1, 1, -1},
    {CUOPT_MIP_KNAPSACK_CUTS, si(-1), -1, 1, -1},
    {CUOPT_MIP_STRONG_CHVATAL_GOMORY_CUTS, si(-1), -1, 1, -1},
    {CUOPT_MIP_REDUCED_COST_STRENGTHENING, si(-1), -1, imax, -1},
    {CUOPT_NUM_GPUS, &p.num_gpus, 1, 2, 1},
    {CUOPT_MIP_BATCH_PDLP_STRONG_BRANCHING, si(0), 0, 1, 0},
    {CUOPT_PRESOLVE, &p.presolve, CUOPT_PRESOLVE_DEFAULT, CUOPT_PRESOLVE_PSLP, CUOPT_PRESOLVE_DEFAULT},
    {CUOPT_MIP_DETERMINISM_MODE, si(0), 0, 1, 0},
    {CUOPT_RANDOM_SEED, &p.random_seed, -1, imax, -1},
    {CUOPT_MIP_RELIABILITY_BRANCHING, si(-1), -1, imax, -1},
    {CUOPT_AMD_PDLP_PRECISION, &p.precision, CUOPT_AMD_PRECISION_FP64, CUOPT_AMD_PRECISION_FP32_MIXED,
     CUOPT_AMD_PRECISION_FP64},
    {CUOPT_AMD_PDLP_REFINEMENT, &p.refinement, 0, 1, 1},
    {CUOPT_AMD_PER_CONSTRAINT_ABSOLUTE, &p.per_constraint_absolute, 0, 1, 0},
    {CUOPT_AMD_SPMV_AUTOTUNE, &p.spmv_autotune, 0, 1, 0},
    {CUOPT_AMD_COLUMN_ORDER, &p.column_order, -1, 1, -1},
    {CUOPT_AMD_STOP_ON_NAN, &p.stop_on_nan, 0, 1, 0},
    {CUOPT_AMD_LONG_ROW_MERGE, &p.long_row_merge, -1, 1, -1},
  };
  // Bool parameters (solver_settings.cu:109-123).
  bools_ = {
    {CUOPT_INFEASIBILITY_DETECTION, &p.infeasibility_detection, false},
    {CUOPT_STRICT_INFEASIBILITY, &p.strict_infeasibility, false},
    {CUOPT_PER_CONSTRAINT_RESIDUAL, &p.per_constraint_residual, false},
    {CUOPT_SAVE_BEST_PRIMAL_SO_FAR, &p.save_best_primal_so_far, false},
    {CUOPT_FIRST_PRIMAL_FEASIBLE, &p.first_primal_feasible, false},
    {CUOPT_MIP_SCALING, sb(true), true},
    {CUOPT_MIP_HEURISTICS_ONLY, sb(false), false},
    {CUOPT_LOG_TO_CONSOLE, &p.log_to_console, true},
    {CUOPT_CROSSOVER, &p.crossover, false},
    {CUOPT_ELIMINATE_DENSE_COLUMNS, sb(true), true},
    {CUOPT_CUDSS_DETERMINISTIC, sb(false), false},
    {CUOPT_DUAL_POSTSOLVE, &p.dual_postsolve, true},
  };
  // String parameters (solver_settings.cu:125-132).
  strs_ = {
    {CUOPT_LOG_FILE, &p.log_file},
    {CUOPT_SOLUTION_FILE, &p.solution_file},
    {CUOPT_USER_PROBLEM_FILE, &p.user_problem_file},
  };
  (void)ss;
}

void SolverSettingsRegistry::set_from_string(const std::string& name, const std::string& value)
{
  for (auto& e : ints_)
    if (e.name == name) {
      int v;
      if (!parse_int(value, v)) throw std::invalid_argument("Parameter " + name + " value " + value + " is not an integer");
      if (v < e.lo || v > e.hi) throw std::invalid_argument("Parameter " + name + " value " + value + " out of range");
      *e.ptr = v;
      return;
    }
  for (auto& e : floats_)
    if (e.name == name) {
      double v;
      if (!parse_float(value, v)) throw std::invalid_argument("Parameter " + name + " value " + value + " is not a float");
      if (v < e.lo || v > e.hi) throw std::invalid_argument("Parameter " + name + " value " + value + " out of range");
      *e.ptr = v;
      return;
    }
  for (auto& e : bools_)
    if (e.name == name) {
      bool v;
      if (!parse_bool(value, v)) throw std::invalid_argument("Parameter " + name + " value " + value + " must be true or false");
      *e.ptr = v;
      return;
    }
  for (auto& e : strs_)
    if (e.name == name) {
      *e.ptr = value;
      return;
    }
  throw std::invalid_argument("Parameter " + name + " not found");
}

void SolverSettingsRegistry::set_int(const std::string& name, int value)
{
  for (auto& e : ints_)
    if (e.name == name) {
      if (value < e.lo || value > e.hi) throw std::invalid_argument("Parameter " + name + " out of range");
      *e.ptr = value;
      return;
    }
  throw std::invalid_argument("Parameter " + name + " not found");
}

void SolverSettingsRegistry::set_float(const std::string& name, double value)
{
  for (auto& e : floats_)
    if (e.name == name) {
      if (value < e.lo || value > e.hi) throw std::invalid_argument("Parameter " + name + " out of range");
      *e.ptr = value;
      return;
    }
  throw std::invalid_argument("Parameter " + name + " not found");
}

void SolverSettingsRegistry::set_bool(const std::string& name, bool value)
{
  for (auto& e : bools_)
    if (e.name == name) {
      *e.ptr = value;
      return;
    }
  throw std::invalid_argument("Parameter " + name + " not found");
}

int SolverSettingsRegistry::get_int(const std::string& name) const
{
  for (auto& e : ints_)
    if (e.name == name) return *e.ptr;
  throw std::invalid_argument("Parameter " + name + " not found");
}

double SolverSettingsRegistry::get_float(const std::string& name) const
{
  for (auto& e : floats_)
    if (e.name == name) return *e.ptr;
  throw std::invalid_argument("Parameter " + name + " not found");
}

bool SolverSettingsRegistry::get_bool(const std::string& name) const
{
  for (auto& e : bools_)
    if (e.name == name) return *e.ptr;
  throw std::invalid_argument("Parameter " + name + " not found");
}

std::string SolverSettingsRegistry::get_as_string(const std::string& name) const
{
  for (auto& e : ints_)
    if (e.name == name) return std::to_string(*e.ptr);
  for (auto& e : floats_)
    if (e.name == name) return std::to_string(*e.ptr);
  for (auto& e : bools_)
    if (e.name == name) return *e.ptr ? "true" : "false";
  for (auto& e : strs_)
    if (e.name == name) return *e.ptr;
  throw std::invalid_argument("Parameter " + name + " not found");
}

}  // namespace cuopt_amd
