/*
 * cuopt_c.h -- C ABI of the MI355X-native PDLP LP engine (libcuopt_amd.so).
 *
 * This header restates the LP subset of the cuOpt C API so that existing
 * callers drop in unchanged: same function names, argument order and meaning,
 * same opaque handle types, same status / parameter constants.  Each
 * declaration names the reference interface it replaces
 * (cpp/include/cuopt/linear_programming/cuopt_c.h and constants.h of the
 * reference repository).
 *
 * Types follow the reference build (CUOPT_INSTANTIATE_DOUBLE 1 / INT32 1,
 * constants.h:17-20): cuopt_float_t = double, cuopt_int_t = int32_t.
 *
 * Ownership: create functions copy the caller's host arrays; getters copy into
 * caller-sized host buffers; destroy functions free and set the handle to NULL.
 * Solver exceptions never cross this boundary: they become error codes.
 */
#ifndef CUOPT_AMD_CUOPT_C_H
#define CUOPT_AMD_CUOPT_C_H

#include <stdint.h>

#ifdef __cplusplus
#include <limits>
extern "C" {
#else
#include <math.h>
#endif

/* ---- constants (constants.h:17-143) ------------------------------------- */
#define CUOPT_INSTANTIATE_FLOAT  0
#define CUOPT_INSTANTIATE_DOUBLE 1
#define CUOPT_INSTANTIATE_INT32  1
#define CUOPT_INSTANTIATE_INT64  0

/* Parameter names (constants.h:22-76).  The PDLP engine acts on the LP ones;
 * the MIP, barrier and work-limit names are accepted with the reference's
 * ranges and defaults and have no effect on an LP solve. */
#define CUOPT_ABSOLUTE_DUAL_TOLERANCE     "absolute_dual_tolerance"
#define CUOPT_RELATIVE_DUAL_TOLERANCE     "relative_dual_tolerance"
#define CUOPT_ABSOLUTE_PRIMAL_TOLERANCE   "absolute_primal_tolerance"
#define CUOPT_RELATIVE_PRIMAL_TOLERANCE   "relative_primal_tolerance"
#define CUOPT_ABSOLUTE_GAP_TOLERANCE      "absolute_gap_tolerance"
#define CUOPT_RELATIVE_GAP_TOLERANCE      "relative_gap_tolerance"
#define CUOPT_INFEASIBILITY_DETECTION     "infeasibility_detection"
#define CUOPT_STRICT_INFEASIBILITY        "strict_infeasibility"
#define CUOPT_PRIMAL_INFEASIBLE_TOLERANCE "primal_infeasible_tolerance"
#define CUOPT_DUAL_INFEASIBLE_TOLERANCE   "dual_infeasible_tolerance"
#define CUOPT_ITERATION_LIMIT             "iteration_limit"
#define CUOPT_TIME_LIMIT                  "time_limit"
#define CUOPT_WORK_LIMIT                  "work_limit"
#define CUOPT_PDLP_SOLVER_MODE            "pdlp_solver_mode"
#define CUOPT_METHOD                      "method"
#define CUOPT_PER_CONSTRAINT_RESIDUAL     "per_constraint_residual"
#define CUOPT_SAVE_BEST_PRIMAL_SO_FAR     "save_best_primal_so_far"
#define CUOPT_FIRST_PRIMAL_FEASIBLE       "first_primal_feasible"
#define CUOPT_LOG_FILE                    "log_file"
#define CUOPT_LOG_TO_CONSOLE              "log_to_console"
#define CUOPT_CROSSOVER                   "crossover"
#define CUOPT_FOLDING                     "folding"
#define CUOPT_AUGMENTED                   "augmented"
#define CUOPT_DUALIZE                     "dualize"
#define CUOPT_ORDERING                    "ordering"
#define CUOPT_BARRIER_DUAL_INITIAL_POINT  "barrier_dual_initial_point"
#define CUOPT_ELIMINATE_DENSE_COLUMNS     "eliminate_dense_columns"
#define CUOPT_CUDSS_DETERMINISTIC         "cudss_deterministic"
#define CUOPT_PRESOLVE                    "presolve"
#define CUOPT_DUAL_POSTSOLVE              "dual_postsolve"
#define CUOPT_MIP_DETERMINISM_MODE            "mip_determinism_mode"
#define CUOPT_MIP_ABSOLUTE_TOLERANCE          "mip_absolute_tolerance"
#define CUOPT_MIP_RELATIVE_TOLERANCE          "mip_relative_tolerance"
#define CUOPT_MIP_INTEGRALITY_TOLERANCE       "mip_integrality_tolerance"
#define CUOPT_MIP_ABSOLUTE_GAP                "mip_absolute_gap"
#define CUOPT_MIP_RELATIVE_GAP                "mip_relative_gap"
#define CUOPT_MIP_HEURISTICS_ONLY             "mip_heuristics_only"
#define CUOPT_MIP_SCALING                     "mip_scaling"
#define CUOPT_MIP_PRESOLVE                    "mip_presolve"
#define CUOPT_MIP_RELIABILITY_BRANCHING       "mip_reliability_branching"
#define CUOPT_MIP_CUT_PASSES                  "mip_cut_passes"
#define CUOPT_MIP_MIXED_INTEGER_ROUNDING_CUTS "mip_mixed_integer_rounding_cuts"
#define CUOPT_MIP_MIXED_INTEGER_GOMORY_CUTS   "mip_mixed_integer_gomory_cuts"
#define CUOPT_MIP_KNAPSACK_CUTS               "mip_knapsack_cuts"
#define CUOPT_MIP_STRONG_CHVATAL_GOMORY_CUTS  "mip_strong_chvatal_gomory_cuts"
#define CUOPT_MIP_REDUCED_COST_STRENGTHENING  "mip_reduced_cost_strengthening"
#define CUOPT_MIP_CUT_CHANGE_THRESHOLD        "mip_cut_change_threshold"
#define CUOPT_MIP_CUT_MIN_ORTHOGONALITY       "mip_cut_min_orthogonality"
#define CUOPT_MIP_BATCH_PDLP_STRONG_BRANCHING "mip_batch_pdlp_strong_branching"
#define CUOPT_SOLUTION_FILE               "solution_file"
#define CUOPT_NUM_CPU_THREADS             "num_cpu_threads"
#define CUOPT_NUM_GPUS                    "num_gpus"
#define CUOPT_USER_PROBLEM_FILE           "user_problem_file"
#define CUOPT_RANDOM_SEED                 "random_seed"

/* MIP determinism mode values (constants.h:78-80) */
#define CUOPT_MODE_OPPORTUNISTIC 0
#define CUOPT_MODE_DETERMINISTIC 1

#define CUOPT_TERIMINATION_STATUS_NO_TERMINATION   0
#define CUOPT_TERIMINATION_STATUS_OPTIMAL          1
#define CUOPT_TERIMINATION_STATUS_INFEASIBLE       2
#define CUOPT_TERIMINATION_STATUS_UNBOUNDED        3
#define CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT  4
#define CUOPT_TERIMINATION_STATUS_TIME_LIMIT       5
#define CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR  6
#define CUOPT_TERIMINATION_STATUS_PRIMAL_FEASIBLE  7
#define CUOPT_TERIMINATION_STATUS_FEASIBLE_FOUND   8
#define CUOPT_TERIMINATION_STATUS_CONCURRENT_LIMIT 9
#define CUOPT_TERIMINATION_STATUS_WORK_LIMIT       10

#define CUOPT_MINIMIZE 1
#define CUOPT_MAXIMIZE -1

#define CUOPT_LESS_THAN    'L'
#define CUOPT_GREATER_THAN 'G'
#define CUOPT_EQUAL        'E'

#define CUOPT_CONTINUOUS 'C'
#define CUOPT_INTEGER    'I'

#ifdef __cplusplus
#define CUOPT_INFINITY std::numeric_limits<double>::infinity()
#else
#define CUOPT_INFINITY INFINITY
#endif

#define CUOPT_PDLP_SOLVER_MODE_STABLE1     0
#define CUOPT_PDLP_SOLVER_MODE_STABLE2     1
#define CUOPT_PDLP_SOLVER_MODE_METHODICAL1 2
#define CUOPT_PDLP_SOLVER_MODE_FAST1       3
#define CUOPT_PDLP_SOLVER_MODE_STABLE3     4

#define CUOPT_METHOD_CONCURRENT   0
#define CUOPT_METHOD_PDLP         1
#define CUOPT_METHOD_DUAL_SIMPLEX 2
#define CUOPT_METHOD_BARRIER      3

#define CUOPT_FILE_FORMAT_MPS 0

#define CUOPT_SUCCESS          0
#define CUOPT_INVALID_ARGUMENT 1
#define CUOPT_MPS_FILE_ERROR   2
#define CUOPT_MPS_PARSE_ERROR  3
#define CUOPT_VALIDATION_ERROR 4
#define CUOPT_OUT_OF_MEMORY    5
#define CUOPT_RUNTIME_ERROR    6

#define CUOPT_PRESOLVE_DEFAULT -1
#define CUOPT_PRESOLVE_OFF     0
#define CUOPT_PRESOLVE_PAPILO  1
#define CUOPT_PRESOLVE_PSLP    2

/* ---- types (cuopt_c.h:25-72) -------------------------------------------- */
typedef void* cuOptOptimizationProblem;
typedef void* cuOptSolverSettings;
typedef void* cuOptSolution;
typedef double cuopt_float_t;
typedef int32_t cuopt_int_t;

/* cuopt_c.h:79 / :84 */
int8_t cuOptGetFloatSize(void);
int8_t cuOptGetIntSize(void);
/* cuopt_c.h:98 */
cuopt_int_t cuOptGetVersion(cuopt_int_t* version_major,
                            cuopt_int_t* version_minor,
                            cuopt_int_t* version_patch);

/* ---- MPS files (cuopt_c.h:112, :124) --------------------------------------- */
/* Reads a free- or fixed-format MPS file into a new problem.  Returns
 * CUOPT_MPS_FILE_ERROR when the file cannot be opened and CUOPT_MPS_PARSE_ERROR
 * for malformed content (problem_ptr is then NULL).  cuopt_c.h:102-112 */
cuopt_int_t cuOptReadProblem(const char* filename, cuOptOptimizationProblem* problem_ptr);
/* Writes the problem as free-format MPS; format must be CUOPT_FILE_FORMAT_MPS.
 * cuopt_c.h:114-126 */
cuopt_int_t cuOptWriteProblem(cuOptOptimizationProblem problem, const char* filename, cuopt_int_t format);

/* ---- problem (cuopt_c.h:171, :240, :414) ---------------------------------- */
/* minimize/maximize c^T x + offset  s.t.  A x {<=,>=,=} rhs,  lb <= x <= ub.
 * A is CSR with num_constraints+1 row offsets.  cuopt_c.h:171-184 */
cuopt_int_t cuOptCreateProblem(cuopt_int_t num_constraints,
                               cuopt_int_t num_variables,
                               cuopt_int_t objective_sense,
                               cuopt_float_t objective_offset,
                               const cuopt_float_t* objective_coefficients,
                               const cuopt_int_t* constraint_matrix_row_offsets,
                               const cuopt_int_t* constraint_matrix_column_indices,
                               const cuopt_float_t* constraint_matrix_coefficent_values,
                               const char* constraint_sense,
                               const cuopt_float_t* rhs,
                               const cuopt_float_t* lower_bounds,
                               const cuopt_float_t* upper_bounds,
                               const char* variable_types,
                               cuOptOptimizationProblem* problem_ptr);

/* bl <= A x <= bu form.  cuopt_c.h:240-253 */
cuopt_int_t cuOptCreateRangedProblem(cuopt_int_t num_constraints,
                                     cuopt_int_t num_variables,
                                     cuopt_int_t objective_sense,
                                     cuopt_float_t objective_offset,
                                     const cuopt_float_t* objective_coefficients,
                                     const cuopt_int_t* constraint_matrix_row_offsets,
                                     const cuopt_int_t* constraint_matrix_column_indices,
                                     const cuopt_float_t* constraint_matrix_coefficients,
                                     const cuopt_float_t* constraint_lower_bounds,
                                     const cuopt_float_t* constraint_upper_bounds,
                                     const cuopt_float_t* variable_lower_bounds,
                                     const cuopt_float_t* variable_upper_bounds,
                                     const char* variable_types,
                                     cuOptOptimizationProblem* problem_ptr);

/* Quadratic-objective forms (cuopt_c.h:309-325 and 391-407).  The reference
 * sends problems with a nonzero quadratic matrix to its barrier method
 * (solve.cu:1142-1152), which this LP engine does not run: they are refused
 * with CUOPT_INVALID_ARGUMENT.  An empty quadratic matrix (no nonzeros) gives
 * the plain LP, with continuous variables. */
cuopt_int_t cuOptCreateQuadraticProblem(cuopt_int_t num_constraints,
                                        cuopt_int_t num_variables,
                                        cuopt_int_t objective_sense,
                                        cuopt_float_t objective_offset,
                                        const cuopt_float_t* objective_coefficients,
                                        const cuopt_int_t* quadratic_objective_matrix_row_offsets,
                                        const cuopt_int_t* quadratic_objective_matrix_column_indices,
                                        const cuopt_float_t* quadratic_objective_matrix_coefficent_values,
                                        const cuopt_int_t* constraint_matrix_row_offsets,
                                        const cuopt_int_t* constraint_matrix_column_indices,
                                        const cuopt_float_t* constraint_matrix_coefficent_values,
                                        const char* constraint_sense,
                                        const cuopt_float_t* rhs,
                                        const cuopt_float_t* lower_bounds,
                                        const cuopt_float_t* upper_bounds,
                                        cuOptOptimizationProblem* problem_ptr);
cuopt_int_t cuOptCreateQuadraticRangedProblem(cuopt_int_t num_constraints,
                                              cuopt_int_t num_variables,
                                              cuopt_int_t objective_sense,
                                              cuopt_float_t objective_offset,
                                              const cuopt_float_t* objective_coefficients,
                                              const cuopt_int_t* quadratic_objective_matrix_row_offsets,
                                              const cuopt_int_t* quadratic_objective_matrix_column_indices,
                                              const cuopt_float_t* quadratic_objective_matrix_coefficent_values,
                                              const cuopt_int_t* constraint_matrix_row_offsets,
                                              const cuopt_int_t* constraint_matrix_column_indices,
                                              const cuopt_float_t* constraint_matrix_coefficients,
                                              const cuopt_float_t* constraint_lower_bounds,
                                              const cuopt_float_t* constraint_upper_bounds,
                                              const cuopt_float_t* variable_lower_bounds,
                                              const cuopt_float_t* variable_upper_bounds,
                                              cuOptOptimizationProblem* problem_ptr);

/* cuopt_c.h:414 */
void cuOptDestroyProblem(cuOptOptimizationProblem* problem_ptr);

/* ---- problem getters (cuopt_c.h:425-592) --------------------------------- */
cuopt_int_t cuOptGetNumConstraints(cuOptOptimizationProblem problem, cuopt_int_t* num_constraints_ptr);
cuopt_int_t cuOptGetNumVariables(cuOptOptimizationProblem problem, cuopt_int_t* num_variables_ptr);
cuopt_int_t cuOptGetObjectiveSense(cuOptOptimizationProblem problem, cuopt_int_t* objective_sense_ptr);
cuopt_int_t cuOptGetObjectiveOffset(cuOptOptimizationProblem problem, cuopt_float_t* objective_offset_ptr);
cuopt_int_t cuOptGetObjectiveCoefficients(cuOptOptimizationProblem problem,
                                          cuopt_float_t* objective_coefficients_ptr);
cuopt_int_t cuOptGetNumNonZeros(cuOptOptimizationProblem problem, cuopt_int_t* num_non_zeros_ptr);
cuopt_int_t cuOptGetConstraintMatrix(cuOptOptimizationProblem problem,
                                     cuopt_int_t* constraint_matrix_row_offsets_ptr,
                                     cuopt_int_t* constraint_matrix_column_indices_ptr,
                                     cuopt_float_t* constraint_matrix_coefficients_ptr);
cuopt_int_t cuOptGetConstraintSense(cuOptOptimizationProblem problem, char* constraint_sense_ptr);
cuopt_int_t cuOptGetConstraintRightHandSide(cuOptOptimizationProblem problem, cuopt_float_t* rhs_ptr);
cuopt_int_t cuOptGetConstraintLowerBounds(cuOptOptimizationProblem problem, cuopt_float_t* lower_bounds_ptr);
cuopt_int_t cuOptGetConstraintUpperBounds(cuOptOptimizationProblem problem, cuopt_float_t* upper_bounds_ptr);
cuopt_int_t cuOptGetVariableLowerBounds(cuOptOptimizationProblem problem, cuopt_float_t* lower_bounds_ptr);
cuopt_int_t cuOptGetVariableUpperBounds(cuOptOptimizationProblem problem, cuopt_float_t* upper_bounds_ptr);
cuopt_int_t cuOptGetVariableTypes(cuOptOptimizationProblem problem, char* variable_types_ptr);

/* ---- settings (cuopt_c.h:601-798) ------------------------------------------ */
cuopt_int_t cuOptCreateSolverSettings(cuOptSolverSettings* settings_ptr);
void cuOptDestroySolverSettings(cuOptSolverSettings* settings_ptr);
cuopt_int_t cuOptSetParameter(cuOptSolverSettings settings,
                              const char* parameter_name,
                              const char* parameter_value);
cuopt_int_t cuOptGetParameter(cuOptSolverSettings settings,
                              const char* parameter_name,
                              cuopt_int_t parameter_value_size,
                              char* parameter_value);
cuopt_int_t cuOptSetIntegerParameter(cuOptSolverSettings settings,
                                     const char* parameter_name,
                                     cuopt_int_t parameter_value);
cuopt_int_t cuOptGetIntegerParameter(cuOptSolverSettings settings,
                                     const char* parameter_name,
                                     cuopt_int_t* parameter_value);
cuopt_int_t cuOptSetFloatParameter(cuOptSolverSettings settings,
                                   const char* parameter_name,
                                   cuopt_float_t parameter_value);
cuopt_int_t cuOptGetFloatParameter(cuOptSolverSettings settings,
                                   const char* parameter_name,
                                   cuopt_float_t* parameter_value);
cuopt_int_t cuOptSetInitialPrimalSolution(cuOptSolverSettings settings,
                                          const cuopt_float_t* primal_solution,
                                          cuopt_int_t num_variables);
cuopt_int_t cuOptSetInitialDualSolution(cuOptSolverSettings settings,
                                        const cuopt_float_t* dual_solution,
                                        cuopt_int_t num_constraints);

/* MIP-only settings (cuopt_c.h:711-765 callbacks, :818 MIP start).  As in the
 * reference, they are accepted and kept with the settings (same argument
 * checks, CUOPT_SUCCESS) and an LP solve never uses them. */
typedef void (*cuOptMIPGetSolutionCallback)(const cuopt_float_t* solution,
                                            const cuopt_float_t* objective_value,
                                            const cuopt_float_t* solution_bound,
                                            void* user_data);
typedef void (*cuOptMIPSetSolutionCallback)(cuopt_float_t* solution,
                                            cuopt_float_t* objective_value,
                                            const cuopt_float_t* solution_bound,
                                            void* user_data);
cuopt_int_t cuOptSetMIPGetSolutionCallback(cuOptSolverSettings settings,
                                           cuOptMIPGetSolutionCallback callback,
                                           void* user_data);
cuopt_int_t cuOptSetMIPSetSolutionCallback(cuOptSolverSettings settings,
                                           cuOptMIPSetSolutionCallback callback,
                                           void* user_data);
cuopt_int_t cuOptAddMIPStart(cuOptSolverSettings settings,
                             const cuopt_float_t* solution,
                             cuopt_int_t num_variables);

/* ---- solve (cuopt_c.h:831-853) -------------------------------------------- */
cuopt_int_t cuOptIsMIP(cuOptOptimizationProblem problem, cuopt_int_t* is_mip_ptr);
/* Returns the solution's error status (cuopt_c.cpp:896-897). */
cuopt_int_t cuOptSolve(cuOptOptimizationProblem problem,
                       cuOptSolverSettings settings,
                       cuOptSolution* solution_ptr);
void cuOptDestroySolution(cuOptSolution* solution_ptr);

/* ---- solution (cuopt_c.h:864-978) ------------------------------------------ */
cuopt_int_t cuOptGetTerminationStatus(cuOptSolution solution, cuopt_int_t* termination_status_ptr);
cuopt_int_t cuOptGetErrorStatus(cuOptSolution solution, cuopt_int_t* error_status_ptr);
cuopt_int_t cuOptGetErrorString(cuOptSolution solution,
                                char* error_string_ptr,
                                cuopt_int_t error_string_size);
cuopt_int_t cuOptGetPrimalSolution(cuOptSolution solution, cuopt_float_t* solution_values);
cuopt_int_t cuOptGetObjectiveValue(cuOptSolution solution, cuopt_float_t* objective_value_ptr);
cuopt_int_t cuOptGetSolveTime(cuOptSolution solution, cuopt_float_t* solve_time_ptr);
cuopt_int_t cuOptGetDualSolution(cuOptSolution solution, cuopt_float_t* dual_solution_ptr);
/* MIP solution getters (cuopt_c.h:933, 944): an LP solution has neither, and
 * both return CUOPT_INVALID_ARGUMENT for it, as in the reference
 * (cuopt_c.cpp:1050-1082). */
cuopt_int_t cuOptGetMIPGap(cuOptSolution solution, cuopt_float_t* mip_gap_ptr);
cuopt_int_t cuOptGetSolutionBound(cuOptSolution solution, cuopt_float_t* solution_bound_ptr);
cuopt_int_t cuOptGetDualObjectiveValue(cuOptSolution solution, cuopt_float_t* dual_objective_value_ptr);
cuopt_int_t cuOptGetReducedCosts(cuOptSolution solution, cuopt_float_t* reduced_cost_ptr);

#ifdef __cplusplus
}
#endif

#endif /* CUOPT_AMD_CUOPT_C_H */
