/*
 * cuopt_amd_ext.h -- additions to the cuOpt C API exported by libcuopt_amd.so.
 *
 * None of these exists in the reference C API; each says what it mirrors.
 * A caller that only uses cuopt_c.h never needs this header.
 */
#ifndef CUOPT_AMD_EXT_H
#define CUOPT_AMD_EXT_H

#include <stddef.h>

#include "cuopt_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Integer parameter selecting the PDHG iterate arithmetic (no reference
 * counterpart; the reference LP path is FP64 only, constants.h:17-18):
 *   0 = FP64 iterates and matrices (default, the reference arithmetic);
 *   1 = FP32 iterates and scaled matrices with compensated FP32 row sums,
 *       FP64 scaling factors, FP64 termination / restart arithmetic. */
#define CUOPT_AMD_PDLP_PRECISION "pdlp_precision"
#define CUOPT_AMD_PRECISION_FP64       0
#define CUOPT_AMD_PRECISION_FP32_MIXED 1

/* Integer parameter (0/1, default 1) enabling FP64 iterative refinement of
 * FP32-mixed solves (no reference counterpart).  Once the relative KKT error is
 * below 1e-2, every PDLP restart recentres the FP32 iterates: the solve
 * continues on the original LP written in the coordinates (x - x*) * Dp,
 * (y - y*) * Dd around the restart point's FP64 values (x*, y*), with Dp, Dd the
 * inverse primal / dual residual norms.  FP32 iterates then carry only the
 * unresolved part of the solution, so FP32-mixed solves reach FP64-level
 * tolerances (e.g. 1e-8) and do not stall when a step size becomes small.
 * Optimality is always tested in FP64 on the original problem.  Ignored by
 * FP64 solves. */
#define CUOPT_AMD_PDLP_REFINEMENT "pdlp_refinement"

/* Integer parameter (0/1, default 0) enabling the SpMV autotuner (no reference
 * counterpart; the reference's only tuner times SpMM batch sizes,
 * optimal_batch_size_handler.cu:196-435).  At setup each of A and A^T is tried
 * with tile capacities of 512, 1 024 and 2 048 nonzeros and with the planner's
 * lanes per row G, G / 2 and 2 G (the row binning), and, when rows longer than a
 * tile exist, their merge inside the SpMV launch or in a launch of its own; each
 * candidate's SpMV is run once untimed and then ten times between HIP events on
 * the real matrix, and the fastest is used unless the untuned plan is within 3 %
 * of it (then the untuned plan stays, so that timing noise does not change the
 * choice from run to run).  The choice is cached per matrix (sizes, precision and a
 * hash of the row offsets) for the life of the process.  Candidates sum a row in
 * different orders, so with the tuner on, results can differ from run to run in
 * the last bits; off (the default), every solve uses capacity 1 024 and the
 * planner's G and is bitwise reproducible. */
#define CUOPT_AMD_SPMV_AUTOTUNE "pdlp_spmv_autotune"

/* Integer parameter (-1, 0 or 1, default -1) choosing the solver's internal column
 * order (no reference counterpart).  1 numbers the columns in the order in which
 * rows of at most one tile (1 020 nonzeros) first touch them, rows in index
 * order, so that the rows one workgroup reduces together gather x-bar from nearby
 * addresses; 0 keeps the given order; -1 (the default) uses 1 for matrices of at
 * least 2^20 nonzeros.  The LP is the same LP: inputs and outputs stay in the
 * caller's column order; only the order of the sums in a row changes. */
#define CUOPT_AMD_COLUMN_ORDER "pdlp_column_order"

/* Integer parameter (0/1, default 0; no reference counterpart).  1 ends a
 * single-LP solve with CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR at the first
 * termination check whose residuals or gap are NaN.  0 keeps the reference's
 * Stable3 behaviour: such a solve runs on to its iteration or time limit.
 * cuOptAmdBatchSolve always stops an LP at a NaN check, since its persistent
 * launch must end. */
#define CUOPT_AMD_STOP_ON_NAN "pdlp_stop_on_nan"

/* Integer parameter (-1/0/1, default -1; no reference counterpart).  Rows longer
 * than an SpMV tile are summed in chunks.  0 merges the chunks in a launch of its
 * own after the SpMV; 1 lets the last chunk of each row merge them inside the SpMV
 * launch (device-scope atomics on a per-row counter); -1 picks 1 when no row has
 * more than 64 chunks, and with pdlp_spmv_autotune times both.  The results are
 * bitwise the same either way. */
#define CUOPT_AMD_LONG_ROW_MERGE "pdlp_long_row_merge"

/* Integer parameter (0/1, default 0) changing the per_constraint_residual test
 * (no reference counterpart).  0 keeps the reference's signed form
 * (relative_residual_t, utils.cuh:576-590): max(0, max_i r_i - rel * w_i) with the
 * signed residual r_i and w_i the row's larger finite bound magnitude, and
 * max(0, max_j r_j - rel * c_j) with the signed dual residual and signed c_j, so a
 * residual below the bound side passes.  1 takes |r_i| - rel * w_i and
 * |r_j| - rel * |c_j|, which bounds every residual in magnitude. */
#define CUOPT_AMD_PER_CONSTRAINT_ABSOLUTE "pdlp_per_constraint_absolute"

/* Slots of cuOptAmdGetTerminationInfo.  Slots 0-9 mirror
 * additional_termination_information_t (pdlp/solver_solution.hpp:53-93). */
#define CUOPT_AMD_INFO_STEPS               0  /* number_of_steps_taken */
#define CUOPT_AMD_INFO_L2_PRIMAL           1  /* l2_primal_residual */
#define CUOPT_AMD_INFO_REL_PRIMAL          2  /* l2_relative_primal_residual */
#define CUOPT_AMD_INFO_L2_DUAL             3  /* l2_dual_residual */
#define CUOPT_AMD_INFO_REL_DUAL            4  /* l2_relative_dual_residual */
#define CUOPT_AMD_INFO_PRIMAL_OBJ          5
#define CUOPT_AMD_INFO_DUAL_OBJ            6
#define CUOPT_AMD_INFO_GAP                 7
#define CUOPT_AMD_INFO_REL_GAP             8
#define CUOPT_AMD_INFO_SOLVE_TIME          9  /* seconds, from solve entry */
#define CUOPT_AMD_INFO_SETUP_TIME          10 /* seconds spent before the first PDHG step */
#define CUOPT_AMD_INFO_LOOP_TIME           11 /* seconds in the PDHG loop (incl. checks); batch solves: the
                                                 persistent launch's device time (HIP events) */
#define CUOPT_AMD_INFO_ETA0                12 /* initial step size */
#define CUOPT_AMD_INFO_OMEGA0              13 /* initial primal weight */
#define CUOPT_AMD_INFO_OMEGA               14 /* final primal weight */
#define CUOPT_AMD_INFO_RESTARTS            15
#define CUOPT_AMD_INFO_POWER_ITERS         16
#define CUOPT_AMD_INFO_SIGMA_MAX_SQ        17
#define CUOPT_AMD_INFO_BETA                18 /* bound rescaling */
#define CUOPT_AMD_INFO_GAMMA               19 /* objective rescaling */
#define CUOPT_AMD_INFO_PRECISION           20
#define CUOPT_AMD_INFO_CHECKS              21 /* termination evaluations */
#define CUOPT_AMD_INFO_STRUCTURE_TIME      22 /* setup: CSR sort, transpose, tile plans, uploads (s) */
#define CUOPT_AMD_INFO_SCALING_TIME        23 /* setup: Ruiz, Pock-Chambolle, bound/objective rescaling (s) */
#define CUOPT_AMD_INFO_POWER_TIME          24 /* setup: power iteration for the step size (s) */
#define CUOPT_AMD_INFO_REFINE_ROUNDS       25 /* FP64 recentrings of a FP32-mixed solve */
#define CUOPT_AMD_INFO_TILE_A              26 /* SpMV tile capacity over A (nonzeros) */
#define CUOPT_AMD_INFO_LANES_BIAS_A        27 /* lanes-per-row shift over A (-1, 0, +1: G / 2, G, 2 G) */
#define CUOPT_AMD_INFO_TILE_AT             28 /* the same over A^T */
#define CUOPT_AMD_INFO_LANES_BIAS_AT       29
#define CUOPT_AMD_INFO_TUNE_TIME           30 /* setup: SpMV autotuning (s; 0 when off or cached) */
#define CUOPT_AMD_INFO_MERGE_A             31 /* rows of A longer than a tile: 1 merged by their last chunk
                                                 in the SpMV launch, 0 by a launch of their own, -1 none */
#define CUOPT_AMD_INFO_MERGE_AT            32 /* the same over A^T */
#define CUOPT_AMD_INFO_LP_TIME             33 /* batch solves: this LP's device time, from its workgroup's
                                                 start to its end (s); single solves: 0 */
#define CUOPT_AMD_INFO_LP_START            34 /* batch solves: when this LP's workgroup started, after the
                                                 first one of the launch (s); single solves: 0 */
#define CUOPT_AMD_INFO_COUNT               35

/* Copies up to n slots of the last solve's statistics into info. */
cuopt_int_t cuOptAmdGetTerminationInfo(cuOptSolution solution, cuopt_float_t* info, cuopt_int_t n);

/* Per-check trace of the solve: rows of CUOPT_AMD_TRACE_WIDTH doubles
 * (k, l2 primal residual, l2 dual residual, primal obj, dual obj, gap,
 * restarted, primal weight after the check, fixed-point error, iterations since
 * restart).  Writes min(cap, recorded) rows and the recorded count to *len. */
#define CUOPT_AMD_TRACE_WIDTH 10
cuopt_int_t cuOptAmdGetTrace(cuOptSolution solution, cuopt_float_t* rows, cuopt_int_t cap, cuopt_int_t* len);

/* Reads num_solutions solutions in one call (no reference counterpart: for
 * callers that read thousands of small solutions, such as a batch, where one
 * call per getter costs more than the copies).  For solution i, with n variables
 * and m constraints, x (n), y (m), the reduced costs (n) and the termination
 * info (CUOPT_AMD_INFO_COUNT values, as cuOptAmdGetTerminationInfo) are written
 * one after another from store + offsets[i], and must fill exactly
 * [offsets[i], offsets[i + 1]) (offsets has num_solutions + 1 entries); arrays a
 * failed solve did not produce are written as zeros.  status[i] and
 * error_status[i] receive the termination and error statuses, and
 * objectives[3 i .. 3 i + 2] the primal objective, dual objective and solve time.
 * Every size is checked before anything is written; the copies run on up to 16
 * host threads.  The handles stay valid (destroy each with cuOptDestroySolution). */
cuopt_int_t cuOptAmdGetSolutions(cuopt_int_t num_solutions,
                                 const cuOptSolution* solutions,
                                 cuopt_float_t* store,
                                 const int64_t* offsets,
                                 cuopt_int_t* status,
                                 cuopt_int_t* error_status,
                                 cuopt_float_t* objectives);

/* y = A x on the GPU with the solver's tiled row-block SpMV (the kernel that
 * replaces the cusparseSpMV calls of pdhg.cu:285-347).  Host arrays in and
 * out.  precision: CUOPT_AMD_PRECISION_FP64 (FP64 products and sums) or
 * CUOPT_AMD_PRECISION_FP32_MIXED (values and x rounded to FP32, compensated
 * FP32 row sums).  Returns CUOPT_SUCCESS or an error code. */
cuopt_int_t cuOptAmdSpmv(cuopt_int_t num_rows,
                         cuopt_int_t num_cols,
                         const cuopt_int_t* row_offsets,
                         const cuopt_int_t* column_indices,
                         const cuopt_float_t* values,
                         const cuopt_float_t* x,
                         cuopt_float_t* y,
                         cuopt_int_t precision);

/* cuOptAmdSpmv without device allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (tight loops, code that must not call
 * hipMalloc).  Same arguments and result as cuOptAmdSpmv, bitwise; every device
 * buffer of the call (the staged CSR arrays, x, y, the tile plan and the long-row
 * partials) lives in one fixed device array per GPU of cuOptAmdSpmvScratchBytes
 * (64 MiB), held until the process ends.  A problem whose staged data exceed it
 * is refused with CUOPT_INVALID_ARGUMENT: (m + 1) * 4 + (nnz + 16) * (4 + s) +
 * n * s + m * 8 bytes plus at most 44 bytes per 1 020 nonzeros, s = 8 (FP64) or
 * 4 (FP32-mixed), each part rounded up to 256 bytes.  Runs on the calling
 * thread's default stream and returns after y is written.  Calls that overlap on
 * one GPU share the array and race. */
cuopt_int_t cuOptAmdSpmvNoAlloc(cuopt_int_t num_rows,
                                cuopt_int_t num_cols,
                                const cuopt_int_t* row_offsets,
                                const cuopt_int_t* column_indices,
                                const cuopt_float_t* values,
                                const cuopt_float_t* x,
                                cuopt_float_t* y,
                                cuopt_int_t precision);
/* Size in bytes of the fixed device array of cuOptAmdSpmvNoAlloc. */
cuopt_int_t cuOptAmdSpmvScratchBytes(size_t* bytes);

/* Benchmark of the PDHG loop (no reference counterpart).  Runs the full solve
 * setup and `warmup` PDHG iterations, then times `steps` iterations between two
 * device synchronisations on the production path (graph-captured iteration
 * blocks, termination checks at the reference cadence included), then runs up
 * to 400 more iterations with direct launches bracketed by HIP events on the
 * solver's stream to time the primal-step and dual-step kernels, and stops.
 * Window edges are moved up to the next termination-check iteration.
 * out (n slots): [0] timed window in ms, [1] mean primal-step kernel ms,
 * [2] mean dual-step kernel ms (long-row launch included), [3] iterations in
 * the timed window, [4] setup seconds, [5] termination checks inside the window,
 * [6] termination status, [7] precision used, [8] iterations with events. */
#define CUOPT_AMD_BENCH_COUNT 9
cuopt_int_t cuOptAmdBenchmarkSteps(cuOptOptimizationProblem problem,
                                   cuOptSolverSettings settings,
                                   cuopt_int_t warmup,
                                   cuopt_int_t steps,
                                   cuopt_float_t* out,
                                   cuopt_int_t n);

/* Names of a problem read from an MPS file (the reference's data model keeps
 * them: data_model.py get_variable_names / get_row_names / get_objective_name /
 * get_problem_name, mps_parser.cpp).  The name is copied NUL-terminated into
 * name_buffer, truncated to buffer_size - 1 characters; a problem built from
 * arrays has empty names.  CUOPT_INVALID_ARGUMENT for a null pointer, an index
 * out of range or buffer_size < 1. */
cuopt_int_t cuOptAmdGetVariableName(cuOptOptimizationProblem problem,
                                    cuopt_int_t index,
                                    char* name_buffer,
                                    cuopt_int_t buffer_size);
cuopt_int_t cuOptAmdGetConstraintName(cuOptOptimizationProblem problem,
                                      cuopt_int_t index,
                                      char* name_buffer,
                                      cuopt_int_t buffer_size);
cuopt_int_t cuOptAmdGetObjectiveName(cuOptOptimizationProblem problem, char* name_buffer, cuopt_int_t buffer_size);
cuopt_int_t cuOptAmdGetProblemName(cuOptOptimizationProblem problem, char* name_buffer, cuopt_int_t buffer_size);

/* Sets the names of a problem built from arrays, as the reference's data model
 * does (data_model.py set_variable_names / set_row_names / set_objective_name /
 * set_problem_name).  They name the variables of the solution_file output and
 * the rows and columns of user_problem_file and cuOptWriteProblem.  count must
 * be the number of variables (constraints) and every entry non-null; the
 * strings are copied.  CUOPT_INVALID_ARGUMENT otherwise. */
cuopt_int_t cuOptAmdSetVariableNames(cuOptOptimizationProblem problem, const char* const* names, cuopt_int_t count);
cuopt_int_t cuOptAmdSetConstraintNames(cuOptOptimizationProblem problem, const char* const* names, cuopt_int_t count);
cuopt_int_t cuOptAmdSetObjectiveName(cuOptOptimizationProblem problem, const char* name);
cuopt_int_t cuOptAmdSetProblemName(cuOptOptimizationProblem problem, const char* name);

/* Number of GPUs visible to the library (0 when no HIP device is usable). */
cuopt_int_t cuOptAmdDeviceCount(void);

/* ---- differentiable unrolled PDHG ------------------------------------------------
 * K steps of PDHG (no restarts) for B objectives at once on one constraint
 * matrix, min c_b^T x + ridge/2 ||x||^2 s.t. lc <= A x <= uc, lb <= x <= ub,
 * with step size 0.998 / sigma_max(A), and the vector-Jacobian product of
 * (x_K, y_K) with respect to (c, lc, uc) through checkpointed recomputation.
 * No reference counterpart (the reference LP path is not differentiable);
 * it serves the torch.autograd.Function in cuopt_amd/autograd.py. */
typedef void* cuOptAmdUnrolled;
/* Host CSR, host bounds; precision as CUOPT_AMD_PDLP_PRECISION. */
cuopt_int_t cuOptAmdUnrolledCreate(cuopt_int_t num_constraints,
                                   cuopt_int_t num_variables,
                                   const cuopt_int_t* row_offsets,
                                   const cuopt_int_t* column_indices,
                                   const cuopt_float_t* values,
                                   const cuopt_float_t* lower_bounds,
                                   const cuopt_float_t* upper_bounds,
                                   cuopt_float_t ridge,
                                   cuopt_int_t precision,
                                   cuOptAmdUnrolled* handle_ptr);
void cuOptAmdUnrolledDestroy(cuOptAmdUnrolled* handle_ptr);
cuopt_int_t cuOptAmdUnrolledStepSize(cuOptAmdUnrolled handle, cuopt_float_t* eta);
/* Device pointers, batch-interleaved (v[i * batch + b]).  c: n x batch;
 * lc, uc: m x batch if bounds_per_batch else m; x_out: n x batch; y_out:
 * m x batch or NULL.  Checkpoints every `checkpoint_every` steps.  Work is
 * queued on `stream` (a hipStream_t, NULL for the default stream). */
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
                                    void* stream);
/* VJP of the last forward call: gx (n x batch), gy (m x batch or NULL) ->
 * gc (n x batch), glc and guc (m x batch), accumulated in FP64 (FP32 handles sum
 * each checkpoint segment in FP32 and add it into the FP64 totals). */
cuopt_int_t cuOptAmdUnrolledBackward(cuOptAmdUnrolled handle,
                                     const cuopt_float_t* gx,
                                     const cuopt_float_t* gy,
                                     cuopt_float_t* gc,
                                     cuopt_float_t* glc,
                                     cuopt_float_t* guc,
                                     void* stream);

/* Batch size for a same-matrix batch on this handle's matrix: times
 * Y = A X plus X = A^T Y (batch-interleaved SpMM, one untimed run and the mean
 * of five) at the sizes visited by the search of the reference's batch-size
 * handler (optimal_batch_size_handler.cu:196-435, called from solve.cu:788):
 * start at the largest power of two <= min(128, max_batch_size), move by
 * factors of two in the direction whose time per batch member falls, at most
 * four steps, then try the size halfway back.  *batch_size gets a value in
 * [1, max_batch_size]; callers split larger batches into chunks of that size. */
cuopt_int_t cuOptAmdUnrolledOptimalBatchSize(cuOptAmdUnrolled handle,
                                             cuopt_int_t max_batch_size,
                                             cuopt_int_t* batch_size);
/* The same search over a caller-supplied cost per batch member, cost(b, context)
 * (for callers that time their own SpMM, and for tests of the search). */
cuopt_int_t cuOptAmdBatchSizeSearch(cuopt_int_t max_batch_size,
                                    cuopt_float_t (*cost)(cuopt_int_t batch_size, void* context),
                                    void* context,
                                    cuopt_int_t* batch_size);

/* Solves num_problems independent LPs together (the heterogeneous batch of the
 * reference's Python BatchSolve, cython_solve.cu:286-352) in one persistent
 * launch: one workgroup per LP runs setup and the whole Stable3 loop on the
 * device.  Uses the settings' tolerances, limits and pdlp_precision; initial
 * solutions and refinement are not used.  solutions[i] receives a solution
 * handle for problems[i] (destroy each with cuOptDestroySolution); every
 * solution reports the wall time of the whole batch as its solve time.
 * Returns the first non-success error status, or CUOPT_SUCCESS. */
cuopt_int_t cuOptAmdBatchSolve(cuopt_int_t num_problems,
                               const cuOptOptimizationProblem* problems,
                               cuOptSolverSettings settings,
                               cuOptSolution* solutions);

#ifdef __cplusplus
}
#endif

#endif /* CUOPT_AMD_EXT_H */
