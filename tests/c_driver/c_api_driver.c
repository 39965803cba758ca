/*
 * c_api_driver.c -- a plain C99 caller of the cuOpt C API, compiled against
 * include/cuopt_c.h and linked to libcuopt_amd.so, as a reference C program
 * would be after switching its include and link lines.
 *
 * Each case restates one of the reference's C API checks
 * (cpp/tests/linear_programming/c_api_tests/c_api_test.c with the expectations of
 * c_api_tests.cpp) and exits 0 when it holds:
 *
 *   sizes                int / float sizes 4 and 8 (c_api_tests.cpp:21-23)
 *   bad_parameter        every setter and getter refuses an unknown name with
 *                        CUOPT_INVALID_ARGUMENT (c_api_test.c:81-135)
 *   missing_file         cuOptReadProblem of a missing file gives
 *                        CUOPT_MPS_FILE_ERROR (c_api_test.c:70-79)
 *   reference_parameters the MIP, barrier and work-limit names and the
 *                        CUOPT_MODE_* values the reference's tests use are
 *                        accepted and read back (c_api_test.c:1616-1660)
 *   write_read IN OUT    read IN, write OUT, read OUT back: same sizes, matrix,
 *                        objective and bounds (c_api_test.c:1380-1459)
 *   infeasible           the 9-row infeasible LP reports Infeasible
 *                        (c_api_test.c:791-925; settled by presolve, no GPU)
 *   ranged               max 5x + 8y over a ranged row: Optimal, objective 32
 *                        within 1e-3 (c_api_test.c:927-1041, c_api_tests.cpp:105-112)
 *   maximize_duals       the maximisation LP of c_api_test.c:1461-1614 under
 *                        Concurrent and PDLP: Optimal, primal and dual objectives
 *                        within 1e-2; Dual simplex is refused
 *   afiro FILE           afiro: Optimal with a 60 s limit, IterationLimit with a
 *                        limit of 1 (c_api_tests.cpp:25-36, 71-78)
 *
 * The last three solve on the GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cuopt_c.h"

#define EXPECT(cond)                                                        \
  do {                                                                      \
    if (!(cond)) {                                                          \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      return 1;                                                             \
    }                                                                       \
  } while (0)

static int case_sizes(void)
{
  EXPECT(cuOptGetIntSize() == 4);
  EXPECT(cuOptGetFloatSize() == 8);
  EXPECT(sizeof(cuopt_int_t) == 4 && sizeof(cuopt_float_t) == 8);
  return 0;
}

static int case_bad_parameter(void)
{
  cuOptSolverSettings s = NULL;
  char buf[64];
  cuopt_int_t iv;
  cuopt_float_t fv;
  EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
  EXPECT(cuOptSetParameter(s, "bad_parameter_name", "1") == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptGetParameter(s, "bad_parameter_name", (cuopt_int_t)sizeof buf, buf) == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptSetIntegerParameter(s, "bad_parameter_name", 1) == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptSetFloatParameter(s, "bad_parameter_name", 1.0) == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptGetIntegerParameter(s, "bad_parameter_name", &iv) == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptGetFloatParameter(s, "bad_parameter_name", &fv) == CUOPT_INVALID_ARGUMENT);
  /* a known name with an out-of-range value is refused too */
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_METHOD, 7) == CUOPT_INVALID_ARGUMENT);
  cuOptDestroySolverSettings(&s);
  EXPECT(s == NULL);
  return 0;
}

static int case_missing_file(void)
{
  cuOptOptimizationProblem p = NULL;
  EXPECT(cuOptReadProblem("missing_file.mps", &p) == CUOPT_MPS_FILE_ERROR);
  cuOptDestroyProblem(&p);
  return 0;
}

static int case_reference_parameters(void)
{
  cuOptSolverSettings s = NULL;
  cuopt_int_t iv     = -5;
  cuopt_float_t fv   = -5.0;
  char buf[64];
  EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
  EXPECT(cuOptGetIntegerParameter(s, CUOPT_MIP_DETERMINISM_MODE, &iv) == CUOPT_SUCCESS);
  EXPECT(iv == CUOPT_MODE_OPPORTUNISTIC);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_MIP_DETERMINISM_MODE, CUOPT_MODE_DETERMINISTIC) == CUOPT_SUCCESS);
  EXPECT(cuOptGetIntegerParameter(s, CUOPT_MIP_DETERMINISM_MODE, &iv) == CUOPT_SUCCESS);
  EXPECT(iv == CUOPT_MODE_DETERMINISTIC);
  EXPECT(cuOptSetFloatParameter(s, CUOPT_WORK_LIMIT, 3.5) == CUOPT_SUCCESS);
  EXPECT(cuOptGetFloatParameter(s, CUOPT_WORK_LIMIT, &fv) == CUOPT_SUCCESS && fv == 3.5);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_FOLDING, 1) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_AUGMENTED, 0) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_DUALIZE, -1) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_ORDERING, 1) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_BARRIER_DUAL_INITIAL_POINT, 0) == CUOPT_SUCCESS);
  EXPECT(cuOptSetParameter(s, CUOPT_ELIMINATE_DENSE_COLUMNS, "false") == CUOPT_SUCCESS);
  EXPECT(cuOptSetParameter(s, CUOPT_CUDSS_DETERMINISTIC, "true") == CUOPT_SUCCESS);
  EXPECT(cuOptGetParameter(s, CUOPT_CUDSS_DETERMINISTIC, (cuopt_int_t)sizeof buf, buf) == CUOPT_SUCCESS);
  EXPECT(strcmp(buf, "true") == 0);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_MIP_CUT_PASSES, 3) == CUOPT_SUCCESS);
  EXPECT(cuOptSetFloatParameter(s, CUOPT_MIP_RELATIVE_GAP, 1e-3) == CUOPT_SUCCESS);
  EXPECT(cuOptSetFloatParameter(s, CUOPT_MIP_ABSOLUTE_TOLERANCE, 1e-5) == CUOPT_SUCCESS);
  EXPECT(cuOptSetFloatParameter(s, CUOPT_MIP_CUT_MIN_ORTHOGONALITY, 0.25) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_MIP_BATCH_PDLP_STRONG_BRANCHING, 1) == CUOPT_SUCCESS);
  EXPECT(cuOptSetParameter(s, CUOPT_MIP_HEURISTICS_ONLY, "true") == CUOPT_SUCCESS);
  /* ranges of the reference registry: cut_min_orthogonality <= 1, determinism mode <= 1 */
  EXPECT(cuOptSetFloatParameter(s, CUOPT_MIP_CUT_MIN_ORTHOGONALITY, 1.5) == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_MIP_DETERMINISM_MODE, 2) == CUOPT_INVALID_ARGUMENT);
  /* mip_presolve is a name constant without a registry entry in the reference */
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_MIP_PRESOLVE, 1) == CUOPT_INVALID_ARGUMENT);
  EXPECT(CUOPT_INSTANTIATE_DOUBLE == 1 && CUOPT_INSTANTIATE_INT32 == 1);
  cuOptDestroySolverSettings(&s);
  return 0;
}

static int read_all(cuOptOptimizationProblem p, cuopt_int_t* m, cuopt_int_t* n, cuopt_int_t* nnz, double** c,
                    cuopt_int_t** ptr, cuopt_int_t** idx, double** val, double** lc, double** uc, double** lb,
                    double** ub)
{
  EXPECT(cuOptGetNumConstraints(p, m) == CUOPT_SUCCESS);
  EXPECT(cuOptGetNumVariables(p, n) == CUOPT_SUCCESS);
  EXPECT(cuOptGetNumNonZeros(p, nnz) == CUOPT_SUCCESS);
  *c   = malloc(sizeof(double) * (size_t)(*n + 1));
  *ptr = malloc(sizeof(cuopt_int_t) * (size_t)(*m + 1));
  *idx = malloc(sizeof(cuopt_int_t) * (size_t)(*nnz + 1));
  *val = malloc(sizeof(double) * (size_t)(*nnz + 1));
  *lc  = malloc(sizeof(double) * (size_t)(*m + 1));
  *uc  = malloc(sizeof(double) * (size_t)(*m + 1));
  *lb  = malloc(sizeof(double) * (size_t)(*n + 1));
  *ub  = malloc(sizeof(double) * (size_t)(*n + 1));
  EXPECT(cuOptGetObjectiveCoefficients(p, *c) == CUOPT_SUCCESS);
  EXPECT(cuOptGetConstraintMatrix(p, *ptr, *idx, *val) == CUOPT_SUCCESS);
  EXPECT(cuOptGetConstraintLowerBounds(p, *lc) == CUOPT_SUCCESS);
  EXPECT(cuOptGetConstraintUpperBounds(p, *uc) == CUOPT_SUCCESS);
  EXPECT(cuOptGetVariableLowerBounds(p, *lb) == CUOPT_SUCCESS);
  EXPECT(cuOptGetVariableUpperBounds(p, *ub) == CUOPT_SUCCESS);
  return 0;
}

static int same_doubles(const double* a, const double* b, cuopt_int_t k)
{
  for (cuopt_int_t i = 0; i < k; ++i)
    if (!(a[i] == b[i] || (isinf(a[i]) && isinf(b[i]) && (a[i] > 0) == (b[i] > 0)))) return 0;
  return 1;
}

static int case_write_read(const char* in, const char* out)
{
  cuOptOptimizationProblem a = NULL, b = NULL;
  cuopt_int_t m1, n1, z1, m2, n2, z2;
  double *c1, *v1, *lc1, *uc1, *lb1, *ub1, *c2, *v2, *lc2, *uc2, *lb2, *ub2;
  cuopt_int_t *p1, *i1, *p2, *i2;
  EXPECT(cuOptReadProblem(in, &a) == CUOPT_SUCCESS);
  EXPECT(cuOptWriteProblem(a, out, CUOPT_FILE_FORMAT_MPS) == CUOPT_SUCCESS);
  EXPECT(cuOptWriteProblem(a, out, 7) == CUOPT_INVALID_ARGUMENT);
  EXPECT(cuOptReadProblem(out, &b) == CUOPT_SUCCESS);
  EXPECT(read_all(a, &m1, &n1, &z1, &c1, &p1, &i1, &v1, &lc1, &uc1, &lb1, &ub1) == 0);
  EXPECT(read_all(b, &m2, &n2, &z2, &c2, &p2, &i2, &v2, &lc2, &uc2, &lb2, &ub2) == 0);
  EXPECT(m1 == m2 && n1 == n2 && z1 == z2);
  EXPECT(memcmp(p1, p2, sizeof(cuopt_int_t) * (size_t)(m1 + 1)) == 0);
  EXPECT(memcmp(i1, i2, sizeof(cuopt_int_t) * (size_t)z1) == 0);
  EXPECT(same_doubles(v1, v2, z1) && same_doubles(c1, c2, n1));
  EXPECT(same_doubles(lc1, lc2, m1) && same_doubles(uc1, uc2, m1));
  EXPECT(same_doubles(lb1, lb2, n1) && same_doubles(ub1, ub2, n1));
  free(c1); free(p1); free(i1); free(v1); free(lc1); free(uc1); free(lb1); free(ub1);
  free(c2); free(p2); free(i2); free(v2); free(lc2); free(uc2); free(lb2); free(ub2);
  cuOptDestroyProblem(&a);
  cuOptDestroyProblem(&b);
  EXPECT(a == NULL && b == NULL);
  return 0;
}

static int solve_status(cuOptOptimizationProblem p, cuOptSolverSettings s, cuopt_int_t* term, double* obj,
                        double* dobj)
{
  cuOptSolution sol = NULL;
  cuopt_int_t err   = -1;
  double t          = -1.0;
  EXPECT(cuOptSolve(p, s, &sol) == CUOPT_SUCCESS);
  EXPECT(cuOptGetErrorStatus(sol, &err) == CUOPT_SUCCESS && err == CUOPT_SUCCESS);
  EXPECT(cuOptGetTerminationStatus(sol, term) == CUOPT_SUCCESS);
  EXPECT(cuOptGetObjectiveValue(sol, obj) == CUOPT_SUCCESS);
  if (dobj) EXPECT(cuOptGetDualObjectiveValue(sol, dobj) == CUOPT_SUCCESS);
  EXPECT(cuOptGetSolveTime(sol, &t) == CUOPT_SUCCESS && t >= 0.0);
  cuOptDestroySolution(&sol);
  EXPECT(sol == NULL);
  return 0;
}

static int case_infeasible(void)
{
  /* rows: -0.5x1 + x2 >= 0.5, 2x1 - x2 >= 3, 3x1 + x2 <= 6, x5 <= 2, 3x4 - x5 <= 2,
   * x4 >= 5, x1 + x5 <= 10, x1 + 2x2 + x4 <= 14, x2 + x4 >= 1; variables (x1, x2, x4, x5) >= 0 */
  const cuopt_int_t ptr[] = {0, 2, 4, 6, 7, 9, 10, 12, 15, 17};
  const cuopt_int_t idx[] = {0, 1, 0, 1, 0, 1, 3, 2, 3, 2, 0, 3, 0, 1, 2, 1, 2};
  const double val[]      = {-0.5, 1, 2, -1, 3, 1, 1, 3, -1, 1, 1, 1, 1, 2, 1, 1, 1};
  const double rhs[]      = {0.5, 3, 6, 2, 2, 5, 10, 14, 1};
  const char sense[]      = {'G', 'G', 'L', 'L', 'L', 'G', 'L', 'L', 'G'};
  const double c[4] = {0, 0, 0, 0}, lb[4] = {0, 0, 0, 0};
  const double ub[4] = {CUOPT_INFINITY, CUOPT_INFINITY, CUOPT_INFINITY, CUOPT_INFINITY};
  const char types[4] = {CUOPT_CONTINUOUS, CUOPT_CONTINUOUS, CUOPT_CONTINUOUS, CUOPT_CONTINUOUS};
  cuOptOptimizationProblem p = NULL;
  cuOptSolverSettings s      = NULL;
  cuopt_int_t term = -1, mip = -1;
  double obj;
  EXPECT(cuOptCreateProblem(9, 4, CUOPT_MINIMIZE, 0.0, c, ptr, idx, val, sense, rhs, lb, ub, types, &p) ==
         CUOPT_SUCCESS);
  EXPECT(cuOptIsMIP(p, &mip) == CUOPT_SUCCESS && mip == 0);
  EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
  EXPECT(solve_status(p, s, &term, &obj, NULL) == 0);
  EXPECT(term == CUOPT_TERIMINATION_STATUS_INFEASIBLE);
  cuOptDestroySolverSettings(&s);
  cuOptDestroyProblem(&p);
  return 0;
}

static int case_ranged(void)
{
  /* max 5x + 8y  s.t.  2x + 3y <= 12,  3x + y <= 6,  2 <= x + 2y <= 8,  0 <= x, y <= 10 */
  const cuopt_int_t ptr[] = {0, 2, 4, 6};
  const cuopt_int_t idx[] = {0, 1, 0, 1, 0, 1};
  const double val[]      = {2, 3, 3, 1, 1, 2};
  const double lc[]       = {-CUOPT_INFINITY, -CUOPT_INFINITY, 2};
  const double uc[]       = {12, 6, 8};
  const double c[]        = {5, 8};
  const double lb[] = {0, 0}, ub[] = {10, 10};
  const char types[] = {CUOPT_CONTINUOUS, CUOPT_CONTINUOUS};
  cuOptOptimizationProblem p = NULL;
  cuOptSolverSettings s      = NULL;
  cuopt_int_t term = -1;
  double obj       = 0.0;
  EXPECT(cuOptCreateRangedProblem(3, 2, CUOPT_MAXIMIZE, 0.0, c, ptr, idx, val, lc, uc, lb, ub, types, &p) ==
         CUOPT_SUCCESS);
  EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
  EXPECT(cuOptSetFloatParameter(s, CUOPT_TIME_LIMIT, 60.0) == CUOPT_SUCCESS);
  EXPECT(solve_status(p, s, &term, &obj, NULL) == 0);
  printf("ranged: status %d objective %.9f\n", (int)term, obj);
  EXPECT(term == CUOPT_TERIMINATION_STATUS_OPTIMAL);
  EXPECT(fabs(obj - 32.0) <= 1e-3);
  cuOptDestroySolverSettings(&s);
  cuOptDestroyProblem(&p);
  return 0;
}

static int case_maximize_duals(void)
{
  /* max 4x1 + x2 + 5x3 + 3x4  s.t.  x1 - x2 - x3 + 3x4 <= 1,  5x1 + x2 + 3x3 + 8x4 <= 55,
   * -x1 + 2x2 + 3x3 - 5x4 <= 3,  x >= 0 */
  const cuopt_int_t ptr[] = {0, 4, 8, 12};
  const cuopt_int_t idx[] = {0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3};
  const double val[]      = {1, -1, -1, 3, 5, 1, 3, 8, -1, 2, 3, -5};
  const double rhs[]      = {1, 55, 3};
  const char sense[]      = {'L', 'L', 'L'};
  const double c[]        = {4, 1, 5, 3};
  const double lb[]       = {0, 0, 0, 0};
  const double ub[] = {CUOPT_INFINITY, CUOPT_INFINITY, CUOPT_INFINITY, CUOPT_INFINITY};
  const char types[] = {CUOPT_CONTINUOUS, CUOPT_CONTINUOUS, CUOPT_CONTINUOUS, CUOPT_CONTINUOUS};
  const cuopt_int_t methods[] = {CUOPT_METHOD_CONCURRENT, CUOPT_METHOD_PDLP};
  cuOptOptimizationProblem p = NULL;
  EXPECT(cuOptCreateProblem(3, 4, CUOPT_MAXIMIZE, 0.0, c, ptr, idx, val, sense, rhs, lb, ub, types, &p) ==
         CUOPT_SUCCESS);
  for (int k = 0; k < 2; ++k) {
    cuOptSolverSettings s = NULL;
    cuopt_int_t term      = -1;
    double obj = 0.0, dobj = 0.0;
    EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
    EXPECT(cuOptSetIntegerParameter(s, CUOPT_METHOD, methods[k]) == CUOPT_SUCCESS);
    EXPECT(cuOptSetFloatParameter(s, CUOPT_TIME_LIMIT, 60.0) == CUOPT_SUCCESS);
    EXPECT(solve_status(p, s, &term, &obj, &dobj) == 0);
    printf("maximize_duals method %d: status %d objective %.9f dual objective %.9f\n", (int)methods[k], (int)term,
           obj, dobj);
    EXPECT(term == CUOPT_TERIMINATION_STATUS_OPTIMAL);
    EXPECT(fabs(obj - dobj) <= 1e-2);
    cuOptDestroySolverSettings(&s);
  }
  {
    /* this engine runs no dual simplex: the method is refused with the reason */
    cuOptSolverSettings s = NULL;
    cuOptSolution sol     = NULL;
    char why[256];
    EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
    EXPECT(cuOptSetIntegerParameter(s, CUOPT_METHOD, CUOPT_METHOD_DUAL_SIMPLEX) == CUOPT_SUCCESS);
    EXPECT(cuOptSolve(p, s, &sol) == CUOPT_INVALID_ARGUMENT);
    EXPECT(cuOptGetErrorString(sol, why, (cuopt_int_t)sizeof why) == CUOPT_SUCCESS && strlen(why) > 0);
    cuOptDestroySolution(&sol);
    cuOptDestroySolverSettings(&s);
  }
  cuOptDestroyProblem(&p);
  return 0;
}

static int case_afiro(const char* path)
{
  cuOptOptimizationProblem p = NULL;
  cuOptSolverSettings s      = NULL;
  cuopt_int_t term = -1;
  double obj;
  EXPECT(cuOptReadProblem(path, &p) == CUOPT_SUCCESS);
  EXPECT(cuOptCreateSolverSettings(&s) == CUOPT_SUCCESS);
  EXPECT(cuOptSetFloatParameter(s, CUOPT_TIME_LIMIT, 60.0) == CUOPT_SUCCESS);
  EXPECT(solve_status(p, s, &term, &obj, NULL) == 0);
  printf("afiro: status %d objective %.9f\n", (int)term, obj);
  EXPECT(term == CUOPT_TERIMINATION_STATUS_OPTIMAL);
  EXPECT(fabs(obj + 464.7531) <= 1e-2 * 464.7531);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_ITERATION_LIMIT, 1) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_METHOD, CUOPT_METHOD_PDLP) == CUOPT_SUCCESS);
  EXPECT(cuOptSetIntegerParameter(s, CUOPT_PRESOLVE, CUOPT_PRESOLVE_OFF) == CUOPT_SUCCESS);
  EXPECT(solve_status(p, s, &term, &obj, NULL) == 0);
  EXPECT(term == CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT);
  cuOptDestroySolverSettings(&s);
  cuOptDestroyProblem(&p);
  return 0;
}

int main(int argc, char** argv)
{
  if (argc < 2) {
    fprintf(stderr, "usage: %s CASE [ARGS]\n", argv[0]);
    return 2;
  }
  const char* c = argv[1];
  int rc        = 2;
  if (!strcmp(c, "sizes")) rc = case_sizes();
  else if (!strcmp(c, "bad_parameter")) rc = case_bad_parameter();
  else if (!strcmp(c, "missing_file")) rc = case_missing_file();
  else if (!strcmp(c, "reference_parameters")) rc = case_reference_parameters();
  else if (!strcmp(c, "write_read") && argc == 4) rc = case_write_read(argv[2], argv[3]);
  else if (!strcmp(c, "infeasible")) rc = case_infeasible();
  else if (!strcmp(c, "ranged")) rc = case_ranged();
  else if (!strcmp(c, "maximize_duals")) rc = case_maximize_duals();
  else if (!strcmp(c, "afiro") && argc == 3) rc = case_afiro(argv[2]);
  else fprintf(stderr, "unknown case %s\n", c);
  if (rc == 0) printf("%s: ok\n", c);
  return rc;
}
