// pdlp_oracle.cpp -- CPU FP64 restatement of the reference PDLP "Stable3" LP path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product library (cuopt_amd/), which must fail loudly if its HIP build is
// missing instead of falling back to anything here.
//
// Every function below cites the reference file:line it restates.  Paths are
// relative to the reference repository root (cpp/src/...).
//
// How each part is pinned (see tests/test_oracle.py):
//   * transpose / bounds mapping ........ exact integer identities (A^T^T == A),
//                                          closed-form row-type mapping
//   * Ruiz + Pock-Chambolle + max-abs step + primal-weight formula
//                                        .. reference KAT on afiro in Methodical1 mode
//                                          (eta0 = 1.4893, omega0 = 0.0141652,
//                                          cpp/tests/linear_programming/pdlp_test.cu:243-287)
//   * power iteration .................... sigma_max(A~)^2 against a dense SVD (numpy)
//   * KKT / termination record ........... closed forms at a constructed optimum (x0,y0,s0):
//                                          r_p = r_d = gap = 0; objective c.x0 == b.y0
//   * PID primal weight, restart decision, fixed-point error, beta/gamma
//                                        .. hand-worked values (tests/test_oracle_pins.py):
//                                          a PID sequence through the fall-back branch, a
//                                          decision table at every threshold boundary, the
//                                          PDHG-metric norm from an explicit dense matrix, a
//                                          diagonal LP with E/L/G/ranged/free rows; each check
//                                          fails on an oracle build with a planted mistake
//   * end-to-end solve ................... closed-form optima of the generated LPs,
//                                          afiro -464.7531, good-max 17, max_offset 0,
//                                          HiGHS (scipy) optima committed as numbers,
//                                          fixed-point property of PDHG at the optimum.
//   Nothing here is "parity unpinned" except the exact iteration trajectory of the
//   reference CUDA build (cuSPARSE/cuBLAS summation order), which no offline check can pin.
//
// Arithmetic follows the reference expression by expression (same operand order,
// separate roundings where the reference uses separate transforms); it is compiled
// with -ffp-contract=off so no FMA is formed.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// Termination status codes (cpp/include/cuopt/linear_programming/constants.h:83-93).
enum Status {
  NoTermination   = 0,
  Optimal         = 1,
  PrimalInfeasible = 2,
  DualInfeasible  = 3,
  IterationLimit  = 4,
  TimeLimit       = 5,
  NumericalError  = 6,
  PrimalFeasible  = 7,
};

struct Csr {
  int nrows = 0, ncols = 0;
  std::vector<int> ptr, idx;
  std::vector<double> val;
};

// Hyper parameters.  Stable3 values: cpp/src/pdlp/solve.cu:167-207 over the
// defaults of cpp/include/cuopt/linear_programming/pdlp/pdlp_hyper_params.cuh:12-56.
// The Methodical1 subset (solve.cu:210-245) is kept only for the initial
// step-size / primal-weight known-answer test.
struct Hyper {
  int ruiz_iters                = 10;
  double pc_alpha               = 1.0;
  bool bound_objective_rescaling = true;
  bool step_from_max_singular   = true;   // initial_step_size_max_singular_value
  bool weight_before_scaling    = true;   // compute_initial_primal_weight_before_scaling
  bool weight_combined_bounds   = false;  // initial_primal_weight_combined_bounds
  double primal_importance      = 1.0;
  double initial_step_scaling   = 1.0;
  int major_iteration           = 200;
  double artificial_threshold   = 0.36;
  double sufficient_reduction   = 0.2;
  double necessary_reduction    = 0.8;
  double k_p = 0.99, k_i = 0.01, k_d = 0.0, i_smooth = 0.3;
};

int g_threads = 1;

// y = A x, each row summed left to right from 0.0 (row-parallel, deterministic).
void spmv(const Csr& A, const double* x, double* y)
{
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int i = 0; i < A.nrows; ++i) {
    double s = 0.0;
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[k] * x[A.idx[k]];
    y[i] = s;
  }
}

double dot(const double* a, const double* b, size_t n)
{
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

double nrm2(const double* a, size_t n) { return std::sqrt(dot(a, a, n)); }

// Stable counting-sort transpose: within each output row the entries keep the
// input row order (compute_transpose_of_problem, cpp/src/mip_heuristics/problem/problem.cu:472-489,
// csr2csc of a row-sorted CSR).
Csr transpose(const Csr& A)
{
  Csr T;
  T.nrows = A.ncols;
  T.ncols = A.nrows;
  T.ptr.assign(A.ncols + 1, 0);
  const int nnz = A.ptr[A.nrows];
  T.idx.resize(nnz);
  T.val.resize(nnz);
  for (int k = 0; k < nnz; ++k) T.ptr[A.idx[k] + 1]++;
  for (int j = 0; j < A.ncols; ++j) T.ptr[j + 1] += T.ptr[j];
  std::vector<int> cur(T.ptr.begin(), T.ptr.end() - 1);
  for (int i = 0; i < A.nrows; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int d = cur[A.idx[k]]++;
      T.idx[d]    = i;
      T.val[d]    = A.val[k];
    }
  return T;
}

inline double fin0(double v) { return std::isfinite(v) ? v : 0.0; }

// compute_sum_bounds / bound_objective_rescaling reduction
// (cpp/src/pdlp/utils.cuh:251-296, initial_scaling.cu:94-149).
double sum_bounds_sq(const std::vector<double>& lo, const std::vector<double>& up)
{
  double s = 0.0;
  for (size_t i = 0; i < lo.size(); ++i) {
    double t = 0.0;
    if (std::isfinite(lo[i]) && lo[i] != up[i]) t += lo[i] * lo[i];
    if (std::isfinite(up[i])) t += up[i] * up[i];
    s += t;
  }
  return s;
}

// combine_finite_abs_bounds (cpp/src/pdlp/utils.cuh:176-185).
double combined_bounds_norm(const std::vector<double>& lo, const std::vector<double>& up)
{
  double s = 0.0;
  for (size_t i = 0; i < lo.size(); ++i) {
    double v = 0.0;
    if (std::isfinite(up[i])) v = std::max(v, std::fabs(up[i]));
    if (std::isfinite(lo[i])) v = std::max(v, std::fabs(lo[i]));
    s += v * v;
  }
  return std::sqrt(s);
}

// Problem after problem_t construction (cpp/src/mip_heuristics/problem/problem.cu:47-100):
// constraint bounds from row types, default variable bounds [0, inf), maximisation
// folded into the sign of c with objective scaling factor -1 and negated offset
// (problem_helpers.cuh:132-151).
struct Problem {
  int m = 0, n = 0;
  Csr A, AT;
  std::vector<double> c, lc, uc, lb, ub;
  double obj_sf = 1.0, obj_off = 0.0;
  bool maximize = false;
};

struct Scaling {
  std::vector<double> Dr, Dc;  // cummulative_constraint_matrix_scaling_, cummulative_variable_scaling_
  double beta = 1.0, gamma = 1.0;
};

// Ruiz L-inf equilibration (initial_scaling.cu:151-220) then Pock-Chambolle
// (initial_scaling.cu:233-364).  Row and column norms of one sweep are taken
// from the same (D_r, D_c) and applied together.
void compute_scaling_vectors(const Problem& P, const Hyper& H, Scaling& S)
{
  const int m = P.m, n = P.n;
  S.Dr.assign(m, 1.0);
  S.Dc.assign(n, 1.0);
  std::vector<double> rowv(m), colv(n);
  for (int it = 0; it < H.ruiz_iters; ++it) {
    std::fill(rowv.begin(), rowv.end(), 0.0);
    std::fill(colv.begin(), colv.end(), 0.0);
    for (int i = 0; i < m; ++i)
      for (int k = P.A.ptr[i]; k < P.A.ptr[i + 1]; ++k) {
        const int j  = P.A.idx[k];
        const double a = std::fabs((P.A.val[k] * S.Dr[i]) * S.Dc[j]);
        rowv[i]        = std::max(rowv[i], a);
        colv[j]        = std::max(colv[j], a);
      }
    // a_divides_sqrt_b_bounded (utils.cuh:154-161)
    for (int i = 0; i < m; ++i) S.Dr[i] = rowv[i] > 0.0 ? S.Dr[i] / std::sqrt(rowv[i]) : S.Dr[i];
    for (int j = 0; j < n; ++j) S.Dc[j] = colv[j] > 0.0 ? S.Dc[j] / std::sqrt(colv[j]) : S.Dc[j];
  }
  // Pock-Chambolle: row sums over A, column sums over A^T, each in storage order.
  for (int i = 0; i < m; ++i) {
    double s = 0.0;
    for (int k = P.A.ptr[i]; k < P.A.ptr[i + 1]; ++k) {
      const double a = std::fabs((P.A.val[k] * S.Dr[i]) * S.Dc[P.A.idx[k]]);
      s += std::pow(a, H.pc_alpha);
    }
    rowv[i] = s;
  }
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int k = P.AT.ptr[j]; k < P.AT.ptr[j + 1]; ++k) {
      const double a = std::fabs((P.AT.val[k] * S.Dr[P.AT.idx[k]]) * S.Dc[j]);
      s += std::pow(a, 2.0 - H.pc_alpha);
    }
    colv[j] = s;
  }
  for (int i = 0; i < m; ++i) S.Dr[i] = rowv[i] > 0.0 ? S.Dr[i] / std::sqrt(rowv[i]) : S.Dr[i];
  for (int j = 0; j < n; ++j) S.Dc[j] = colv[j] > 0.0 ? S.Dc[j] / std::sqrt(colv[j]) : S.Dc[j];
}

// scale_problem (initial_scaling.cu:416-568).  Returns the scaled problem.
// A~ = (a * D_r) * D_c; the transposed copy multiplies in the order of
// scale_transposed_problem_kernel, (a * D_c) * D_r.
Problem scale_problem(const Problem& P, const Hyper& H, Scaling& S)
{
  Problem Q = P;
  for (int i = 0; i < P.m; ++i)
    for (int k = P.A.ptr[i]; k < P.A.ptr[i + 1]; ++k)
      Q.A.val[k] = P.A.val[k] * S.Dr[i] * S.Dc[P.A.idx[k]];
  for (int j = 0; j < P.n; ++j)
    for (int k = P.AT.ptr[j]; k < P.AT.ptr[j + 1]; ++k)
      Q.AT.val[k] = P.AT.val[k] * S.Dc[j] * S.Dr[P.AT.idx[k]];
  for (int j = 0; j < P.n; ++j) {
    Q.c[j] = P.c[j] * S.Dc[j];
    // divide_check_zero (utils.cuh:270-280)
    if (S.Dc[j] == 0.0) {
      Q.lb[j] = 0.0;
      Q.ub[j] = 0.0;
    } else {
      Q.lb[j] = P.lb[j] / S.Dc[j];
      Q.ub[j] = P.ub[j] / S.Dc[j];
    }
  }
  for (int i = 0; i < P.m; ++i) {
    Q.lc[i] = P.lc[i] * S.Dr[i];
    Q.uc[i] = P.uc[i] * S.Dr[i];
  }
  S.beta = 1.0;
  S.gamma = 1.0;
  if (H.bound_objective_rescaling) {
    // bound_objective_rescaling (initial_scaling.cu:94-149), on the scaled bounds and c.
    S.beta  = 1.0 / (std::sqrt(sum_bounds_sq(Q.lc, Q.uc)) + 1.0);
    double cs = 0.0;
    for (int j = 0; j < P.n; ++j) cs += Q.c[j] * Q.c[j] * 1.0;
    S.gamma = 1.0 / (std::sqrt(cs) + 1.0);
    for (int i = 0; i < P.m; ++i) {
      Q.lc[i] = Q.lc[i] * S.beta;
      Q.uc[i] = Q.uc[i] * S.beta;
    }
    for (int j = 0; j < P.n; ++j) {
      Q.lb[j] = Q.lb[j] * S.beta;
      Q.ub[j] = Q.ub[j] * S.beta;
      Q.c[j]  = Q.c[j] * S.gamma;
    }
  }
  return Q;
}

// scale_solutions / unscale_solutions (initial_scaling.cu:570-770): one rounding
// per transform, the second transform multiplying by beta, gamma or their
// host-computed reciprocals.
void scale_primal(const Scaling& S, bool bor, std::vector<double>& x)
{
  for (size_t j = 0; j < x.size(); ++j) {
    double v = S.Dc[j] != 0.0 ? x[j] / S.Dc[j] : x[j];
    if (bor) v = v * S.beta;
    x[j] = v;
  }
}
void scale_dual(const Scaling& S, bool bor, std::vector<double>& y)
{
  for (size_t i = 0; i < y.size(); ++i) {
    double v = S.Dr[i] != 0.0 ? y[i] / S.Dr[i] : y[i];
    if (bor) v = v * S.gamma;
    y[i] = v;
  }
}
void scale_slack(const Scaling& S, bool bor, std::vector<double>& z)
{
  for (size_t j = 0; j < z.size(); ++j) {
    double v = z[j] * S.Dc[j];
    if (bor) v = v * S.gamma;
    z[j] = v;
  }
}
void unscale_primal(const Scaling& S, bool bor, std::vector<double>& x)
{
  const double ib = 1.0 / S.beta;
  for (size_t j = 0; j < x.size(); ++j) {
    double v = x[j] * S.Dc[j];
    if (bor) v = v * ib;
    x[j] = v;
  }
}
void unscale_dual(const Scaling& S, bool bor, std::vector<double>& y)
{
  const double ig = 1.0 / S.gamma;
  for (size_t i = 0; i < y.size(); ++i) {
    double v = y[i] * S.Dr[i];
    if (bor) v = v * ig;
    y[i] = v;
  }
}
void unscale_slack(const Scaling& S, bool bor, std::vector<double>& z)
{
  const double ig = 1.0 / S.gamma;
  for (size_t j = 0; j < z.size(); ++j) {
    double v = S.Dc[j] != 0.0 ? z[j] / S.Dc[j] : z[j];
    if (bor) v = v * ig;
    z[j] = v;
  }
}

// compute_initial_step_size, power-iteration branch (cpp/src/pdlp/pdlp.cu:2587-2739).
// Returns sigma_max^2 and the iteration count.
double power_iteration(const Problem& Q, int* iters_out)
{
  const int m = Q.m, n = Q.n;
  std::vector<double> z(m), q(m), atq(n);
  std::mt19937 gen(1);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (int i = 0; i < m; ++i) z[i] = dist(gen);
  double sigma_sq = 0.0;
  int it          = 0;
  for (int i = 0; i < 5000; ++i) {
    ++it;
    q              = z;
    const double nq = nrm2(q.data(), m);
    for (int r = 0; r < m; ++r) q[r] = q[r] / nq;
    spmv(Q.AT, q.data(), atq.data());
    spmv(Q.A, atq.data(), z.data());
    sigma_sq = dot(q.data(), z.data(), m);
    for (int r = 0; r < m; ++r) q[r] = q[r] * -sigma_sq + z[r];
    const double res = nrm2(q.data(), m);
    if (res < 1e-4) break;
  }
  if (iters_out) *iters_out = it;
  return sigma_sq;
}

// One evaluation of convergence_information_t::compute_convergence_information
// in reflected mode (convergence_information.cu:326-483, 546-559, 574-606,
// 649-690, 700-810) on the unscaled problem for (x, y, slack).
// linf_primal / linf_dual: the per_constraint_residual maxima (see below).
struct Kkt {
  double l2_primal = 0, l2_dual = 0, pobj = 0, dobj = 0, gap = 0, absobj = 0;
  double linf_primal = 0, linf_dual = 0;
};

// combine_finite_abs_bounds (cpp/src/pdlp/utils.cuh:180-188): the larger finite
// bound magnitude of a row, 0 when both bounds are infinite.
inline double combined_abs_bound(double lo, double up)
{
  double v = 0.0;
  if (std::isfinite(up)) v = std::max(v, std::fabs(up));
  if (std::isfinite(lo)) v = std::max(v, std::fabs(lo));
  return v;
}

// rel_primal / rel_dual: the relative primal / dual tolerances, used for the
// per_constraint_residual maxima max(0, max_i (r_i - rel * w_i)) of
// convergence_information.cu:390-416 (primal: w = combined_abs_bound) and :453-470
// (dual: w = c_j), with the signed residuals and signed c_j of relative_residual_t
// (utils.cuh:576-590) and 0 as the neutral element of the maximum.
Kkt convergence_information(const Problem& P,
                            const std::vector<double>& x,
                            const std::vector<double>& y,
                            const std::vector<double>& slack,
                            std::vector<double>& tmp_m,
                            std::vector<double>& tmp_n,
                            double rel_primal = 0.0,
                            double rel_dual   = 0.0)
{
  Kkt K;
  spmv(P.A, x.data(), tmp_m.data());
  double rp2 = 0.0, psum = 0.0, linf_p = 0.0;
  for (int i = 0; i < P.m; ++i) {
    const double Ax  = tmp_m[i];
    const double cl  = std::max(P.lc[i], std::min(Ax, P.uc[i]));
    const double r   = Ax - cl;
    rp2 += r * r;
    linf_p = std::max(linf_p, r - rel_primal * combined_abs_bound(P.lc[i], P.uc[i]));
    psum += std::max(y[i], 0.0) * fin0(P.lc[i]) + std::min(y[i], 0.0) * fin0(P.uc[i]);
  }
  K.linf_primal = linf_p;
  double pobj = dot(x.data(), P.c.data(), P.n);
  if (P.obj_sf != 1.0 || P.obj_off != 0.0) pobj = P.obj_sf * (pobj + P.obj_off);
  K.l2_primal = std::sqrt(rp2);
  spmv(P.AT, y.data(), tmp_n.data());
  double rd2 = 0.0, linf_d = 0.0;
  for (int j = 0; j < P.n; ++j) {
    const double g = P.c[j] - tmp_n[j];
    const double r = g - slack[j];
    rd2 += r * r;
    linf_d = std::max(linf_d, r - rel_dual * P.c[j]);
  }
  K.l2_dual   = std::sqrt(rd2);
  K.linf_dual = linf_d;
  double dobj = dot(slack.data(), x.data(), P.n) + psum;
  if (P.obj_sf != 1.0 || P.obj_off != 0.0) dobj = P.obj_sf * (dobj + P.obj_off);
  K.pobj   = pobj;
  K.dobj   = dobj;
  K.gap    = std::fabs(pobj - dobj);
  K.absobj = std::fabs(pobj) + std::fabs(dobj);
  return K;
}

// check_termination_criteria_kernel (termination_strategy.cu:222-375): the l2
// branch, or with per_constraint the L-inf branch (:317-332), whose maxima already
// hold the relative terms.
int termination_status(const Kkt& K, const double* tol, double b_norm, double c_norm, bool per_constraint = false)
{
  // tol: abs_primal, rel_primal, abs_dual, rel_dual, abs_gap, rel_gap
  const bool optimal_gap     = K.gap <= tol[4] + tol[5] * K.absobj;
  const bool primal_feasible = per_constraint ? K.linf_primal <= tol[0] : K.l2_primal <= tol[0] + tol[1] * b_norm;
  const bool dual_feasible   = per_constraint ? K.linf_dual <= tol[2] : K.l2_dual <= tol[2] + tol[3] * c_norm;
  if (dual_feasible && primal_feasible && optimal_gap) return Optimal;
  if (primal_feasible) return PrimalFeasible;
  return NoTermination;
}

// conditional_major (cpp/src/pdlp/utils.cuh:62-81).
int64_t conditional_major(int64_t k)
{
  int64_t step = 10, threshold = 1000;
  while (k >= threshold) {
    step *= 10;
    threshold *= 10;
  }
  return step;
}

// cupdlpx_new_primal_weight_computation (pdlp_restart_strategy.cu:757-835).
struct Pid {
  double err_sum = 0.0, last_err = 0.0, best_gap = kInf;
};
void pid_update(double primal_distance,
                double dual_distance,
                double rel_dual,
                double rel_primal,
                double step_size,
                const Hyper& H,
                Pid& pid,
                double& omega,
                double& best_omega,
                double& tau,
                double& sigma)
{
  const double l2p   = std::sqrt(primal_distance);
  const double l2d   = std::sqrt(dual_distance);
  const double ratio = (rel_primal == 0.0) ? kInf : rel_dual / rel_primal;
  if (l2p > 1e-16 && l2d > 1e-16 && l2p < 1e12 && l2d < 1e12 && ratio > 1e-8 && ratio < 1e8) {
    const double cur = omega;
    const double err = std::log(l2d) - std::log(l2p) - std::log(cur);
    const double s   = pid.err_sum * H.i_smooth + err;
    pid.err_sum      = s;
    const double de  = err - pid.last_err;
    const double w   = std::exp(H.k_p * err + H.k_i * s + H.k_d * de) * cur;
    omega            = w;
    pid.last_err     = err;
    tau              = step_size / w;
    sigma            = step_size * w;
  } else {
    omega        = best_omega;
    pid.err_sum  = 0.0;
    pid.last_err = 0.0;
    tau          = step_size / best_omega;
    sigma        = step_size * best_omega;
  }
  const double g = std::fabs(std::log10(ratio));
  if (g < pid.best_gap) {
    pid.best_gap = g;
    best_omega   = omega;
  }
}

// should_cupdlpx_restart (pdlp_restart_strategy.cu:688-755) for one LP, with the
// artificial criterion of should_do_artificial_restart (:1397-1419).  k counts all
// iterations, k_r those since the last restart; last_trial is the FPE of the
// previous major check.
bool restart_decision(int64_t k, int64_t k_r, double fpe, double fpe0, double last_trial,
                      const Hyper& H)
{
  if (k == H.major_iteration) return true;  // forced at the first major
  if (k < H.major_iteration) return false;
  if (fpe <= H.sufficient_reduction * fpe0) return true;
  if (fpe <= H.necessary_reduction * fpe0 && fpe > last_trial) return true;
  return (double)k_r >= H.artificial_threshold * (double)k;
}

// compute_fixed_error (pdlp.cu:1748-1855): dx = x_bar - x, dy = y_bar - y, the
// interaction (A~^T y_bar - A~^T y) . dx of compute_interaction_and_movement
// (adaptive_step_size_strategy.cu:351-517), and fixed_error_computation
// (pdlp.cu:1401-1420): sqrt(omega |dx|^2 + |dy|^2 / omega + 2 eta interaction).
// aty holds A~^T y; tmp_n (n) receives A~^T y_bar.
double fixed_point_error(const Csr& AT, const std::vector<double>& x, const std::vector<double>& xbar,
                         const std::vector<double>& y, const std::vector<double>& ybar,
                         const std::vector<double>& aty, double omega, double eta,
                         std::vector<double>& tmp_n)
{
  spmv(AT, ybar.data(), tmp_n.data());
  double inter = 0.0, nx = 0.0, ny = 0.0;
  for (size_t j = 0; j < x.size(); ++j) {
    const double dx = xbar[j] - x[j];
    const double t  = tmp_n[j] - aty[j];
    inter += t * dx;
    nx += dx * dx;
  }
  for (size_t i = 0; i < y.size(); ++i) {
    const double dy = ybar[i] - y[i];
    ny += dy * dy;
  }
  const double movement = nx * omega + ny / omega;
  const double ci       = 2.0 * inter * eta;
  return std::sqrt(movement + ci);
}

// Parameter and info slots shared with oracle/oracle.py.
enum ParamSlot {
  P_ABS_PRIMAL = 0, P_REL_PRIMAL, P_ABS_DUAL, P_REL_DUAL, P_ABS_GAP, P_REL_GAP,
  P_ITER_LIMIT, P_TIME_LIMIT, P_THREADS, P_MODE, P_PER_CONSTRAINT, P_NPARAMS
};
enum InfoSlot {
  I_STATUS = 0, I_ITERS, I_L2_PRIMAL, I_REL_PRIMAL, I_L2_DUAL, I_REL_DUAL, I_POBJ, I_DOBJ,
  I_GAP, I_REL_GAP, I_SOLVE_TIME, I_ETA0, I_OMEGA0, I_BETA, I_GAMMA, I_OMEGA, I_RESTARTS,
  I_POWER_ITERS, I_SIGMA_SQ, I_B_NORM, I_C_NORM, I_SETUP_TIME, I_NINFO
};
constexpr int kTraceWidth = 10;  // k, l2p, l2d, pobj, dobj, gap, restarted, omega, fpe, k_r

Hyper hyper_for_mode(int mode)
{
  Hyper H;
  if (mode == 1) {  // Methodical1 subset (solve.cu:210-245)
    H.ruiz_iters              = 5;
    H.bound_objective_rescaling = false;
    H.step_from_max_singular  = false;
    H.weight_before_scaling   = false;
    H.weight_combined_bounds  = true;
  }
  return H;
}

bool build_problem(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                   const double* c, const double* lc, const double* uc, const double* lb,
                   const double* ub, int maximize, double offset, Problem& P)
{
  P.m = m;
  P.n = n;
  P.A.nrows = m;
  P.A.ncols = n;
  P.A.ptr.assign(rowptr, rowptr + m + 1);
  const int nnz = rowptr[m];
  P.A.idx.assign(colidx, colidx + nnz);
  P.A.val.assign(vals, vals + nnz);
  // Column indices sorted within each row (stable), as problem_t keeps a sorted CSR.
  for (int i = 0; i < m; ++i) {
    const int b = rowptr[i], e = rowptr[i + 1];
    bool sorted = true;
    for (int k = b + 1; k < e; ++k) sorted &= P.A.idx[k - 1] <= P.A.idx[k];
    if (!sorted) {
      std::vector<std::pair<int, double>> row;
      for (int k = b; k < e; ++k) row.emplace_back(P.A.idx[k], P.A.val[k]);
      std::stable_sort(row.begin(), row.end(),
                       [](const auto& a, const auto& b2) { return a.first < b2.first; });
      for (int k = b; k < e; ++k) {
        P.A.idx[k] = row[k - b].first;
        P.A.val[k] = row[k - b].second;
      }
    }
  }
  P.AT = transpose(P.A);
  P.c.assign(c, c + n);
  P.lc.assign(lc, lc + m);
  P.uc.assign(uc, uc + m);
  P.lb.assign(n, 0.0);
  P.ub.assign(n, kInf);
  if (lb) P.lb.assign(lb, lb + n);
  if (ub) P.ub.assign(ub, ub + n);
  P.maximize = maximize != 0;
  P.obj_off  = offset;
  if (P.maximize) {
    for (double& v : P.c) v = -v;
    P.obj_sf  = -1.0;
    P.obj_off = -offset;
  }
  for (int j = 0; j < n; ++j)
    if (!(P.ub[j] >= P.lb[j])) return false;
  for (int i = 0; i < m; ++i)
    if (!(P.uc[i] >= P.lc[i])) return false;
  return true;
}

struct Setup {
  Scaling S;
  Problem Q;        // scaled problem
  double eta = 1.0, omega = 1.0, best_omega = 1.0, sigma_sq = 0.0;
  int power_iters = 0;
};

// Initial primal weight (pdlp.cu:2775-2832), scaling, initial step size, in the
// order of run_solver (pdlp.cu:2026-2040).
void setup(const Problem& P, const Hyper& H, Setup& U)
{
  auto primal_weight = [&](const Problem& X) {
    double cn = 0.0;
    for (double v : X.c) cn += v * v * 1.0;
    cn = std::sqrt(cn);
    double bn;
    if (H.weight_combined_bounds) {
      bn = combined_bounds_norm(X.lc, X.uc);
    } else if (H.bound_objective_rescaling) {
      return 1.0;
    } else {
      bn = std::sqrt(sum_bounds_sq(X.lc, X.uc));
    }
    if (bn > 0.0 && cn > 0.0) return H.primal_importance * (cn / bn);
    return H.primal_importance;
  };
  if (H.weight_before_scaling) U.omega = primal_weight(P);
  compute_scaling_vectors(P, H, U.S);
  U.Q = scale_problem(P, H, U.S);
  if (H.step_from_max_singular) {
    U.sigma_sq = power_iteration(U.Q, &U.power_iters);
    U.eta      = 0.998 / std::sqrt(U.sigma_sq);
  } else {
    double mx = 0.0;
    for (double v : U.Q.A.val) mx = std::fabs(mx) < std::fabs(v) ? std::fabs(v) : std::fabs(mx);
    U.eta = mx != 0.0 ? H.initial_step_scaling / mx : H.initial_step_scaling;
  }
  if (!H.weight_before_scaling) U.omega = primal_weight(U.Q);
  U.best_omega = U.omega;
}

}  // namespace

extern "C" {

int oracle_num_params() { return P_NPARAMS; }
int oracle_num_info() { return I_NINFO; }
int oracle_trace_width() { return kTraceWidth; }

// y = A x in FP64 (row sums left to right).
void oracle_spmv(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                 const double* x, double* y)
{
  Csr A;
  A.nrows = m;
  A.ncols = n;
  A.ptr.assign(rowptr, rowptr + m + 1);
  A.idx.assign(colidx, colidx + rowptr[m]);
  A.val.assign(vals, vals + rowptr[m]);
  spmv(A, x, y);
}

// Transpose to CSR of A^T (stable).
void oracle_transpose(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                      int* t_ptr, int* t_idx, double* t_val)
{
  Csr A;
  A.nrows = m;
  A.ncols = n;
  A.ptr.assign(rowptr, rowptr + m + 1);
  A.idx.assign(colidx, colidx + rowptr[m]);
  A.val.assign(vals, vals + rowptr[m]);
  Csr T = transpose(A);
  std::copy(T.ptr.begin(), T.ptr.end(), t_ptr);
  std::copy(T.idx.begin(), T.idx.end(), t_idx);
  std::copy(T.val.begin(), T.val.end(), t_val);
}

// Setup quantities: D_r (m), D_c (n), out[0..7] = beta, gamma, eta, omega,
// sigma_sq, power_iters, b_norm, c_norm.  mode 0 = Stable3, 1 = Methodical1.
int oracle_setup(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                 const double* c, const double* lc, const double* uc, const double* lb,
                 const double* ub, int maximize, int mode, double* Dr, double* Dc, double* out,
                 double* scaled_vals)
{
  Problem P;
  if (!build_problem(m, n, rowptr, colidx, vals, c, lc, uc, lb, ub, maximize, 0.0, P)) return -1;
  Hyper H = hyper_for_mode(mode);
  Setup U;
  setup(P, H, U);
  if (Dr) std::copy(U.S.Dr.begin(), U.S.Dr.end(), Dr);
  if (Dc) std::copy(U.S.Dc.begin(), U.S.Dc.end(), Dc);
  if (scaled_vals) std::copy(U.Q.A.val.begin(), U.Q.A.val.end(), scaled_vals);
  out[0] = U.S.beta;
  out[1] = U.S.gamma;
  out[2] = U.eta;
  out[3] = U.omega;
  out[4] = U.sigma_sq;
  out[5] = U.power_iters;
  out[6] = std::sqrt(sum_bounds_sq(P.lc, P.uc));
  double cn = 0.0;
  for (double v : P.c) cn += v * v;
  out[7] = std::sqrt(cn);
  return 0;
}

// KKT record on the unscaled problem for a given (x, y, slack):
// out = l2_primal, l2_dual, pobj, dobj, gap, absobj, b_norm, c_norm, linf_primal, linf_dual.
// out[8], out[9] = the per_constraint_residual maxima for the relative
// tolerances rel_primal, rel_dual.
void oracle_kkt(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                const double* c, const double* lc, const double* uc, int maximize, double offset,
                const double* x, const double* y, const double* slack, double rel_primal,
                double rel_dual, double* out)
{
  Problem P;
  build_problem(m, n, rowptr, colidx, vals, c, lc, uc, nullptr, nullptr, maximize, offset, P);
  std::vector<double> xv(x, x + n), yv(y, y + m), sv(slack, slack + n), tm(m), tn(n);
  Kkt K   = convergence_information(P, xv, yv, sv, tm, tn, rel_primal, rel_dual);
  out[8]  = K.linf_primal;
  out[9]  = K.linf_dual;
  out[0]  = K.l2_primal;
  out[1]  = K.l2_dual;
  out[2]  = K.pobj;
  out[3]  = K.dobj;
  out[4]  = K.gap;
  out[5]  = K.absobj;
  out[6]  = std::sqrt(sum_bounds_sq(P.lc, P.uc));
  double cn = 0.0;
  for (double v : P.c) cn += v * v;
  out[7] = std::sqrt(cn);
}

// One PID primal-weight update (pid_update above) on an explicit state with the
// Stable3 gains.  state (in/out, 7 doubles): omega, best_omega, error_sum,
// last_error, best_residual_gap, tau, sigma.
void oracle_pid_update(double primal_distance, double dual_distance, double rel_dual,
                       double rel_primal, double step_size, double* state)
{
  const Hyper H = hyper_for_mode(0);
  Pid pid;
  pid.err_sum  = state[2];
  pid.last_err = state[3];
  pid.best_gap = state[4];
  double omega = state[0], best = state[1], tau = state[5], sigma = state[6];
  pid_update(primal_distance, dual_distance, rel_dual, rel_primal, step_size, H, pid, omega, best,
             tau, sigma);
  state[0] = omega;
  state[1] = best;
  state[2] = pid.err_sum;
  state[3] = pid.last_err;
  state[4] = pid.best_gap;
  state[5] = tau;
  state[6] = sigma;
}

// The Stable3 restart decision at a major check (restart_decision above): 1 = restart.
int oracle_restart_decision(long long k, long long k_r, double fpe, double fpe0, double last_trial)
{
  return restart_decision(k, k_r, fpe, fpe0, last_trial, hyper_for_mode(0)) ? 1 : 0;
}

// Fixed-point error of the step (x, y) -> (x_bar, y_bar) under the CSR matrix A
// (in a solve, the scaled matrix), primal weight omega and step size eta.
double oracle_fpe(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                  const double* x, const double* xbar, const double* y, const double* ybar,
                  double omega, double eta)
{
  Csr A;
  A.nrows = m;
  A.ncols = n;
  A.ptr.assign(rowptr, rowptr + m + 1);
  A.idx.assign(colidx, colidx + rowptr[m]);
  A.val.assign(vals, vals + rowptr[m]);
  const Csr AT = transpose(A);
  std::vector<double> xv(x, x + n), xb(xbar, xbar + n), yv(y, y + m), yb(ybar, ybar + m);
  std::vector<double> aty(n), tmp(n);
  spmv(AT, yv.data(), aty.data());
  return fixed_point_error(AT, xv, xb, yv, yb, aty, omega, eta, tmp);
}

// Full Stable3 solve (run_pdlp -> pdlp_solver_t::run_solver, pdlp.cu:2013-2458).
// Returns the termination status.  x_out (n), y_out (m), rc_out (n) receive the
// unscaled returned iterate (maximisation: y and reduced costs negated, solve.cu:583-588).
// trace (optional): one row of kTraceWidth doubles per termination check.
int oracle_solve(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                 const double* c, const double* lc, const double* uc, const double* lb,
                 const double* ub, int maximize, double offset, const double* params,
                 const double* x_init, const double* y_init, double* x_out, double* y_out,
                 double* rc_out, double* info, double* trace, int trace_cap, int* trace_len)
{
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed  = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  for (int i = 0; i < I_NINFO; ++i) info[i] = std::nan("");
  if (trace_len) *trace_len = 0;
  g_threads = std::max(1, (int)params[P_THREADS]);
  const double tol[6] = {params[P_ABS_PRIMAL], params[P_REL_PRIMAL], params[P_ABS_DUAL],
                         params[P_REL_DUAL],   params[P_ABS_GAP],    params[P_REL_GAP]};
  const bool per_constraint     = params[P_PER_CONSTRAINT] != 0.0;
  const int64_t iteration_limit = (int64_t)params[P_ITER_LIMIT];
  const double time_limit       = params[P_TIME_LIMIT];

  Problem P;
  if (!build_problem(m, n, rowptr, colidx, vals, c, lc, uc, lb, ub, maximize, offset, P)) {
    info[I_STATUS] = PrimalInfeasible;
    info[I_ITERS]  = 0;
    return PrimalInfeasible;  // has_crossing_bounds (solve.cu:1176-1179)
  }
  if (m == 0) {  // run_pdlp_solver (solve.cu:555-572)
    info[I_STATUS] = NumericalError;
    info[I_ITERS]  = 0;
    return NumericalError;
  }
  const Hyper H = hyper_for_mode(0);
  const double b_norm = std::sqrt(sum_bounds_sq(P.lc, P.uc));
  double c_norm       = 0.0;
  for (double v : P.c) c_norm += v * v;
  c_norm = std::sqrt(c_norm);

  Setup U;
  setup(P, H, U);
  const Problem& Q = U.Q;
  const Scaling& S = U.S;
  const bool bor   = H.bound_objective_rescaling;
  double omega = U.omega, best_omega = U.best_omega, eta = U.eta;
  double tau = eta / omega, sigma = eta * omega;
  info[I_ETA0]        = eta;
  info[I_OMEGA0]      = omega;
  info[I_BETA]        = S.beta;
  info[I_GAMMA]       = S.gamma;
  info[I_POWER_ITERS] = U.power_iters;
  info[I_SIGMA_SQ]    = U.sigma_sq;
  info[I_B_NORM]      = b_norm;
  info[I_C_NORM]      = c_norm;
  info[I_SETUP_TIME]  = elapsed();

  std::vector<double> x(n, 0.0), y(m, 0.0), xhat(n, 0.0), yhat(m, 0.0), slack(n, 0.0);
  std::vector<double> x0(n, 0.0), y0(m, 0.0), xbar(n), ybar(m), aty(n), ax(m);
  std::vector<double> tmp_m(m), tmp_n(n);
  // update_primal_dual_solutions (pdlp.cu:1236-1393): the initial point seeds
  // x, x^ and the restart anchor, then all three are scaled.
  if (x_init) {
    x.assign(x_init, x_init + n);
    xhat = x;
    x0   = x;
    scale_primal(S, bor, x);
    scale_primal(S, bor, xhat);
    scale_primal(S, bor, x0);
  }
  if (y_init) {
    y.assign(y_init, y_init + m);
    yhat = y;
    y0   = y;
    scale_dual(S, bor, y);
    scale_dual(S, bor, yhat);
    scale_dual(S, bor, y0);
  }
  // project_initial_primal (pdlp.cu:2129-2154): only the current iterate.
  for (int j = 0; j < n; ++j) x[j] = std::min(std::max(x[j], Q.lb[j]), Q.ub[j]);

  const double nan = std::numeric_limits<double>::quiet_NaN();
  double fpe = nan, fpe0 = nan, last_trial = kInf;
  Pid pid;
  int64_t k = 0, k_r = 0;
  int restarts = 0;
  int status   = NoTermination;
  Kkt K;
  auto finish = [&](int st) {
    // fill_return_problem_solution (termination_strategy.cu:518-634) + run_pdlp sign flip.
    for (int j = 0; j < n; ++j) x_out[j] = xhat[j];
    for (int i = 0; i < m; ++i) y_out[i] = P.maximize ? -yhat[i] : yhat[i];
    for (int j = 0; j < n; ++j) rc_out[j] = P.maximize ? -slack[j] : slack[j];
    info[I_STATUS]     = st;
    info[I_ITERS]      = (double)k;
    info[I_L2_PRIMAL]  = K.l2_primal;
    info[I_REL_PRIMAL] = K.l2_primal / (1.0 + b_norm);
    info[I_L2_DUAL]    = K.l2_dual;
    info[I_REL_DUAL]   = K.l2_dual / (1.0 + c_norm);
    info[I_POBJ]       = K.pobj;
    info[I_DOBJ]       = K.dobj;
    info[I_GAP]        = K.gap;
    info[I_REL_GAP]    = K.gap / (1.0 + K.absobj);
    info[I_SOLVE_TIME] = elapsed();
    info[I_OMEGA]      = omega;
    info[I_RESTARTS]   = restarts;
    return st;
  };

  while (true) {
    const bool is_major = ((k % H.major_iteration) == 0 && k > 0) || k <= 0;
    const bool is_cond  = (k % conditional_major(k)) == 0;
    bool restarted      = false;
    if (is_major || is_cond) {
      unscale_primal(S, bor, xhat);
      unscale_dual(S, bor, yhat);
      unscale_slack(S, bor, slack);
      K = convergence_information(P, xhat, yhat, slack, tmp_m, tmp_n, tol[1], tol[3]);
      const int st = termination_status(K, tol, b_norm, c_norm, per_constraint);
      double* rec  = nullptr;
      if (trace && trace_len && *trace_len < trace_cap) {
        rec = trace + (size_t)(*trace_len) * kTraceWidth;
        rec[0] = (double)k; rec[1] = K.l2_primal; rec[2] = K.l2_dual; rec[3] = K.pobj;
        rec[4] = K.dobj; rec[5] = K.gap; rec[6] = 0.0; rec[7] = omega; rec[8] = fpe;
        rec[9] = (double)k_r;
        ++*trace_len;
      }
      // check_termination (pdlp.cu:854-1180) then check_limits (pdlp.cu:336-494).
      if (k > 1 && st == Optimal) return finish(Optimal);
      if (elapsed() >= time_limit) return finish(TimeLimit);
      if (k >= iteration_limit) return finish(IterationLimit);
      scale_primal(S, bor, xhat);
      scale_dual(S, bor, yhat);
      scale_slack(S, bor, slack);
      if (is_major) {
        const bool do_restart = restart_decision(k, k_r, fpe, fpe0, last_trial, H);
        last_trial = fpe;
        if (do_restart) {
          // cupdlpx_restart (pdlp_restart_strategy.cu:867-985)
          double dp = 0.0, dd = 0.0;
          for (int j = 0; j < n; ++j) {
            const double t = x0[j] - 1.0 * xhat[j];
            dp += t * t;
          }
          for (int i = 0; i < m; ++i) {
            const double t = y0[i] - 1.0 * yhat[i];
            dd += t * t;
          }
          pid_update(dp, dd, K.l2_dual / (1.0 + c_norm), K.l2_primal / (1.0 + b_norm), eta, H,
                     pid, omega, best_omega, tau, sigma);
          x0 = xhat;
          x  = xhat;
          y0 = yhat;
          y  = yhat;
          k_r        = 0;
          last_trial = kInf;
          restarted  = true;
          ++restarts;
        }
      }
      if (rec) {
        rec[6] = restarted ? 1.0 : 0.0;
        rec[7] = omega;
      }
    }
    // take_step -> compute_next_primal_dual_solution_reflected (pdhg.cu:785-1030)
    const int64_t k2         = k + 2;
    const bool should_major  = ((k + 1) % H.major_iteration == 0) || (k2 % conditional_major(k2)) == 0;
    spmv(Q.AT, y.data(), aty.data());
    for (int j = 0; j < n; ++j) {
      // primal_reflected_(major_)projection (pdhg.cu:431-475)
      const double next    = x[j] - tau * (Q.c[j] - aty[j]);
      const double clamped = std::max(std::min(next, Q.ub[j]), Q.lb[j]);
      xbar[j]              = 2.0 * clamped - x[j];
      if (should_major) {
        xhat[j]  = clamped;
        slack[j] = (clamped - next) / tau;
      }
    }
    spmv(Q.A, xbar.data(), ax.data());
    for (int i = 0; i < m; ++i) {
      // dual_reflected_(major_)projection (pdhg.cu:488-532)
      const double t    = y[i] / sigma - ax[i];
      const double proj = std::max(-Q.uc[i], std::min(t, -Q.lc[i]));
      const double nd   = (t - proj) * sigma;
      ybar[i]           = 2.0 * nd - y[i];
      if (should_major) yhat[i] = nd;
    }
    if (((k + 1) % H.major_iteration) == 0 || restarted) {
      fpe = fixed_point_error(Q.AT, x, xbar, y, ybar, aty, omega, eta, tmp_n);
      if (restarted) fpe0 = fpe;
    }
    // halpern_update (pdlp.cu:2505-2563) with reflection coefficient 1.
    const double w = double(k_r + 1) / double(k_r + 2);
    for (int j = 0; j < n; ++j) x[j] = w * xbar[j] + (1.0 - w) * x0[j];
    for (int i = 0; i < m; ++i) y[i] = w * ybar[i] + (1.0 - w) * y0[i];
    ++k;
    ++k_r;
  }
  return status;
}

}  // extern "C"
