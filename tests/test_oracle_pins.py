"""Hand-worked checks of the oracle's restart, primal-weight, fixed-point-error and
bound/objective-rescaling functions, and proof that each check catches the
plausible mistakes in them.

The expected values below are worked by hand from the reference formulas:
- PID primal weight: cupdlpx_new_primal_weight_computation,
  cpp/src/pdlp/restart_strategy/pdlp_restart_strategy.cu:757-835, Stable3 gains
  k_p = 0.99, k_i = 0.01, k_d = 0, i_smooth = 0.3 (cpp/src/pdlp/solve.cu:167-207).
- Restart decision: should_cupdlpx_restart (pdlp_restart_strategy.cu:688-755) and
  should_do_artificial_restart (:1397-1419); thresholds 0.2 / 0.8 / 0.36, first
  major at 200.
- Fixed-point error: fixed_error_computation (cpp/src/pdlp/pdlp.cu:1401-1420) over
  the interaction (A^T y_bar - A^T y) . (x_bar - x) of compute_interaction_and_movement
  (cpp/src/pdlp/step_size_strategy/adaptive_step_size_strategy.cu:351-517).  It is the
  norm of (dx, dy) in the metric [[w I, eta A^T], [eta A, I / w]], checked against
  that explicit dense matrix.
- beta / gamma: bound_objective_rescaling
  (cpp/src/pdlp/initial_scaling_strategy/initial_scaling.cu:94-149) on the scaled
  constraint bounds and objective of a diagonal matrix whose Ruiz scaling is exact.

The mutation tests compile the oracle with one planted mistake each (a swapped gain,
a dropped smoothing factor, a wrong threshold, a dropped interaction term, l = u rows
counted twice, ...) and assert that the matching check fails on that build.
"""

import math
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import ctypes
import numpy as np
import pytest

from cuopt_amd import generators as G
from oracle import oracle as O

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "pdlp_oracle.cpp")
E = math.e
INF = float("inf")


# ---------------------------------------------------------------- PID primal weight

def check_pid(L=None):
    st = dict(omega=1.0, best_omega=1.0, error_sum=0.0, last_error=0.0, best_gap=INF, tau=0.0, sigma=0.0)
    eta = 0.5
    # Step 1: |x^ - x0| = 1, |y^ - y0| = e, relative residuals 1e-2 (dual) / 1e-4 (primal).
    #   error = ln e - ln 1 - ln 1 = 1;  S = 0.3 * 0 + 1 = 1;  w = exp(0.99 + 0.01) * 1 = e.
    #   ratio = 100, |log10| = 2 < inf: best weight e.
    st = O.pid_update(st, 1.0, E * E, 1e-2, 1e-4, eta, L)
    ok = [
        math.isclose(st["omega"], E, rel_tol=1e-14),
        math.isclose(st["error_sum"], 1.0, rel_tol=1e-14),
        math.isclose(st["last_error"], 1.0, rel_tol=1e-14),
        math.isclose(st["best_omega"], E, rel_tol=1e-14),
        math.isclose(st["best_gap"], 2.0, rel_tol=1e-14),
        math.isclose(st["tau"], eta / E, rel_tol=1e-14),
        math.isclose(st["sigma"], eta * E, rel_tol=1e-14),
    ]
    # Step 2: distances 2 and 2 e^3.  error = ln(2 e^3) - ln 2 - ln e = 2;
    #   S = 0.3 * 1 + 2 = 2.3;  w = exp(0.99 * 2 + 0.01 * 2.3) * e = exp(3.003).
    #   ratio = 1e-3 / 1 -> |log10| = 3 > 2: the best weight stays e.
    st = O.pid_update(st, 4.0, 4.0 * E ** 6, 1e-3, 1.0, eta, L)
    ok += [
        math.isclose(st["omega"], math.exp(3.003), rel_tol=1e-13),
        math.isclose(st["error_sum"], 2.3, rel_tol=1e-14),
        math.isclose(st["last_error"], 2.0, rel_tol=1e-14),
        math.isclose(st["best_omega"], E, rel_tol=1e-14),
        math.isclose(st["tau"], eta / math.exp(3.003), rel_tol=1e-13),
    ]
    # Step 3: a zero primal distance fails the 1e-16 guard: fall back to the best weight e,
    #   reset S and the last error.  ratio = 1 -> |log10| = 0 < 2: best stays e, gap 0.
    st = O.pid_update(st, 0.0, 1.0, 1.0, 1.0, eta, L)
    ok += [
        math.isclose(st["omega"], E, rel_tol=1e-14),
        st["error_sum"] == 0.0 and st["last_error"] == 0.0,
        math.isclose(st["tau"], eta / E, rel_tol=1e-14),
        st["best_gap"] == 0.0,
    ]
    # Step 4: distances 1 and e^2 from the reset state: error = 2 - 0 - 1 = 1;
    #   S = 0.3 * 0 + 1 = 1;  w = exp(1) * e = e^2.  ratio 10: best stays e.
    st = O.pid_update(st, 1.0, E ** 4, 1e-1, 1e-2, eta, L)
    ok += [
        math.isclose(st["omega"], E * E, rel_tol=1e-14),
        math.isclose(st["error_sum"], 1.0, rel_tol=1e-14),
        math.isclose(st["best_omega"], E, rel_tol=1e-14),
    ]
    # Step 5: a zero relative primal residual makes the ratio infinite: fall back to e,
    #   and |log10 inf| = inf never becomes the best gap.
    st = O.pid_update(st, 1.0, 1.0, 1.0, 0.0, eta, L)
    ok += [math.isclose(st["omega"], E, rel_tol=1e-14), st["best_gap"] == 0.0]
    # Step 6: a ratio of exactly 1e8 is outside the open range (1e-8, 1e8): fall back.
    st2 = dict(st, omega=3.0, best_omega=2.0)
    st2 = O.pid_update(st2, 1.0, 1.0, 1.0, 1e-8, eta, L)
    ok += [st2["omega"] == 2.0]
    return all(ok), ok


def test_pid_update_hand_worked_sequence():
    good, flags = check_pid()
    assert good, flags


# ---------------------------------------------------------------- restart decision

# (k, k_r, fpe, fpe0, last_trial, restart?, why)
RESTART_TABLE = [
    (0, 0, float("nan"), float("nan"), INF, False, "no decision before the first major"),
    (200, 200, 1.0, float("nan"), INF, True, "forced at the first major"),
    (200, 200, 5.0, 1.0, 0.1, True, "forced even without any reduction"),
    # In a solve k_r = k = 200 at the first major, where the artificial rule also holds;
    # k_r = 50 pins the forced rule on its own.
    (200, 50, 5.0, 1.0, 0.1, True, "forced whatever k_r"),
    (400, 100, 0.2, 1.0, 0.1, True, "sufficient: fpe <= 0.2 fpe0 (boundary)"),
    (400, 100, 0.15, 1.0, 0.2, True, "sufficient although the error is still falling"),
    (400, 100, 0.22, 1.0, 0.3, False, "0.22 > 0.2 fpe0 and falling: no restart"),
    (600, 200, 0.5, 1.0, 0.4, True, "necessary: <= 0.8 fpe0 and rising"),
    (600, 200, 0.8, 1.0, 0.7, True, "necessary at the 0.8 boundary"),
    (600, 200, 0.85, 1.0, 0.7, False, "0.85 > 0.8 fpe0: not necessary"),
    (600, 200, 0.5, 1.0, 0.5, False, "necessary needs a strict rise over the last trial"),
    (600, 200, 0.5, 1.0, 0.6, False, "falling error: no necessary restart"),
    (1000, 360, 0.9, 1.0, 0.5, True, "artificial: k_r >= 0.36 k (boundary 360)"),
    (1000, 359, 0.9, 1.0, 0.5, False, "k_r = 359 < 360"),
    (1000, 355, 0.9, 1.0, 0.5, False, "k_r = 355 < 360"),
    (2000, 720, 0.9, 1.0, 0.95, True, "artificial at 720 of 2000"),
    (2000, 719, 0.9, 1.0, 0.95, False, "k_r = 719 < 720"),
    (400, 400, 0.9, 1.0, 0.95, True, "artificial: no restart since the start"),
]


def check_restart(L=None):
    got = [O.restart_decision(k, kr, f, f0, lt, L) for (k, kr, f, f0, lt, _, _) in RESTART_TABLE]
    want = [row[5] for row in RESTART_TABLE]
    return got == want, [(row[6], g) for row, g in zip(RESTART_TABLE, got) if g != row[5]]


def test_restart_decision_table():
    good, wrong = check_restart()
    assert good, wrong


# ---------------------------------------------------------------- fixed-point error

A_DENSE = np.array([[1.0, 0.0, -2.0, 0.5],
                    [0.0, 3.0, 0.0, -1.0],
                    [2.0, -1.0, 0.0, 0.0]])


def _csr(D):
    ptr, idx, val = [0], [], []
    for row in D:
        nz = np.nonzero(row)[0]
        idx += nz.tolist()
        val += row[nz].tolist()
        ptr.append(len(idx))
    return np.array(ptr, np.int32), np.array(idx, np.int32), np.array(val)


def check_fpe(L=None):
    ok = []
    # 1 x 1 by hand: A = [2], dx = 1, dy = 1, w = 1, eta = 1/4:
    #   fpe^2 = 1 * 1 + 1 / 1 + 2 * (1 * 2 * 1) * 1/4 = 3.
    f = O.fpe(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.0]), 1,
              [0.0], [1.0], [0.0], [1.0], 1.0, 0.25, L)
    ok.append(math.isclose(f, math.sqrt(3.0), rel_tol=1e-15))
    # 1 x 1 with w = 4: fpe^2 = 4 * 1 + 1 / 4 + 1 = 5.25 (w weights the primal movement).
    f = O.fpe(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.0]), 1,
              [0.0], [1.0], [0.0], [1.0], 4.0, 0.25, L)
    ok.append(math.isclose(f, math.sqrt(5.25), rel_tol=1e-15))
    # 3 x 4: the metric norm of z = (dx, dy) under M = [[w I, eta A^T], [eta A, I / w]].
    rng = np.random.default_rng(3)
    ptr, idx, val = _csr(A_DENSE)
    for w, eta in [(1.7, 0.3), (0.2, 0.25), (5.0, 0.1)]:
        x, xb = rng.standard_normal(4), rng.standard_normal(4)
        y, yb = rng.standard_normal(3), rng.standard_normal(3)
        M = np.block([[w * np.eye(4), eta * A_DENSE.T], [eta * A_DENSE, np.eye(3) / w]])
        z = np.concatenate([xb - x, yb - y])
        want = math.sqrt(float(z @ M @ z))
        got = O.fpe(ptr, idx, val, 4, x, xb, y, yb, w, eta, L)
        ok.append(math.isclose(got, want, rel_tol=1e-13))
    return all(ok), ok


def test_fixed_point_error_is_the_pdhg_metric_norm():
    good, flags = check_fpe()
    assert good, flags


def test_fpe_metric_is_positive_definite_at_the_solver_step():
    # With eta * sigma_max(A) < 1 (the solver takes eta = 0.998 / sigma_max) the metric is
    # positive definite, so the square root of the fixed-point error is always real.
    s = np.linalg.svd(A_DENSE, compute_uv=False)[0]
    eta = 0.998 / s
    for w in (0.1, 1.0, 10.0):
        M = np.block([[w * np.eye(4), eta * A_DENSE.T], [eta * A_DENSE, np.eye(3) / w]])
        assert np.linalg.eigvalsh(M).min() > 0


# ---------------------------------------------------------------- beta / gamma

def _diag_lp():
    # A = diag(4, 16, 1/4, 1, 4, 1): one Ruiz sweep gives D_r = D_c = a^(-1/2) =
    # (1/2, 1/4, 2, 1, 1/2, 1) exactly; later sweeps and Pock-Chambolle keep it.
    a = np.array([4.0, 16.0, 0.25, 1.0, 4.0, 1.0])
    n = a.size
    # Rows: E b = 6, L u = 8, G l = 1.5, ranged [-1, 1], L u = 2, free.
    lo = np.array([6.0, -INF, 1.5, -1.0, -INF, -INF])
    up = np.array([6.0, 8.0, INF, 1.0, 2.0, INF])
    c = np.array([2.0, 8.0, 1.0, 0.0, 0.0, 0.0])
    return G.LP(m=n, n=n, row_offsets=np.arange(n + 1, dtype=np.int32), col_indices=np.arange(n, dtype=np.int32),
                values=a, c=c, con_lb=lo, con_ub=up, var_lb=np.zeros(n), var_ub=np.full(n, 10.0))


def check_beta_gamma(L=None):
    s = O.setup(_diag_lp(), L=L)
    d = np.array([0.5, 0.25, 2.0, 1.0, 0.5, 1.0])
    # Scaled constraint bounds: E [3, 3] -> 9 (l = u counts once), L (-inf, 2] -> 4,
    # G [3, inf) -> 9, ranged [-1, 1] -> 1 + 1, L (-inf, 1] -> 1, free -> 0.
    # Sum 25, beta = 1 / (5 + 1).  Variable bounds do not enter.
    # Scaled objective c D_c = (1, 2, 2, 0, 0, 0): norm 3, gamma = 1 / (3 + 1).
    ok = [
        np.array_equal(s["Dr"], d),
        np.array_equal(s["Dc"], d),
        math.isclose(s["beta"], 1.0 / 6.0, rel_tol=1e-15),
        math.isclose(s["gamma"], 0.25, rel_tol=1e-15),
    ]
    return all(ok), ok


def test_bound_objective_rescaling_hand_worked():
    good, flags = check_beta_gamma()
    assert good, flags


# ---------------------------------------------------------------- per-constraint residuals

def _per_constraint_lp():
    # A = I (3 x 3).  Rows: G with l = 2 (w = 2), L with u = 3 (w = 3), ranged [-1, 4]
    # (w = max(|-1|, |4|) = 4).  c = (1, -1, 2).
    n = 3
    return G.LP(m=n, n=n, row_offsets=np.arange(n + 1, dtype=np.int32), col_indices=np.arange(n, dtype=np.int32),
                values=np.ones(n), c=np.array([1.0, -1.0, 2.0]), con_lb=np.array([2.0, -INF, -1.0]),
                con_ub=np.array([INF, 3.0, 4.0]), var_lb=np.zeros(n), var_ub=np.full(n, 10.0))


def check_per_constraint(L=None):
    """convergence_information.cu:390-470 with relative_residual_t (utils.cuh:576-590):
    max(0, max_i r_i - rel w_i) over the signed residuals, signed c_j on the dual side."""
    lp = _per_constraint_lp()
    rel = 0.1
    y0 = np.zeros(3)

    def k(x, s):
        return O.kkt(lp, np.array(x), y0, np.array(s), rel_primal=rel, rel_dual=rel, L=L)

    ok = []
    # x = (1, 3, 4): Ax - clamp = (-1, 0, 0); r - 0.1 w = (-1.2, -0.3, -0.4): every term is
    # negative, so the maximum is the neutral 0 -- the row below its lower bound passes.
    ok.append(k([1.0, 3.0, 4.0], [0, 0, 0])["linf_primal"] == 0.0)
    # x = (2, 3, 4.5): r = (0, 0, 0.5); r - 0.1 w = (-0.2, -0.3, 0.1): 0.1 (ranged row, w = 4).
    ok.append(math.isclose(k([2.0, 3.0, 4.5], [0, 0, 0])["linf_primal"], 0.1, rel_tol=1e-12))
    # x = (1, 5, 4.5): r = (-1, 2, 0.5) -> (-1.2, 1.7, 0.1): 1.7.
    ok.append(math.isclose(k([1.0, 5.0, 4.5], [0, 0, 0])["linf_primal"], 1.7, rel_tol=1e-12))
    # Dual, y = 0: r_d = c - s.  s = (1, 0, 2): r_d = (0, -1, 0); r_d - 0.1 c = (-0.1, -0.9, -0.2): 0.
    ok.append(k([2.0, 3.0, 4.0], [1.0, 0.0, 2.0])["linf_dual"] == 0.0)
    # s = (1, -1.05, 2): r_d = (0, 0.05, 0); r_d - 0.1 c = (-0.1, 0.15, -0.2): 0.15 (signed c = -1).
    ok.append(math.isclose(k([2.0, 3.0, 4.0], [1.0, -1.05, 2.0])["linf_dual"], 0.15, rel_tol=1e-12))
    return all(ok), ok


def test_per_constraint_residual_hand_worked():
    good, flags = check_per_constraint()
    assert good, flags


def _signed_maxima(lp, x, y, s, tol):
    import scipy.sparse as sp

    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    ax = A @ x
    r = ax - np.clip(ax, lp.con_lb, lp.con_ub)
    w = np.maximum(np.where(np.isfinite(lp.con_ub), np.abs(lp.con_ub), 0.0),
                   np.where(np.isfinite(lp.con_lb), np.abs(lp.con_lb), 0.0))
    rd = lp.c - A.T @ y - s
    return max(0.0, np.max(r - tol * w)), max(0.0, np.max(rd - tol * lp.c))


def test_per_constraint_solve_stops_at_the_first_check_passing_the_signed_rule():
    # termination_strategy.cu:317-332: with per_constraint_residual the solve stops at the
    # first check whose signed L-inf maxima (and the gap test) pass.  Transportation costs
    # are positive, so the dual side -rel c_j loosens every column's test.
    lp = G.transportation_lp(sources=200, sinks=500, arcs=50, seed=3)
    tol = 1e-4
    per = O.solve(lp, tol=tol, per_constraint_residual=True, iteration_limit=100000, trace_cap=100000)
    l2 = O.solve(lp, tol=tol, iteration_limit=100000)
    assert per["status"] == l2["status"] == 1
    p_inf, d_inf = _signed_maxima(lp, per["x"], per["y"], per["reduced_cost"], tol)
    assert p_inf <= tol and d_inf <= tol
    assert per["iterations"] != l2["iterations"]


def test_per_constraint_signed_rule_cannot_pass_a_cost_below_minus_abs_over_rel():
    # With the signed c_j of relative_residual_t a column with c_j < -abs / rel needs
    # r_j <= abs + rel c_j < 0 ... while r_j -> 0 at the optimum: the dual test cannot
    # pass, so such an LP runs to its limit even though the l2 solve converges.  This is
    # the reference's rule, kept for result identity.
    lp = G.random_lp(m=200, n=500, seed=3)
    tol = 1e-4
    assert lp.c.min() < -1.0
    l2 = O.solve(lp, tol=tol, iteration_limit=5000)
    per = O.solve(lp, tol=tol, per_constraint_residual=True, iteration_limit=3000)
    assert l2["status"] == 1 and l2["iterations"] < 3000
    assert per["status"] == 4 and per["iterations"] == 3000
    _, d_inf = _signed_maxima(lp, per["x"], per["y"], per["reduced_cost"], tol)
    assert d_inf >= tol * (-lp.c.min()) - 1e-6


# ---------------------------------------------------------------- planted mistakes

# (name, check, original text, mutated text)
MUTANTS = [
    ("pid_swapped_gains", check_pid,
     "std::exp(H.k_p * err + H.k_i * s + H.k_d * de)", "std::exp(H.k_i * err + H.k_p * s + H.k_d * de)"),
    ("pid_dropped_i_smooth", check_pid, "pid.err_sum * H.i_smooth + err", "pid.err_sum + err"),
    ("pid_log10_error", check_pid,
     "std::log(l2d) - std::log(l2p) - std::log(cur)", "std::log10(l2d) - std::log10(l2p) - std::log10(cur)"),
    ("pid_no_reset_on_fallback", check_pid, "    pid.err_sum  = 0.0;\n", "\n"),
    ("pid_best_weight_not_tracked", check_pid, "    best_omega   = omega;\n", "\n"),
    ("pid_wrong_distance_guard", check_pid, "l2p > 1e-16 && l2d > 1e-16", "l2p >= 0.0 && l2d >= 0.0"),
    ("restart_sufficient_0.25", check_restart,
     "double sufficient_reduction   = 0.2;", "double sufficient_reduction   = 0.25;"),
    ("restart_necessary_0.9", check_restart,
     "double necessary_reduction    = 0.8;", "double necessary_reduction    = 0.9;"),
    ("restart_artificial_0.35", check_restart,
     "double artificial_threshold   = 0.36;", "double artificial_threshold   = 0.35;"),
    ("restart_artificial_0.37", check_restart,
     "double artificial_threshold   = 0.36;", "double artificial_threshold   = 0.37;"),
    ("restart_necessary_without_rise", check_restart,
     "fpe <= H.necessary_reduction * fpe0 && fpe > last_trial", "fpe <= H.necessary_reduction * fpe0"),
    ("restart_no_forced_first", check_restart, "  if (k == H.major_iteration) return true;", ""),
    ("restart_artificial_swapped", check_restart,
     "(double)k_r >= H.artificial_threshold * (double)k", "(double)k >= H.artificial_threshold * (double)k_r"),
    ("fpe_dropped_interaction", check_fpe, "std::sqrt(movement + ci)", "std::sqrt(movement)"),
    ("fpe_interaction_factor_1", check_fpe, "2.0 * inter * eta", "1.0 * inter * eta"),
    ("fpe_weight_swapped", check_fpe, "nx * omega + ny / omega", "nx / omega + ny * omega"),
    ("fpe_interaction_sign", check_fpe, "const double t  = tmp_n[j] - aty[j];", "const double t  = aty[j] - tmp_n[j];"),
    ("beta_counts_fixed_rows_twice", check_beta_gamma,
     "if (std::isfinite(lo[i]) && lo[i] != up[i]) t += lo[i] * lo[i];",
     "if (std::isfinite(lo[i])) t += lo[i] * lo[i];"),
    ("beta_unscaled_bounds", check_beta_gamma,
     "S.beta  = 1.0 / (std::sqrt(sum_bounds_sq(Q.lc, Q.uc)) + 1.0);",
     "S.beta  = 1.0 / (std::sqrt(sum_bounds_sq(P.lc, P.uc)) + 1.0);"),
    ("gamma_unscaled_objective", check_beta_gamma,
     "cs += Q.c[j] * Q.c[j] * 1.0;", "cs += P.c[j] * P.c[j] * 1.0;"),
    ("per_constraint_abs_primal", check_per_constraint,
     "linf_p = std::max(linf_p, r - rel_primal", "linf_p = std::max(linf_p, std::fabs(r) - rel_primal"),
    ("per_constraint_abs_dual", check_per_constraint,
     "linf_d = std::max(linf_d, r - rel_dual * P.c[j]);", "linf_d = std::max(linf_d, std::fabs(r) - rel_dual * P.c[j]);"),
    ("per_constraint_abs_c", check_per_constraint,
     "linf_d = std::max(linf_d, r - rel_dual * P.c[j]);", "linf_d = std::max(linf_d, r - rel_dual * std::fabs(P.c[j]));"),
    ("per_constraint_neutral_minus_inf", check_per_constraint,
     "double rp2 = 0.0, psum = 0.0, linf_p = 0.0;", "double rp2 = 0.0, psum = 0.0, linf_p = -kInf;"),
    ("per_constraint_lower_bound_only", check_per_constraint,
     "  if (std::isfinite(lo)) v = std::max(v, std::fabs(lo));", "  if (std::isfinite(lo)) v = std::fabs(lo);"),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    src = open(ORACLE_SRC).read()
    root = tmp_path_factory.mktemp("oracle_mutants")
    jobs = []
    for name, _, old, new in MUTANTS:
        assert src.count(old) == 1, f"mutation target of {name} not found exactly once"
        path = root / f"{name}.cpp"
        path.write_text(src.replace(old, new))
        jobs.append((name, path, root / f"lib{name}.so"))

    def compile_one(job):
        name, path, so = job
        r = subprocess.run([cxx, "-O1", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                            "-o", str(so), str(path)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        return name, str(so)

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        built = dict(ex.map(compile_one, jobs))
    return {name: ctypes.CDLL(so) for name, so in built.items()}


def test_unmutated_source_builds_and_passes(tmp_path):
    # The mutants are compiled from this source with one edit each: the unedited
    # source compiled the same way must pass every check.
    so = tmp_path / "libclean.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-o", str(so),
                    ORACLE_SRC], check=True)
    L = ctypes.CDLL(str(so))
    for check in (check_pid, check_restart, check_fpe, check_beta_gamma, check_per_constraint):
        assert check(L)[0], check.__name__


@pytest.mark.parametrize("name,check", [(m[0], m[1]) for m in MUTANTS], ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(mutant_libs, name, check):
    good, _ = check(mutant_libs[name])
    assert not good, f"the planted mistake {name} passes {check.__name__}"
