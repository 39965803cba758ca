"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle,
closed-form optima, committed HiGHS optima and reference fixtures.

Tolerances, stated per test:
* FP64 SpMV: |y_gpu - y_ref| <= 64 eps64 * sum_j |a_ij x_j|  (summation order differs)
* FP32 compensated SpMV vs FP64 products of the same FP32-rounded inputs:
  |dy| <= eps32 |y| + 64 eps32^2 sum_j |a_ij x_j|
* FP64 solve vs oracle: same status and iteration count, objective within
  1e-9 relative, iterates within 1e-7 (the same arithmetic, different sum order)
* FP32-mixed solve at tol: objective within 10 tol of the closed-form optimum
"""

import json
import os

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd import generators as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TL = 120.0  # seconds: every solve carries a time limit
EPS64 = np.finfo(np.float64).eps
EPS32 = float(np.finfo(np.float32).eps)


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def abs_products(lp, x):
    import scipy.sparse as sp

    A = sp.csr_matrix((np.abs(lp.values), lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    return A @ np.abs(x)


@pytest.fixture(scope="module")
def highs():
    with open(os.path.join(GOLDEN, "highs_optima.json")) as f:
        return json.load(f)["cases"]


def _spmv_cases():
    rng = np.random.default_rng(0)
    cases = [G.random_lp(m=500, n=1200, seed=1), G.skewed_lp(m=20000, n=40000, seed=2)]
    # edge shapes: empty rows, a row of exactly one tile, one just over, a very long row
    lens = np.array([0, 0, 2048, 2049, 1, 0, 10000, 3, 0, 4096, 0])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    n = 12000
    idx = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens if l] + []).astype(np.int32)
    val = rng.standard_normal(idx.size)
    edge = G.LP(m=len(lens), n=n, row_offsets=ptr, col_indices=idx, values=val, c=np.zeros(n),
                con_lb=np.zeros(len(lens)), con_ub=np.zeros(len(lens)), var_lb=np.zeros(n), var_ub=np.ones(n))
    cases.append(edge)
    return cases


@pytest.mark.parametrize("case", range(3))
def test_spmv_fp64_matches_oracle(case):
    lp = _spmv_cases()[case]
    x = np.random.default_rng(case).standard_normal(lp.n)
    y = c_api.spmv(lp.row_offsets, lp.col_indices, lp.values, x, lp.n, "fp64")
    ref = O.spmv(lp.row_offsets, lp.col_indices, lp.values, x, lp.n)
    bound = 64 * EPS64 * abs_products(lp, x) + 1e-300
    assert np.all(np.abs(y - ref) <= bound)
    # transpose too (the A^T operand of the primal step)
    tp, ti, tv = O.transpose(lp.row_offsets, lp.col_indices, lp.values, lp.n)
    yv = np.random.default_rng(9).standard_normal(lp.m)
    yt = c_api.spmv(tp, ti, tv, yv, lp.m, "fp64")
    reft = O.spmv(tp, ti, tv, yv, lp.m)
    assert np.allclose(yt, reft, rtol=0, atol=np.max(np.abs(reft)) * 1e-12 + 1e-300)


@pytest.mark.parametrize("case", range(3))
def test_spmv_fp32_compensated(case):
    lp = _spmv_cases()[case]
    x = np.random.default_rng(case + 5).standard_normal(lp.n)
    y = c_api.spmv(lp.row_offsets, lp.col_indices, lp.values, x, lp.n, "fp32_mixed")
    v32 = lp.values.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    ref = O.spmv(lp.row_offsets, lp.col_indices, v32, x32, lp.n)  # exact-ish sums of the rounded inputs
    absp = abs_products(lp, x32)
    # compensated sum: one final FP32 rounding of the row value plus O(n eps^2) terms
    bound = EPS32 * np.abs(ref) + 64 * EPS32 * EPS32 * absp + 1e-30
    assert np.all(np.abs(y - ref) <= bound)
    # a plain left-to-right FP32 sum of the same products is far worse on long rows
    ptr = lp.row_offsets
    longest = int(np.argmax(np.diff(ptr)))
    if ptr[longest + 1] - ptr[longest] > 2048:
        seg = slice(ptr[longest], ptr[longest + 1])
        prods = (lp.values[seg].astype(np.float32) * x.astype(np.float32)[lp.col_indices[seg]]).astype(np.float32)
        naive = np.float32(0)
        for p in prods:
            naive = np.float32(naive + p)
        assert abs(y[longest] - ref[longest]) * 8 < abs(float(naive) - ref[longest]) + EPS32 * abs(ref[longest])


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_spmv_full_size_skewed(precision):
    # The skewed LP at its full size (m = 500 000, n = 1 000 000, 4.75 M nonzeros,
    # rows of up to about 474 000 nonzeros split over many tiles), with the
    # tolerances of the cases above.
    lp = G.skewed_lp()
    assert lp.m == 500_000 and lp.n == 1_000_000
    x = np.random.default_rng(21).standard_normal(lp.n)
    y = c_api.spmv(lp.row_offsets, lp.col_indices, lp.values, x, lp.n, precision)
    if precision == "fp64":
        ref = O.spmv(lp.row_offsets, lp.col_indices, lp.values, x, lp.n)
        bound = 64 * EPS64 * abs_products(lp, x) + 1e-300
    else:
        v32 = lp.values.astype(np.float32).astype(np.float64)
        x32 = x.astype(np.float32).astype(np.float64)
        ref = O.spmv(lp.row_offsets, lp.col_indices, v32, x32, lp.n)
        bound = EPS32 * np.abs(ref) + 64 * EPS32 * EPS32 * abs_products(lp, x32) + 1e-30
    worst = int(np.argmax(np.abs(y - ref) - bound))
    assert np.all(np.abs(y - ref) <= bound), (worst, y[worst], ref[worst], bound[worst])


@pytest.mark.parametrize("m,n", [(400, 1000), (20000, 50000)])
def test_setup_matches_oracle(m, n):
    # The FP64 power iteration ends each step in one workgroup up to 16 384 rows and
    # with three reductions above: both forms against the oracle's iteration count.
    lp = G.random_lp(m=m, n=n, seed=3)
    r = c_api.solve(lp, tol=1e-4, iteration_limit=0, time_limit=TL)
    s = O.setup(lp)
    assert r.status == 4
    assert rel(r.info["beta"], s["beta"]) < 1e-14
    assert rel(r.info["gamma"], s["gamma"]) < 1e-14
    assert r.info["power_iterations"] == s["power_iterations"]
    assert abs(r.info["sigma_max_sq"] - s["sigma_max_sq"]) < 1e-10 * s["sigma_max_sq"]
    assert abs(r.info["eta0"] - s["eta"]) < 1e-10 * s["eta"]


def test_fp32_step_size_by_lanczos_matches_svd():
    # FP32-mixed solves estimate sigma_max(A~)^2 by Lanczos on the FP64 scaled
    # matrix (the oracle's scaled values here): checked against a dense SVD, and
    # never below the reference power iteration's estimate
    lp = G.random_lp(m=400, n=1000, seed=3)
    r = c_api.solve(lp, tol=1e-4, iteration_limit=0, time_limit=TL, precision="fp32_mixed")
    s = O.setup(lp)
    import scipy.sparse as sp

    At = sp.csr_matrix((s["scaled_values"], lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
    smax2 = np.linalg.svd(At, compute_uv=False)[0] ** 2
    assert abs(r.info["sigma_max_sq"] - smax2) <= 1e-5 * smax2
    assert r.info["sigma_max_sq"] >= s["sigma_max_sq"] * (1 - 1e-12)
    assert r.info["power_iterations"] < 100
    assert abs(r.info["eta0"] - 0.998 / np.sqrt(smax2)) <= 1e-5 * r.info["eta0"]


def test_device_transpose_unsorted_rows_and_empty_columns():
    # rows given out of column order and columns without nonzeros: FP64 solve
    # still matches the oracle (host row sort, device transpose)
    lp = G.random_lp(m=200, n=500, seed=12)
    n2 = lp.n + 3  # three empty columns at the end
    perm_rows = []
    idx, val = lp.col_indices.copy(), lp.values.copy()
    rng = np.random.default_rng(0)
    for i in range(lp.m):
        b, e = lp.row_offsets[i], lp.row_offsets[i + 1]
        p = rng.permutation(e - b)
        idx[b:e], val[b:e] = idx[b:e][p], val[b:e][p]
    import dataclasses

    lp2 = dataclasses.replace(lp, n=n2, col_indices=idx, values=val, c=np.concatenate([lp.c, np.ones(3)]),
                              var_lb=np.concatenate([lp.var_lb, np.zeros(3)]),
                              var_ub=np.concatenate([lp.var_ub, np.full(3, 10.0)]))
    r = c_api.solve(lp2, tol=1e-6, time_limit=TL, iteration_limit=100000)
    o = O.solve(lp2, tol=1e-6)
    assert r.status == 1 and r.iterations == o["iterations"]
    assert rel(r.primal_objective, o["primal_objective"]) < 1e-9


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_solve_fp64_matches_oracle(tol):
    lp = G.random_lp(m=500, n=1250, seed=7)
    r = c_api.solve(lp, tol=tol, precision="fp64", time_limit=TL, iteration_limit=100000)
    o = O.solve(lp, tol=tol, trace_cap=100000)
    assert r.error == 0 and r.status == 1 and o["status"] == 1
    assert r.iterations == o["iterations"]
    assert rel(r.primal_objective, o["primal_objective"]) < 1e-9
    assert np.max(np.abs(r.x - o["x"])) < 1e-7
    assert np.max(np.abs(r.y - o["y"])) < 1e-7
    assert np.max(np.abs(r.reduced_cost - o["reduced_cost"])) < 1e-7
    assert rel(r.primal_objective, lp.obj_opt) < 50 * tol
    # same restart decisions at the same checks
    tr, ot = r.trace, o["trace"]
    assert tr.shape == ot.shape
    assert np.array_equal(tr[:, 0], ot[:, 0]) and np.array_equal(tr[:, 6], ot[:, 6])
    assert np.allclose(tr[:, 1:3], ot[:, 1:3], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_small_random_lp_2000x5000_fp64_matches_oracle(tol):
    # The small random LP (m = 2 000, n = 5 000, 8 nonzeros per column, 0 <= x <= 10,
    # equality rows, seed 0) in plain FP64: the oracle's iteration count and restart
    # decisions, its objective to 1e-9, and the built-in optimum c.x0 = b.y0.
    lp = G.random_lp(m=2000, n=5000, seed=0)
    print(f"small random LP tol {tol:g}", flush=True)
    r = c_api.solve(lp, tol=tol, precision="fp64", time_limit=TL, iteration_limit=100000)
    o = O.solve(lp, tol=tol, threads=8, trace_cap=100000)
    assert r.error == 0 and r.status == 1 and o["status"] == 1
    assert r.iterations == o["iterations"]
    assert rel(r.primal_objective, o["primal_objective"]) < 1e-9
    assert np.array_equal(r.trace[:, 0], o["trace"][:, 0]) and np.array_equal(r.trace[:, 6], o["trace"][:, 6])
    assert np.max(np.abs(r.x - o["x"])) < 1e-7 and np.max(np.abs(r.y - o["y"])) < 1e-7
    assert rel(r.primal_objective, lp.obj_opt) < 10 * tol
    # FP32-mixed (refined below 1e-4) meets the same test and optimum
    f = c_api.solve(lp, tol=tol, precision="fp32_mixed", time_limit=TL, iteration_limit=100000)
    assert f.status == 1 and rel(f.primal_objective, lp.obj_opt) < 10 * tol


def test_skewed_fp64_matches_oracle_and_closed_form():
    lp = G.skewed_lp(m=20000, n=40000, seed=5)
    r = c_api.solve(lp, tol=1e-6, precision="fp64", time_limit=TL, iteration_limit=200000)
    o = O.solve(lp, tol=1e-6, threads=8)
    assert r.status == 1 and o["status"] == 1
    assert abs(r.iterations - o["iterations"]) <= 200
    assert rel(r.primal_objective, lp.obj_opt) < 1e-4
    assert rel(r.primal_objective, o["primal_objective"]) < 1e-6


@pytest.mark.parametrize("tol", [1e-4])
def test_fp32_mixed_reaches_tolerance(tol):
    lp = G.random_lp(m=500, n=1250, seed=7)
    r = c_api.solve(lp, tol=tol, precision="fp32_mixed", time_limit=TL, iteration_limit=200000)
    assert r.error == 0 and r.status == 1
    assert r.info["precision"] == 1
    assert r.info["l2_primal_residual"] <= tol + tol * O.setup(lp)["b_norm"]
    assert rel(r.primal_objective, lp.obj_opt) < 10 * tol


def test_fp32_mixed_skewed():
    lp = G.skewed_lp(m=20000, n=40000, seed=5)
    r = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=200000)
    assert r.status == 1
    assert rel(r.primal_objective, lp.obj_opt) < 1e-3


@pytest.mark.parametrize("fname", ["afiro_original.mps", "good-max.mps", "max_offset.mps",
                                   "lp_model_with_var_bounds.mps"])
def test_fixtures_fp64(highs, fname):
    # the problem handle cuOptReadProblem returns is the one solved (the drop-in read path)
    prob = c_api.Problem.read_mps(os.path.join(GOLDEN, fname))
    lp = prob.lp
    r = c_api.solve(prob, tol=1e-8, time_limit=TL, iteration_limit=1000000)
    o = O.solve(lp, tol=1e-8)
    assert r.status == 1
    assert r.iterations == o["iterations"]
    assert rel(r.primal_objective, highs[fname]["objective"]) < 1e-6
    assert rel(r.dual_objective, o["dual_objective"]) < 1e-8
    assert np.allclose(r.y, o["y"], atol=1e-7)


def test_reference_ranged_problem():
    # c_api_tests.cpp:105-112 / c_api_test.c:927-1040: max 5x + 8y s.t. 2x + 3y <= 12,
    # 3x + y <= 6, 2 <= x + 2y <= 8, 0 <= x, y <= 10; optimum 32 at (0, 4)
    lp = G.LP(m=3, n=2, row_offsets=np.array([0, 2, 4, 6], np.int32), col_indices=np.array([0, 1] * 3, np.int32),
              values=np.array([2.0, 3.0, 3.0, 1.0, 1.0, 2.0]), c=np.array([5.0, 8.0]),
              con_lb=np.array([-np.inf, -np.inf, 2.0]), con_ub=np.array([12.0, 6.0, 8.0]),
              var_lb=np.zeros(2), var_ub=np.full(2, 10.0), maximize=True, name="ranged")
    r = c_api.solve(lp, time_limit=TL, method=c_api.CUOPT_METHOD_CONCURRENT)
    assert r.status == 1
    assert abs(r.primal_objective - 32.0) < 1e-3


@pytest.mark.parametrize("method", ["CONCURRENT", "PDLP"])
def test_reference_maximize_problem_dual_variables(method):
    # c_api_tests.cpp:149-166 / c_api_test.c:1461-1580: max 4x1 + x2 + 5x3 + 3x4 s.t. three <= rows,
    # x >= 0.  The reference asks objective == dual objective within 1e-2 for PDLP at the
    # default tolerances; the HiGHS optimum is 29 with row duals (11, 0, 6) for the max problem.
    A = np.array([[1, -1, -1, 3], [5, 1, 3, 8], [-1, 2, 3, -5]], float)
    lp = G.LP(m=3, n=4, row_offsets=np.array([0, 4, 8, 12], np.int32), col_indices=np.tile(np.arange(4, dtype=np.int32), 3),
              values=A.ravel(), c=np.array([4.0, 1.0, 5.0, 3.0]), con_lb=np.full(3, -np.inf),
              con_ub=np.array([1.0, 55.0, 3.0]), var_lb=np.zeros(4), var_ub=np.full(4, np.inf), maximize=True,
              name="maxdual")
    r = c_api.solve(lp, time_limit=TL, method=getattr(c_api, f"CUOPT_METHOD_{method}"))
    assert r.status == 1
    assert abs(r.primal_objective - r.dual_objective) < 1e-2
    assert abs(r.primal_objective - 29.0) < 1e-2
    np.testing.assert_allclose(r.y, [11.0, 0.0, 6.0], atol=5e-2)
    # reduced costs are the dual slack of the iterate: c - A^T y up to the 1e-4 tolerance
    np.testing.assert_allclose(r.reduced_cost, lp.c - A.T @ r.y, atol=1e-3)


def _solve_with(lp, tol, **params):
    prob = c_api.Problem(lp)
    st = c_api.make_settings(tol, "fp64", params.pop("iteration_limit", 200000), TL)
    for name, v in params.items():
        assert st.set(name, v) == 0, name
    return c_api.solve_problem(prob, st)


def _residuals(lp, r):
    import scipy.sparse as sp

    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    ax = A @ r.x
    viol = np.maximum(lp.con_lb - ax, 0.0) + np.maximum(ax - lp.con_ub, 0.0)
    w = np.maximum(np.where(np.isfinite(lp.con_lb), np.abs(lp.con_lb), 0.0),
                   np.where(np.isfinite(lp.con_ub), np.abs(lp.con_ub), 0.0))
    return viol, w


def test_per_constraint_residual_signed_matches_oracle():
    # convergence_information.cu:390-470, utils.cuh:576-590, termination_strategy.cu:317-332:
    # the reference's signed rule, max(0, max_i r_i - rel w_i) <= abs and
    # max(0, max_j r_d,j - rel c_j) <= abs.  FP64 GPU solve against the oracle with the same
    # rule: the same stopping iteration; the returned point passes the rule in numpy.
    import scipy.sparse as sp

    lp = G.transportation_lp(sources=200, sinks=500, arcs=50, seed=3)  # costs > 0
    tol = 1e-4
    per = _solve_with(lp, tol, per_constraint_residual=True)
    o = O.solve(lp, tol=tol, per_constraint_residual=True, iteration_limit=200000)
    plain = _solve_with(lp, tol)
    assert per.status == o["status"] == 1
    assert per.iterations == o["iterations"] != plain.iterations
    assert rel(per.primal_objective, o["primal_objective"]) < 1e-9
    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    viol, w = _residuals(lp, per)
    ax = A @ per.x
    r = ax - np.clip(ax, lp.con_lb, lp.con_ub)
    rd = lp.c - A.T @ per.y - per.reduced_cost
    assert max(0.0, np.max(r - tol * w)) <= tol * (1 + 1e-6)
    assert max(0.0, np.max(rd - tol * lp.c)) <= tol * (1 + 1e-6)


def test_per_constraint_residual_absolute_extension():
    # pdlp_per_constraint_absolute = 1 (an extension): max_i |r_p,i| - rel w_i <= abs and
    # max_j |r_d,j| - rel |c_j| <= abs, with w_i the larger finite bound magnitude of row i.
    # The returned point meets both (checked in numpy); the L2-terminated point of the
    # same LP does not, so the option changes the rule.  (This LP has costs below
    # -abs / rel, under which the signed rule's dual test cannot pass.)
    import scipy.sparse as sp

    lp = G.random_lp(m=3000, n=7500, seed=41)
    tol = 1e-4
    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))

    def linf(r):
        viol, w = _residuals(lp, r)
        rd = lp.c - A.T @ r.y - r.reduced_cost
        return np.max(viol - tol * w, initial=0.0), np.max(np.abs(rd) - tol * np.abs(lp.c), initial=0.0)

    plain = _solve_with(lp, tol)
    per = _solve_with(lp, tol, per_constraint_residual=True, pdlp_per_constraint_absolute=1)
    assert plain.status == per.status == 1
    p_inf, d_inf = linf(per)
    assert p_inf <= tol * (1 + 1e-6) and d_inf <= tol * (1 + 1e-6)
    assert max(linf(plain)) > tol
    assert rel(per.primal_objective, lp.obj_opt) < 1e-4


def test_first_primal_feasible():
    # pdlp.cu:928-983: the first check whose iterate passes the primal test, but not
    # the optimality test, ends the solve with PrimalFeasible.  The primal tolerances
    # are looser than the dual and gap ones here, so the primal test passes first.
    lp = G.random_lp(m=1000, n=2500, seed=42)
    tol, ptol = 1e-8, 1e-4
    loose = dict(absolute_primal_tolerance=ptol, relative_primal_tolerance=ptol)
    full = _solve_with(lp, tol, **loose)
    first = _solve_with(lp, tol, first_primal_feasible=True, **loose)
    assert full.status == 1
    assert first.status == c_api.CUOPT_TERIMINATION_STATUS_PRIMAL_FEASIBLE
    assert first.iterations < full.iterations
    viol, _ = _residuals(lp, first)
    b_norm = np.sqrt(np.sum(np.where(np.isfinite(lp.con_lb) & (lp.con_lb != lp.con_ub), lp.con_lb, 0.0) ** 2)
                     + np.sum(np.where(np.isfinite(lp.con_ub), lp.con_ub, 0.0) ** 2))
    assert np.linalg.norm(viol) <= (ptol + ptol * b_norm) * (1 + 1e-6)
    # the previous check was not primal feasible
    tr = first.trace
    assert tr.shape[0] >= 2 and tr[-2, 1] > ptol + ptol * b_norm


def test_save_best_primal_so_far():
    # pdlp.cu:508-625, 394-401: at the iteration limit the best point met at a check
    # is returned (no check is primal feasible this early: the smallest L2 primal
    # residual wins), with status IterationLimit
    lp = G.skewed_lp(m=20000, n=40000, seed=43)
    last = _solve_with(lp, 1e-8, iteration_limit=1500)
    best = _solve_with(lp, 1e-8, iteration_limit=1500, save_best_primal_so_far=True)
    assert last.status == best.status == c_api.CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT
    np.testing.assert_array_equal(last.trace, best.trace)  # the same run
    l2p = best.trace[:, 1]
    assert best.info["l2_primal_residual"] == l2p.min()
    viol, _ = _residuals(lp, best)
    assert abs(np.linalg.norm(viol) - l2p.min()) <= 1e-6 * (1 + l2p.min())
    assert best.info["l2_primal_residual"] <= last.info["l2_primal_residual"]


def test_transport_matches_highs(highs):
    case = highs["transport_small"]
    lp = G.transportation_lp(**case["kwargs"])
    r = c_api.solve(lp, tol=1e-8, time_limit=TL, iteration_limit=1000000)
    assert r.status == 1
    assert rel(r.primal_objective, case["objective"]) < 1e-6


def primal_violation(lp, x):
    """Independent numpy check: ||Ax - clamp(Ax, l_c, u_c)||_2 and the largest bound violation of x."""
    ax = lp.to_scipy() @ x
    rp = ax - np.clip(ax, lp.con_lb, lp.con_ub)
    bv = np.max(np.maximum(lp.var_lb - x, 0.0) + np.maximum(x - lp.var_ub, 0.0), initial=0.0)
    return float(np.linalg.norm(rp)), float(bv)


def bound_norm(lp):
    l, u = lp.con_lb, lp.con_ub
    t = np.where(np.isfinite(l) & (l != u), l * l, 0.0) + np.where(np.isfinite(u), u * u, 0.0)
    return float(np.sqrt(t.sum()))


# FP32-mixed solves refined in FP64: the FP32 loop stops at 1e-4 and correction
# rounds take the original problem's FP64 KKT test to 1e-8.  Pinned by the HiGHS
# optimum (transportation, <= rows and = rows) and the closed form (skewed LP
# with = and <= rows, whose duals on <= rows exercise the dual kink).
def test_fp32_refinement_transport_1e8(highs):
    case = highs["transport_medium"]
    lp = G.transportation_lp(**case["kwargs"])
    r = c_api.solve(lp, tol=1e-8, precision="fp32_mixed", time_limit=TL, iteration_limit=2_000_000)
    assert r.status == 1, (r.status, r.info)
    assert r.info["refine_rounds"] >= 1
    assert rel(r.primal_objective, case["objective"]) < 1e-7
    rp, bv = primal_violation(lp, r.x)
    assert rp <= 1e-8 * (1.0 + bound_norm(lp)) * 1.01 and bv <= 1e-12


def test_fp32_refinement_skewed_1e8():
    lp = G.skewed_lp(m=20000, n=40000, seed=5)
    r = c_api.solve(lp, tol=1e-8, precision="fp32_mixed", time_limit=TL, iteration_limit=2_000_000)
    assert r.status == 1, (r.status, r.info)
    assert r.info["refine_rounds"] >= 1
    assert rel(r.primal_objective, lp.obj_opt) < 1e-7
    rp, bv = primal_violation(lp, r.x)
    assert rp <= 1e-8 * (1.0 + bound_norm(lp)) * 1.01 and bv <= 1e-12


def test_full_size_skewed_lp_fp32_mixed_and_fp64():
    # the headline LP at its full size (m = 500 000, n = 1 000 000, about 4.75 M
    # nonzeros) through size-independent properties: both precisions reach 1e-4
    # relative KKT, the objective is the built-in optimum to 1e-3, an independent
    # numpy check confirms the primal residual, and the two precisions agree to 1e-6
    lp = G.skewed_lp()
    assert lp.m == 500_000 and lp.n == 1_000_000
    r32 = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=200000)
    r64 = c_api.solve(lp, tol=1e-4, precision="fp64", time_limit=TL, iteration_limit=200000)
    assert r32.status == r64.status == 1
    for r in (r32, r64):
        assert rel(r.primal_objective, lp.obj_opt) < 1e-3
        rp, bv = primal_violation(lp, r.x)
        assert rp <= 1e-4 * (1.0 + bound_norm(lp)) * 1.01 and bv <= 1e-12
    assert abs(r32.primal_objective - r64.primal_objective) <= 1e-6 * abs(r64.primal_objective)


def test_full_size_transport_fp32_refined_to_1e8(highs):
    # the transportation LP at its full size (20 000 x 50 000, n = 1 000 000 arcs,
    # m = 70 000): FP32-mixed with FP64 refinement to 1e-8 against the HiGHS optimum
    case = highs["transport_full"]
    lp = G.transportation_lp(**case["kwargs"])
    assert lp.n == 1_000_000 and lp.m == 70_000
    r = c_api.solve(lp, tol=1e-8, precision="fp32_mixed", time_limit=TL, iteration_limit=2_000_000)
    assert r.status == 1, (r.status, r.info)
    assert r.info["refine_rounds"] >= 1
    assert rel(r.primal_objective, case["objective"]) < 1e-7
    rp, bv = primal_violation(lp, r.x)
    assert rp <= 1e-8 * (1.0 + bound_norm(lp)) * 1.01 and bv <= 1e-12


def test_fp32_refinement_off_is_plain_loop():
    # default FP32-mixed solves recentre at restarts once the relative KKT error
    # is below 1e-2; with refinement off the loop is the plain one and never does
    lp = G.random_lp(m=500, n=1250, seed=7)
    a = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=200000)
    b = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=200000, refinement=False)
    assert a.status == b.status == 1
    assert a.info["refine_rounds"] >= 1 and b.info["refine_rounds"] == 0
    assert rel(a.primal_objective, lp.obj_opt) < 1e-3 and rel(b.primal_objective, lp.obj_opt) < 1e-3


def test_fp32_recentring_rescues_stalling_lps():
    # LPs of the batched workload on which plain FP32 iterates freeze (the PID
    # drives one step size below the iterates' rounding): with recentring they
    # converge in about the FP64 iteration count, to 1e-4 and to 1e-8
    lps = G.batch_lps(count=4096, seed=0)
    for i in (1672, 2723):
        lp = lps[i]
        d = c_api.solve(lp, tol=1e-4, precision="fp64", time_limit=TL)
        for tol in (1e-4, 1e-8):
            r = c_api.solve(lp, tol=tol, precision="fp32_mixed", time_limit=TL, iteration_limit=100000)
            assert r.status == 1, (i, tol, r.iterations)
            assert rel(r.primal_objective, lp.obj_opt) < 10 * tol
        plain = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=3 * d.iterations,
                            refinement=False)
        assert plain.status != 1  # the failure the recentring removes


def test_iteration_limit_semantics():
    prob = c_api.Problem.read_mps(os.path.join(GOLDEN, "afiro_original.mps"))
    lp = prob.lp
    r = c_api.solve(prob, tol=0.0, iteration_limit=10, time_limit=TL)
    o = O.solve(lp, tol=0.0, iteration_limit=10)
    assert r.status == 4 and r.iterations == 10
    assert np.any(r.x != 0.0)
    assert np.allclose(r.x, o["x"], atol=1e-12)


def test_time_limit_semantics():
    lp = G.random_lp(m=2000, n=5000, seed=0)
    r = c_api.solve(lp, tol=0.0, time_limit=0.5)
    assert r.status == 5
    assert r.solve_time < 5 * 0.5 + 2.0  # pdlp_test.cu:159-175 allows 5x


def test_initial_point_is_a_fixed_point():
    lp = G.random_lp(m=300, n=800, seed=8)
    r = c_api.solve(lp, tol=1e-8, x_init=lp.x_opt, y_init=lp.y_opt, time_limit=TL, iteration_limit=10000)
    assert r.status == 1 and r.iterations == 10
    assert np.allclose(r.x, lp.x_opt, atol=1e-8)


def test_deterministic_bitwise():
    lp = G.skewed_lp(m=20000, n=40000, seed=6)
    a = c_api.solve(lp, tol=1e-4, precision="fp64", time_limit=TL, iteration_limit=3000)
    b = c_api.solve(lp, tol=1e-4, precision="fp64", time_limit=TL, iteration_limit=3000)
    assert a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    c = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=3000)
    d = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=3000)
    assert np.array_equal(c.x, d.x) and np.array_equal(c.y, d.y)


def test_graph_replays_match_direct_launches(monkeypatch):
    # Iteration blocks between checks are replayed from captured graphs; every
    # replay must give bitwise the result of launching the same kernels directly.
    lp = G.skewed_lp(m=20000, n=40000, seed=7)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("CUOPT_AMD_GRAPHS", flag)
        for prec in ("fp64", "fp32_mixed"):
            out[flag, prec] = c_api.solve(lp, tol=1e-5, precision=prec, time_limit=TL, iteration_limit=50000)
    for prec in ("fp64", "fp32_mixed"):
        a, b = out["1", prec], out["0", prec]
        assert a.status == b.status == 1
        assert a.iterations == b.iterations
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
        assert np.array_equal(a.trace, b.trace, equal_nan=True)  # fpe is NaN before the first one


# ---- batched persistent solve (cuOptAmdBatchSolve) --------------------------------
# Same arithmetic as the single-LP path except the order of row reductions: FP64
# batch results match single-LP solves (status, objective to 1e-7 relative,
# iteration counts within a few checks) and the closed-form optima.
def test_batch_fp64_matches_single_solves(highs):
    lps = [G.random_lp(m=m, n=int(2.5 * m), seed=100 + m) for m in (50, 73, 120, 200, 333, 400)]
    lps.append(c_api.read_lp(os.path.join(GOLDEN, "afiro_original.mps")))
    lps.append(c_api.read_lp(os.path.join(GOLDEN, "good-max.mps")))
    lps.append(G.skewed_lp(m=2000, n=4000, seed=5))
    rs = c_api.batch_solve(lps, tol=1e-8, precision="fp64", time_limit=TL, iteration_limit=500000)
    for lp, r in zip(lps, rs):
        s = c_api.solve(lp, tol=1e-8, precision="fp64", time_limit=TL, iteration_limit=500000)
        assert r.status == s.status == 1, (lp.name, r.status, s.status)
        assert rel(r.primal_objective, s.primal_objective) < 1e-7, lp.name
        assert abs(r.iterations - s.iterations) <= max(200, s.iterations // 10), (lp.name, r.iterations, s.iterations)
        assert r.info["power_iterations"] == s.info["power_iterations"], lp.name
        assert abs(r.info["eta0"] - s.info["eta0"]) <= 1e-10 * s.info["eta0"]
        if lp.obj_opt is not None:
            assert rel(r.primal_objective, lp.obj_opt) < 1e-6
    assert rel(rs[6].primal_objective, highs["afiro_original.mps"]["objective"]) < 1e-6
    assert rel(rs[7].primal_objective, highs["good-max.mps"]["objective"]) < 1e-6


def test_batch_fp32_mixed_config0_style():
    lps = G.batch_lps(count=64, m_min=50, m_max=400, seed=1)
    rs = c_api.batch_solve(lps, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=200000)
    for lp, r in zip(lps, rs):
        assert r.status == 1 and r.info["precision"] == 1
        assert rel(r.primal_objective, lp.obj_opt) < 1e-3
    assert max(r.iterations for r in rs) < 20000  # no stalled member (recentring)


def test_batch_fp32_step_size_by_lanczos():
    # FP32-mixed batch members estimate sigma_max(A~)^2 as FP32-mixed single-LP solves
    # do: Lanczos on A~ A~^T from the same hashed start, with the Ritz value found by
    # Sturm counts.  Only the order of the row sums differs, so the estimate agrees
    # with the single-LP one far inside the 0.2 % margin of eta = 0.998 / sigma_max,
    # and both agree with a dense SVD of the oracle's scaled matrix.
    import scipy.sparse as sp

    lps = [G.random_lp(m=m, n=int(2.5 * m), seed=200 + m) for m in (50, 120, 400)]
    lps.append(G.skewed_lp(m=2000, n=4000, seed=5))
    rs = c_api.batch_solve(lps, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=0)
    for lp, r in zip(lps, rs):
        s = c_api.solve(lp, tol=1e-4, precision="fp32_mixed", time_limit=TL, iteration_limit=0)
        assert abs(r.info["sigma_max_sq"] - s.info["sigma_max_sq"]) <= 1e-10 * s.info["sigma_max_sq"], lp.name
        assert abs(r.info["power_iterations"] - s.info["power_iterations"]) <= 1, lp.name
        assert r.info["power_iterations"] < 100
        o = O.setup(lp)
        At = sp.csr_matrix((o["scaled_values"], lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
        smax2 = np.linalg.svd(At, compute_uv=False)[0] ** 2
        assert abs(r.info["sigma_max_sq"] - smax2) <= 1e-5 * smax2, lp.name
        assert abs(r.info["eta0"] - 0.998 / np.sqrt(r.info["sigma_max_sq"])) <= 1e-14 * r.info["eta0"]


def test_batch_fp32_recentring_to_1e8():
    lps = G.batch_lps(count=4096, seed=0)
    pick = [lps[i] for i in (1672, 2723, 3846, 0, 1, 2)]
    rs = c_api.batch_solve(pick, tol=1e-8, precision="fp32_mixed", time_limit=TL, iteration_limit=200000)
    for lp, r in zip(pick, rs):
        assert r.status == 1 and r.info["refine_rounds"] >= 1
        assert rel(r.primal_objective, lp.obj_opt) < 1e-7


def test_batch_edge_cases_keep_order():
    good = G.random_lp(m=60, n=150, seed=3)
    crossing = G.random_lp(m=60, n=150, seed=4)
    crossing.var_lb = crossing.var_lb.copy()
    crossing.var_lb[0] = 11.0  # above its upper bound of 10
    empty = c_api.read_lp(os.path.join(GOLDEN, "lp-model-no-constraints.mps"))
    rs = c_api.batch_solve([good, crossing, empty, good], tol=1e-6, time_limit=TL, iteration_limit=100000)
    assert [r.status for r in rs] == [1, 2, 6, 1]
    assert np.array_equal(rs[0].x, rs[3].x)


def _with_dense_row_and_column(lp, seed):
    """lp plus one never-binding dense <= row and one dense column (a new variable in
    [0, 10] with an entry in every row): rows and columns longer than the batch
    kernel's 2 552-nonzero row blocks when lp has more than 2 552 rows and columns."""
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    col = rng.standard_normal((lp.m, 1))
    A = sp.hstack([A, sp.csr_matrix(col)]).tocsr()
    row = rng.standard_normal(lp.n + 1)
    A = sp.vstack([A, sp.csr_matrix(row)]).tocsr()
    A.sort_indices()
    ub = np.append(lp.var_ub, 10.0)
    return G.LP(m=lp.m + 1, n=lp.n + 1, row_offsets=A.indptr.astype(np.int32), col_indices=A.indices.astype(np.int32),
                values=A.data.astype(np.float64), c=np.append(lp.c, rng.uniform(-1.0, 1.0)),
                con_lb=np.append(lp.con_lb, -np.inf), con_ub=np.append(lp.con_ub, np.abs(row) @ ub + 1.0),
                var_lb=np.append(lp.var_lb, 0.0), var_ub=ub, name=lp.name + "+dense")


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_batch_rows_longer_than_a_block(precision):
    # A row of 6 001 and a column of 3 001 nonzeros take the batch kernel's
    # in-place path (longer than its 2 552-nonzero row blocks); the single-LP
    # solver is the reference.
    lps = [_with_dense_row_and_column(G.random_lp(m=3000, n=6000, seed=11), 1),
           G.random_lp(m=80, n=200, seed=12)]
    assert np.diff(lps[0].row_offsets).max() == 6001
    rs = c_api.batch_solve(lps, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=500000)
    for lp, r in zip(lps, rs):
        s = c_api.solve(lp, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=500000)
        assert r.status == s.status == 1, (lp.name, r.status, s.status)
        assert rel(r.primal_objective, s.primal_objective) < 1e-5, lp.name


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp64"])
def test_full_size_batch_of_4096_lps(precision):
    # the batched workload at its full size (4 096 LPs, m log-uniform in [50, 2 000])
    # in one launch: every LP meets the 1e-4 relative KKT test it reports, and a
    # sample checked in numpy has its primal residual within the tolerance and its
    # objective near the built-in optimum
    lps = G.batch_lps(count=4096, seed=0)
    rs = c_api.batch_solve(lps, tol=1e-4, precision=precision, time_limit=TL)
    assert all(r.status == 1 for r in rs)
    for r in rs:
        assert r.info["l2_relative_primal_residual"] <= 1e-4 * (1 + 1e-9)
        assert r.info["l2_relative_dual_residual"] <= 1e-4 * (1 + 1e-9)
        assert r.info["relative_gap"] <= 1e-4 * (1 + 1e-9)
    for i in np.random.default_rng(0).choice(len(lps), 32, replace=False):
        lp, r = lps[i], rs[i]
        rp, bv = primal_violation(lp, r.x)
        assert rp <= 1e-4 * (1.0 + bound_norm(lp)) * 1.01 and bv <= 1e-12
        assert abs(r.primal_objective - lp.obj_opt) <= 1e-2 * (1 + abs(lp.obj_opt))


def test_batch_iteration_limit():
    lps = [G.random_lp(m=100, n=250, seed=s) for s in range(4)]
    rs = c_api.batch_solve(lps, tol=0.0, time_limit=TL, iteration_limit=50)
    assert all(r.status == 4 and r.iterations == 50 for r in rs)


def test_batch_time_limit_counts_per_lp_from_its_dispatch():
    # time_limit applies to each LP from the start of its own workgroup (as each of the
    # reference's per-LP solve_lp calls starts its own timer).  A batch four times the
    # resident capacity (two workgroups per CU) of LPs that cannot converge (tolerance 0)
    # ends every LP with TimeLimit, and the call takes about one limit per dispatch wave.
    import torch

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    count, limit = 8 * cus, 0.05
    waves = -(-count // (2 * cus))
    lps = [G.random_lp(m=50, n=125, seed=s) for s in range(count)]
    print(f"batch time limit: {count} LPs, {waves} waves, limit {limit}s", flush=True)
    rs = c_api.batch_solve(lps, tol=0.0, time_limit=limit, iteration_limit=10**9)
    assert all(r.status == c_api.CUOPT_TERIMINATION_STATUS_TIME_LIMIT for r in rs)
    kernel = rs[0].info["loop_time"]
    assert limit <= kernel <= (waves + 1) * limit * 1.5 + 0.5, (kernel, waves)


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_spmv_autotune_picks_a_candidate_caches_it_and_solves(precision):
    # pdlp_spmv_autotune = 1: a tile capacity in {512, 1024, 2048} and a lanes shift in
    # {-1, 0, 1} for each of A and A^T, timed at setup on the first solve of a matrix and
    # cached for the next; the tuned solve reaches the optimum (only the summation order
    # of row sums differs from the untuned solve).
    lp = G.skewed_lp(m=20000, n=40000, seed=9)
    tune = {"pdlp_spmv_autotune": 1}
    first = c_api.solve(lp, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=200000, params=tune)
    again = c_api.solve(lp, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=200000, params=tune)
    plain = c_api.solve(lp, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=200000)
    for k in ("tile_a", "tile_at"):
        assert first.info[k] in (512, 1024, 2048) and again.info[k] == first.info[k]
    for k in ("lanes_bias_a", "lanes_bias_at"):
        assert first.info[k] in (-1, 0, 1) and again.info[k] == first.info[k]
    assert first.info["tune_time"] > 0 and again.info["tune_time"] == 0
    assert plain.info["tile_a"] == plain.info["tile_at"] == 1024 and plain.info["tune_time"] == 0
    assert first.status == again.status == plain.status == 1
    assert np.array_equal(first.x, again.x)  # the same plan: bitwise the same solve
    # both stop within the 1e-6 relative KKT test, on either side of the optimum
    assert rel(first.primal_objective, plain.primal_objective) < 1e-5
    assert rel(first.primal_objective, lp.obj_opt) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_column_relabelling_is_the_same_lp(precision):
    # pdlp_column_order = 1 numbers the columns internally by first touch of the short
    # rows; inputs (c, bounds, the initial point) and outputs (x, reduced costs) stay in
    # the caller's order.  The relabelled solve reaches the same optimum as the plain one,
    # and a start at the built-in optimum is still a fixed point (x returned unchanged).
    import scipy.sparse as sp

    lp = G.skewed_lp(m=20000, n=40000, seed=11)
    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    plain = c_api.solve(lp, tol=1e-8, precision=precision, time_limit=TL, iteration_limit=400000,
                        params={"pdlp_column_order": 0})
    relab = c_api.solve(lp, tol=1e-8, precision=precision, time_limit=TL, iteration_limit=400000,
                        params={"pdlp_column_order": 1})
    assert plain.status == relab.status == 1
    assert rel(relab.primal_objective, plain.primal_objective) < 1e-7
    assert rel(relab.primal_objective, lp.obj_opt) < 1e-6
    ax = A @ relab.x
    b = np.where(np.isfinite(lp.con_ub), lp.con_ub, 0.0)
    assert np.linalg.norm(ax - np.clip(ax, lp.con_lb, lp.con_ub)) <= 1e-6 * (1 + np.linalg.norm(b))
    assert np.all(relab.x >= lp.var_lb - 1e-9) and np.all(relab.x <= lp.var_ub + 1e-9)
    # reduced costs in the caller's order: c - A^T y minus the dual residual
    assert np.linalg.norm(lp.c - A.T @ relab.y - relab.reduced_cost) <= 1e-6 * (1 + np.linalg.norm(lp.c))
    fixed = c_api.solve(lp, tol=1e-8, precision=precision, x_init=lp.x_opt, y_init=lp.y_opt, time_limit=TL,
                        iteration_limit=10000) if precision == "fp64" else None
    if fixed is not None:
        fx = c_api.solve(lp, tol=1e-8, precision=precision, x_init=lp.x_opt, y_init=lp.y_opt, time_limit=TL,
                         iteration_limit=10000, params={"pdlp_column_order": 1})
        assert fx.status == fixed.status == 1 and fx.iterations == fixed.iterations == 10
        assert np.allclose(fx.x, lp.x_opt, atol=1e-8)


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("shape", [(20000, 40000, 5), (100000, 200000, 3)])
def test_long_row_merge_forms_are_bitwise_the_same(precision, shape):
    # pdlp_long_row_merge: rows longer than an SpMV tile are summed in chunks, merged
    # either by a launch of their own (0) or by the row's last chunk inside the SpMV
    # launch through device-scope atomics (1).  The merge tree is the same, so the
    # whole solve (graph-captured iteration blocks included) is bitwise the same.
    # The larger LP has rows of more than 64 chunks, where -1 keeps the merge launch.
    m, n, seed = shape
    lp = G.skewed_lp(m=m, n=n, seed=seed)
    runs = {mode: c_api.solve(lp, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=3000,
                              params={"pdlp_column_order": 1, "pdlp_long_row_merge": mode})
            for mode in (0, 1, -1)}
    assert runs[0].info["merge_a"] == 0 and runs[1].info["merge_a"] == 1
    assert runs[-1].info["merge_a"] == (1 if m == 20000 else 0)
    for mode in (1, -1):
        assert runs[mode].iterations == runs[0].iterations
        for f in ("x", "y", "reduced_cost"):
            assert np.array_equal(getattr(runs[mode], f), getattr(runs[0], f)), (mode, f)


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("tile", [512, 2048])
def test_resident_grid_matches_one_workgroup_per_tile(monkeypatch, precision, tile):
    # The tiled SpMV launches as many workgroups as are resident and loops each over
    # tiles b, b + grid, ... (reusing its LDS buffers after a barrier).  Against one
    # workgroup per tile (CUOPT_AMD_PERSISTENT=0) the whole solve is bitwise the same,
    # at the tile capacities the tuner may choose besides the default 1 024; the LP has
    # rows longer than a tile in both merge forms' range.
    lp = G.skewed_lp(m=20000, n=40000, seed=7)
    monkeypatch.setenv("CUOPT_AMD_SPMV_TILE_A", str(tile))
    monkeypatch.setenv("CUOPT_AMD_SPMV_TILE_AT", str(tile))
    runs = {}
    for persistent in ("0", "1"):
        monkeypatch.setenv("CUOPT_AMD_PERSISTENT", persistent)
        print(f"tile {tile} {precision} persistent={persistent}", flush=True)
        runs[persistent] = c_api.solve(lp, tol=1e-6, precision=precision, time_limit=TL, iteration_limit=200000)
    a, b = runs["0"], runs["1"]
    assert a.info["tile_a"] == a.info["tile_at"] == tile
    assert a.status == b.status == 1 and a.iterations == b.iterations
    for f in ("x", "y", "reduced_cost"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert rel(b.primal_objective, lp.obj_opt) < 1e-5


def _highs(lp):
    """HiGHS (SciPy linprog) on a parsed LP: (status, objective with the offset, in the user's sense)."""
    import scipy.sparse as sp
    from scipy.optimize import linprog

    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    eq = lp.con_lb == lp.con_ub
    up, lo = np.isfinite(lp.con_ub) & ~eq, np.isfinite(lp.con_lb) & ~eq
    sg = -1.0 if lp.maximize else 1.0
    A_ub = sp.vstack([A[up], -A[lo]]) if (up.any() or lo.any()) else None
    b_ub = np.concatenate([lp.con_ub[up], -lp.con_lb[lo]]) if A_ub is not None else None
    ref = linprog(sg * lp.c, A_ub=A_ub, b_ub=b_ub, A_eq=A[eq] if eq.any() else None,
                  b_eq=lp.con_lb[eq] if eq.any() else None, bounds=list(zip(lp.var_lb, lp.var_ub)), method="highs")
    return ref.status, (sg * ref.fun + lp.offset if ref.status == 0 else None)


# The reference parser corpus's LP fixtures that HiGHS solves to an optimum
# (datasets/linear_programming, read by mps_parser_test.cpp): solved through the
# read handle in both precisions.
PARSER_OPTIMAL = ["free-format-mps-1.mps", "good-mps-1.mps", "good-mps-1-clrf.mps", "good-mps-1-comments.mps",
                  "good-mps-fixed-empty-name.mps", "good-mps-fixed-no-name.mps",
                  "good-mps-fixed-two-objectives.mps", "good-mps-fixed-two-objectives-objname.mps",
                  "good-mps-fixed-two-objectives-objname-next-line.mps", "good-mps-rhs-cost.mps",
                  "good-mps-some-var-bounds.mps", "good-mps-upper-bound-inf-var.mps"]


@pytest.mark.parametrize("fname", PARSER_OPTIMAL)
def test_reference_parser_fixtures_solve_to_the_highs_optimum(fname):
    prob = c_api.Problem.read_mps(os.path.join(GOLDEN, "mps", fname))
    lp = prob.lp
    status, obj = _highs(lp)
    assert status == 0
    for precision in ("fp64", "fp32_mixed"):
        r = c_api.solve(prob, tol=1e-8, precision=precision, time_limit=TL, iteration_limit=200000)
        assert r.status == 1, (fname, precision, r.status)
        assert abs(r.primal_objective - obj) <= 1e-6 * (1 + abs(obj)), (fname, precision)
        if precision == "fp64":
            assert r.iterations == O.solve(lp, tol=1e-8)["iterations"], fname
