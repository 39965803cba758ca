"""Pins the CPU FP64 oracle before anything is checked against it.

Each test checks the oracle against something it cannot have copied from itself:
reference known-answer values, closed-form optima of the generated LPs, HiGHS
optima committed under tests/golden/, a dense SVD, or exact integer identities.
A planted mistake test shows the KKT check would catch a dropped term.
"""

import json
import os

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd import generators as G
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


@pytest.fixture(scope="module")
def highs():
    with open(os.path.join(GOLDEN, "highs_optima.json")) as f:
        return json.load(f)["cases"]


def test_transpose_is_an_involution():
    lp = G.skewed_lp(m=3000, n=6000, seed=1)
    tp, ti, tv = O.transpose(lp.row_offsets, lp.col_indices, lp.values, lp.n)
    # rows of A^T are sorted by row index of A (stable counting sort)
    for j in range(0, lp.n, 97):
        seg = ti[tp[j]:tp[j + 1]]
        assert np.all(np.diff(seg) > 0)
    bp, bi, bv = O.transpose(tp, ti, tv, lp.m)
    assert np.array_equal(bp, lp.row_offsets)
    assert np.array_equal(bi, lp.col_indices)
    assert np.array_equal(bv, lp.values)


def test_spmv_matches_scipy():
    lp = G.random_lp(m=300, n=700, seed=2)
    x = np.random.default_rng(0).standard_normal(lp.n)
    y = O.spmv(lp.row_offsets, lp.col_indices, lp.values, x, lp.n)
    ref = lp.to_scipy() @ x
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_range_rows_map_to_closed_form_bounds():
    lp = c_api.read_lp(os.path.join(GOLDEN, "good-mps-fixed-ranges.mps"))
    # Bounds stated in the fixture's own header comment.
    assert np.allclose(lp.con_lb, [1.2, 1.5, 7.9, 3.5, 3.9, 4.9])
    assert np.allclose(lp.con_ub, [5.4, 4.9, 9.5, 6.9, 3.9, 4.9])


def test_methodical1_initial_step_and_weight_known_answer():
    # Reference KAT: cpp/tests/linear_programming/pdlp_test.cu:243-287 (afiro, Methodical1,
    # tolerance 1e-4): initial step size 1.4893, initial primal weight 0.0141652.
    lp = c_api.read_lp(os.path.join(GOLDEN, "afiro_original.mps"))
    s = O.setup(lp, mode=1)
    assert abs(s["eta"] - 1.4893) <= 1e-4
    assert abs(s["omega"] - 0.0141652) <= 1e-4


def test_power_iteration_matches_dense_svd():
    lp = G.random_lp(m=400, n=1000, seed=3)
    s = O.setup(lp)
    import scipy.sparse as sp

    D = sp.csr_matrix((s["scaled_values"], lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
    smax2 = np.linalg.svd(D, compute_uv=False)[0] ** 2
    assert abs(s["sigma_max_sq"] - smax2) / smax2 < 1e-3
    assert abs(s["eta"] - 0.998 / np.sqrt(s["sigma_max_sq"])) < 1e-15


def test_scaling_of_diagonal_matrix_is_unit():
    # Textbook special case: for A = diag(d), one Ruiz sweep gives D_r = D_c = 1/sqrt|d|,
    # so |a~_ii| = 1; later sweeps and Pock-Chambolle (alpha = 1) leave it there.
    d = np.array([4.0, 0.25, -9.0, 100.0, 1e-3])
    n = d.size
    lp = G.LP(m=n, n=n, row_offsets=np.arange(n + 1, dtype=np.int32), col_indices=np.arange(n, dtype=np.int32),
              values=d, c=np.ones(n), con_lb=np.ones(n), con_ub=np.ones(n), var_lb=np.zeros(n),
              var_ub=np.full(n, np.inf))
    s = O.setup(lp)
    assert np.allclose(np.abs(s["scaled_values"]), 1.0, rtol=1e-15, atol=0)
    assert np.allclose(s["Dr"], 1.0 / np.sqrt(np.abs(d)), rtol=1e-15)


def test_scaling_is_permutation_equivariant():
    # Permuting rows and columns must permute D_r and D_c (a swapped index would not).
    lp = G.skewed_lp(m=1500, n=3000, seed=4)
    rng = np.random.default_rng(0)
    pr, pc = rng.permutation(lp.m), rng.permutation(lp.n)
    A = lp.to_scipy()[pr][:, pc].tocsr()
    A.sort_indices()
    lq = G.LP(m=lp.m, n=lp.n, row_offsets=A.indptr.astype(np.int32), col_indices=A.indices.astype(np.int32),
              values=A.data, c=lp.c[pc], con_lb=lp.con_lb[pr], con_ub=lp.con_ub[pr], var_lb=lp.var_lb[pc],
              var_ub=lp.var_ub[pc])
    s, t = O.setup(lp), O.setup(lq)
    assert np.allclose(t["Dr"], s["Dr"][pr], rtol=1e-13)
    assert np.allclose(t["Dc"], s["Dc"][pc], rtol=1e-13)
    assert abs(t["beta"] - s["beta"]) < 1e-13 * s["beta"]
    assert abs(t["sigma_max_sq"] - s["sigma_max_sq"]) < 1e-3 * s["sigma_max_sq"]


@pytest.mark.parametrize("gen", ["random", "skewed"])
def test_kkt_vanishes_at_constructed_optimum(gen):
    lp = G.random_lp(m=300, n=800, seed=5) if gen == "random" else G.skewed_lp(m=2000, n=4000, seed=6)
    k = O.kkt(lp, lp.x_opt, lp.y_opt, lp.s_opt)
    scale = 1.0 + np.abs(lp.values).sum()
    assert k["l2_primal"] < 1e-12 * scale
    assert k["l2_dual"] < 1e-12 * scale
    assert k["gap"] < 1e-10 * (1 + abs(lp.obj_opt))
    assert rel(k["pobj"], lp.obj_opt) < 1e-13
    # obj* = c.x0 = b.y0 (only the equality rows carry y0)
    assert rel(float(lp.con_ub @ lp.y_opt), lp.obj_opt) < 1e-10


def test_kkt_detects_planted_error():
    lp = G.random_lp(m=300, n=800, seed=5)
    y = lp.y_opt.copy()
    y[17] += 1e-3  # wrong dual entry -> dual residual must appear
    k = O.kkt(lp, lp.x_opt, y, lp.s_opt)
    assert k["l2_dual"] > 1e-4
    x = lp.x_opt.copy()
    x[3] += 1e-3  # wrong primal entry -> primal residual must appear
    k = O.kkt(lp, x, lp.y_opt, lp.s_opt)
    assert k["l2_primal"] > 1e-4


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_solve_reaches_closed_form_optimum(tol):
    lp = G.random_lp(m=500, n=1250, seed=7)
    r = O.solve(lp, tol=tol)
    assert r["status"] == 1
    assert r["l2_primal_residual"] <= tol + tol * r["b_norm"]
    assert r["l2_dual_residual"] <= tol + tol * r["c_norm"]
    assert rel(r["primal_objective"], lp.obj_opt) < 50 * tol
    ind = G.kkt_residuals(lp, r["x"], r["y"])
    assert ind["bound_violation"] == 0.0
    assert ind["primal_residual"] <= 2 * (tol + tol * r["b_norm"])


def test_skewed_solve_reaches_closed_form_optimum(highs):
    lp = G.skewed_lp(m=2000, n=4000, seed=5)
    r = O.solve(lp, tol=1e-8)
    assert r["status"] == 1
    assert rel(r["primal_objective"], lp.obj_opt) < 1e-6
    assert rel(lp.obj_opt, highs["skewed_small"]["objective"]) < 1e-9


@pytest.mark.parametrize("name", ["transport_small", "random_small"])
def test_solve_matches_highs(highs, name):
    case = highs[name]
    lp = getattr(G, case["generator"])(**case["kwargs"])
    r = O.solve(lp, tol=1e-8)
    assert r["status"] == 1
    assert rel(r["primal_objective"], case["objective"]) < 1e-6


@pytest.mark.parametrize("fname,tol,bound", [
    ("afiro_original.mps", 1e-8, 1e-6),
    ("good-max.mps", 1e-8, 1e-6),
    ("max_offset.mps", 1e-8, 1e-6),
    ("lp_model_with_var_bounds.mps", 1e-8, 1e-6),
])
def test_mps_fixtures(highs, fname, tol, bound):
    lp = c_api.read_lp(os.path.join(GOLDEN, fname))
    r = O.solve(lp, tol=tol)
    assert r["status"] == 1
    assert rel(r["primal_objective"], highs[fname]["objective"]) < bound
    if fname == "afiro_original.mps":
        assert abs(r["primal_objective"] - (-464.7531)) < 1e-3  # test_lp_solver.py:92-111


def test_maximize_negates_duals():
    lp = c_api.read_lp(os.path.join(GOLDEN, "good-max.mps"))
    r = O.solve(lp, tol=1e-8)
    # max 2x1+3x2+x3 s.t. x1+2x2+x3 <= 10: the constraint dual of the max problem is >= 0
    # in the maximisation convention after the sign flip of solve.cu:583-588.
    assert r["y"][0] > 0
    assert abs(r["dual_objective"] - 17.0) < 1e-5


def test_iteration_limit_returns_current_iterate():
    # pdlp_test.cu:124-147: iteration limit 10 with zero tolerance -> IterationLimit, x != 0.
    lp = c_api.read_lp(os.path.join(GOLDEN, "afiro_original.mps"))
    r = O.solve(lp, tol=0.0, iteration_limit=10)
    assert r["status"] == 4
    assert r["iterations"] == 10
    assert np.any(r["x"] != 0.0)


def test_no_constraints_is_numerical_error_for_pdlp():
    lp = c_api.read_lp(os.path.join(GOLDEN, "lp-model-no-constraints.mps"))
    r = O.solve(lp)
    assert r["status"] == 6  # run_pdlp_solver, solve.cu:563-568


def test_fixed_point_at_optimum():
    # PDHG's fixed points are the saddle points: started at (x*, y*) the first
    # allowed check (k = 10) already reports optimality.
    lp = G.random_lp(m=300, n=800, seed=8)
    r = O.solve(lp, tol=1e-8, x_init=lp.x_opt, y_init=lp.y_opt)
    assert r["status"] == 1
    assert r["iterations"] == 10
    assert np.allclose(r["x"], lp.x_opt, atol=1e-8)


def test_check_schedule_and_restarts():
    lp = G.random_lp(m=500, n=1250, seed=9)
    r = O.solve(lp, tol=1e-10, iteration_limit=3000, trace_cap=200)
    t = r["trace"]
    ks = t[:, 0].astype(int)
    # checks every 10 below 1000, then every 100 (utils.cuh:62-81)
    assert list(ks[:101]) == list(range(0, 1001, 10))
    assert np.all(np.diff(ks[ks >= 1000]) == 100)
    restarted = ks[t[:, 6] == 1]
    assert restarted[0] == 200  # forced first restart (pdlp_restart_strategy.cu:694-698)
    assert np.all(restarted % 200 == 0)
