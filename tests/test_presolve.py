"""LP presolve and postsolve (presolve.cpp).

The reference presolves with a third-party library (PSLP, solve.cu:1183-1325)
whose reductions are not reproducible here; what is pinned is the contract:
the postsolved (x, y, reduced costs) is an optimal primal-dual pair of the
ORIGINAL problem (checked independently in numpy: feasibility, sign
conditions, complementary slackness, objective), infeasible / unbounded
detections are right, and solves with and without presolve agree.  Problems
that presolve solves completely need no GPU."""

import numpy as np
import pytest
import scipy.sparse as sp

from cuopt_amd import c_api
from cuopt_amd import generators as G


def kkt_ok(lp, r, tol=1e-7):
    """Optimality of (x, y, z) for lp in numpy (minimisation sense after the max flip)."""
    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    x, y, z = r.x, r.y, r.reduced_cost
    sgn = -1.0 if lp.maximize else 1.0
    ax = A @ x
    scale = 1.0 + np.abs(ax).max(initial=0.0)
    assert np.all(ax >= lp.con_lb - tol * scale) and np.all(ax <= lp.con_ub + tol * scale)
    assert np.all(x >= lp.var_lb - tol) and np.all(x <= lp.var_ub + tol)
    assert np.allclose(z, lp.c - A.T @ y, atol=1e-9 * (1 + np.abs(lp.c).max(initial=0)))
    zs, ys = sgn * z, sgn * y
    zt = tol * (1 + np.abs(lp.c).max(initial=0.0))
    for j in range(lp.n):
        at_lo = x[j] <= lp.var_lb[j] + tol
        at_hi = x[j] >= lp.var_ub[j] - tol
        if not at_lo:
            assert zs[j] <= zt, ("reduced cost", j, zs[j])
        if not at_hi:
            assert zs[j] >= -zt, ("reduced cost", j, zs[j])
    for i in range(lp.m):
        lo = ax[i] <= lp.con_lb[i] + tol * scale
        hi = ax[i] >= lp.con_ub[i] - tol * scale
        if not lo:
            assert ys[i] <= zt, ("row dual", i, ys[i])
        if not hi:
            assert ys[i] >= -zt, ("row dual", i, ys[i])
    return True


def lp_from(A, c, lc, uc, lb, ub, maximize=False):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return G.LP(m=A.shape[0], n=A.shape[1], row_offsets=A.indptr.astype(np.int32),
                col_indices=A.indices.astype(np.int32), values=A.data.astype(np.float64), c=np.asarray(c, float),
                con_lb=np.asarray(lc, float), con_ub=np.asarray(uc, float), var_lb=np.asarray(lb, float),
                var_ub=np.asarray(ub, float), maximize=maximize, name="presolve_case")


INF = np.inf


def fully_reducible():
    # min x0 + 2 x1 - x2 + 0 x3 with
    #   r0: 2 x0 >= 4 (singleton), r1: -x1 <= -1 (singleton, negative coefficient),
    #   r2: empty row 0 = 0, r3: x2 <= 5 (singleton), r4: x0 + x3 >= 4 with x3 fixed at 3.
    # Optimum x = (2, 1, 5, 3), objective -1; duals y = (1/2, -2, 0, -1, 0), z = 0.
    A = np.array([[2, 0, 0, 0], [0, -1, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0], [1, 0, 0, 1]], float)
    return lp_from(A, [1, 2, -1, 0], [4, -INF, 0, -INF, 4], [INF, -1, 0, 5, INF], [0, 0, 0, 3], [INF, INF, INF, 3])


def test_presolve_solves_singletons_fixed_and_empty_rows_without_gpu():
    lp = fully_reducible()
    r = c_api.solve(lp, tol=1e-8, presolve=True)
    assert r.error == 0 and r.status == 1
    # termination information of the original problem, no NaN
    assert r.info["l2_primal_residual"] < 1e-12 and r.info["l2_dual_residual"] == 0.0
    assert r.info["gap"] == 0.0 and r.info["primal_objective"] == r.primal_objective
    assert np.allclose(r.x, [2, 1, 5, 3]) and abs(r.primal_objective - (-1.0)) < 1e-12
    assert np.allclose(r.y, [0.5, -2, 0, -1, 0]) and np.allclose(r.reduced_cost, 0)
    assert kkt_ok(lp, r)


def test_rows_of_explicit_zeros_free_their_columns():
    # row 0 holds only an explicit zero coefficient on x0 (an empty row in effect);
    # x0, then in no row, must take the bound its cost prefers (5), not 0
    A = sp.csr_matrix((np.array([0.0, 1.0]), np.array([0, 1]), np.array([0, 1, 2])), shape=(2, 2))
    lp = lp_from(A, [1, 1], [-1, 2], [1, INF], [5, 0], [9, 10])
    r = c_api.solve(lp, tol=1e-8, presolve=True)
    assert r.error == 0 and r.status == 1
    assert np.allclose(r.x, [5, 2]) and abs(r.primal_objective - 7.0) < 1e-12
    assert kkt_ok(lp, r)


def test_file_parameters_without_gpu(tmp_path):
    # solve.cu:1281-1284 and 1327-1330 on a problem that presolve solves: the user
    # problem is written as MPS before the solve and reads back as the same LP; the
    # .sol file has the reference's layout (solution_writer.cu:16-40), without
    # variable lines when the problem has no names

    lp = fully_reducible()
    prob = c_api.Problem(lp)
    st = c_api.make_settings(1e-8, presolve=True)
    assert st.set("user_problem_file", str(tmp_path / "user.mps")) == 0
    assert st.set("solution_file", str(tmp_path / "out.sol")) == 0
    r = c_api.solve_problem(prob, st)
    assert r.status == 1
    back = c_api.read_lp(str(tmp_path / "user.mps"))
    assert np.allclose(back.c, lp.c) and np.allclose(back.values, lp.values)
    assert np.allclose(back.var_lb, lp.var_lb) and np.allclose(back.var_ub, lp.var_ub)
    lines = (tmp_path / "out.sol").read_text().splitlines()
    assert lines == ["# Status: Optimal", "# Objective value: -1"]


def test_dual_postsolve_off_returns_no_duals():
    # pdlp_test.cu:903-930: with presolve and dual_postsolve false the solution has no
    # duals (the C getters then fill zeros); the primal solution is unchanged
    lp = fully_reducible()
    prob = c_api.Problem(lp)
    st = c_api.make_settings(1e-8, presolve=True)
    assert st.set("dual_postsolve", False) == 0
    r = c_api.solve_problem(prob, st)
    assert r.status == 1 and np.allclose(r.x, [2, 1, 5, 3])
    assert not np.any(r.y) and not np.any(r.reduced_cost)


def test_presolve_maximisation_and_empty_columns():
    # max 3 x0 - x1 + 2 x2 s.t. x0 <= 4 (singleton), x1 >= -2 via -x1 <= 2; x2 in no row, in [0, 7].
    A = sp.csr_matrix(np.array([[1.0, 0, 0], [0, -1.0, 0]]))
    lp = lp_from(A, [3, -1, 2], [-INF, -INF], [4, 2], [-INF, -INF, 0], [INF, INF, 7], maximize=True)
    r = c_api.solve(lp, tol=1e-8, presolve=True)
    assert r.status == 1 and np.allclose(r.x, [4, -2, 7])
    assert abs(r.primal_objective - (12 + 2 + 14)) < 1e-12
    assert kkt_ok(lp, r)


def test_presolve_detects_infeasible_and_unbounded():
    # empty equality row with b = 1
    lp = lp_from(np.array([[0.0, 0.0], [1.0, 1.0]]), [1, 1], [1, 0], [1, 10], [0, 0], [5, 5])
    assert c_api.solve(lp, presolve=True).status == 2
    # singleton rows giving x0 >= 3 and x0 <= 2
    lp = lp_from(np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 1.0]]), [1, 1], [3, -INF, 0], [INF, 2, 10], [0, 0], [9, 9])
    assert c_api.solve(lp, presolve=True).status == 2
    # a variable fixed at +inf has no finite value
    lp = lp_from(np.array([[1.0, 1.0]]), [1, 1], [0], [10], [0, INF], [5, INF])
    assert c_api.solve(lp, presolve=True).status == 2
    # x1 in no row, cost -1, no upper bound: unbounded (as the reference reports it: DualInfeasible)
    lp = lp_from(np.array([[1.0, 0.0]]), [1, -1], [0], [1], [0, 0], [5, INF])
    assert c_api.solve(lp, presolve=True).status == 3


def test_reference_infeasible_problem_by_row_activity():
    # c_api_test.c:791-900 (solved there by dual simplex, status Infeasible): x4 >= 5
    # and x5 <= 2 are singleton rows, and then 3 x4 - x5 >= 13 > 2 in row 5
    ptr = [0, 2, 4, 6, 7, 9, 10, 12, 15, 17]
    idx = [0, 1, 0, 1, 0, 1, 3, 2, 3, 2, 0, 3, 0, 1, 2, 1, 2]
    val = [-0.5, 1.0, 2.0, -1.0, 3.0, 1.0, 1.0, 3.0, -1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0]
    rhs = np.array([0.5, 3.0, 6.0, 2.0, 2.0, 5.0, 10.0, 14.0, 1.0])
    ge = np.array([1, 1, 0, 0, 0, 1, 0, 0, 1], bool)
    A = sp.csr_matrix((val, idx, ptr), shape=(9, 4))
    lp = lp_from(A, [0, 0, 0, 0], np.where(ge, rhs, -INF), np.where(ge, INF, rhs), [0] * 4, [INF] * 4)
    r = c_api.solve(lp, presolve=True)
    assert r.error == 0 and r.status == 2
    # the activity test alone: one two-variable row that its bounds cannot reach
    lp = lp_from(np.array([[1.0, 1.0]]), [1, 1], [30], [INF], [0, 0], [10, 10])
    assert c_api.solve(lp, presolve=True).status == 2
    # ... and one that they just reach is kept (solved on the GPU when present)
    lp = lp_from(np.array([[1.0, 1.0]]), [1, 1], [20], [INF], [0, 0], [10, 10])
    r = c_api.solve(lp, presolve=True)
    assert r.status != 2


def test_presolve_off_keeps_the_no_constraint_rule():
    lp = fully_reducible()
    r = c_api.solve(lp, presolve=False)
    if c_api.device_count() == 0:
        assert r.error == c_api.CUOPT_RUNTIME_ERROR  # the reduced problem would need the GPU
    else:
        assert r.status == 1


# ---- GPU: problems with a part left for PDLP ---------------------------------------

def augmented(seed):
    """A generated LP with singleton rows, a fixed column, empty rows and columns added."""
    base = G.random_lp(m=120, n=300, seed=seed)
    A = sp.csr_matrix((base.values, base.col_indices, base.row_offsets), shape=(base.m, base.n)).tolil()
    m, n = base.m, base.n
    rng = np.random.default_rng(seed)
    A.resize((m + 4, n + 2))
    lc = list(base.con_lb) + [0.0, -INF, 0.0, 0.5]
    uc = list(base.con_ub) + [0.0, INF, INF, 1.5]
    A[m + 1, 5] = 2.0          # singleton: 2 x5 in (-inf, inf): a free row
    A[m + 2, 7] = -1.0         # singleton: -x7 >= 0 -> x7 <= 0, with x7 >= 0: fixes x7 at 0
    A[m + 3, n] = 1.0          # singleton on a new column: x_n in [0.5, 1.5]
    A[m + 3, n] = 1.0
    c = np.append(base.c, [1.0, -0.5])  # x_n (cost 1) goes to 0.5; x_{n+1} (in no row) to its upper bound 4
    lb = np.append(base.var_lb, [0.0, 0.0])
    ub = np.append(base.var_ub, [10.0, 4.0])
    return lp_from(A.tocsr(), c, lc, uc, lb, ub)


gpu = pytest.mark.gpu


@gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_presolve_matches_the_plain_solve(precision):
    for seed in (1, 2):
        lp = augmented(seed)
        a = c_api.solve(lp, tol=1e-8, precision=precision, time_limit=120.0, presolve=False)
        b = c_api.solve(lp, tol=1e-8, precision=precision, time_limit=120.0, presolve=True)
        assert a.status == b.status == 1
        assert abs(a.primal_objective - b.primal_objective) <= 1e-6 * (1 + abs(a.primal_objective))
        assert kkt_ok(lp, b, tol=1e-5)


def with_slack_columns(base, seed, frac=0.3):
    """base plus a private free column (a signed slack with a random cost) in a random
    share of its rows, which become ranged: presolve substitutes those rows away and the
    GPU solves what is left."""
    rng = np.random.default_rng(seed)
    A0 = sp.csr_matrix((base.values, base.col_indices, base.row_offsets), shape=(base.m, base.n))
    rows = np.flatnonzero(rng.random(base.m) < frac)
    S = sp.csr_matrix((rng.choice([-1.0, 1.0], len(rows)) * rng.uniform(0.5, 2.0, len(rows)),
                       (rows, np.arange(len(rows)))), shape=(base.m, len(rows)))
    A = sp.hstack([A0, S]).tocsr()
    lc, uc = base.con_lb.copy(), base.con_ub.copy()
    lc[rows] -= rng.uniform(0.1, 1.0, len(rows))
    uc[rows] += rng.uniform(0.1, 1.0, len(rows))
    c = np.concatenate([base.c, rng.uniform(-0.05, 0.05, len(rows))])
    lb = np.concatenate([base.var_lb, np.full(len(rows), -INF)])
    ub = np.concatenate([base.var_ub, np.full(len(rows), INF)])
    return lp_from(A, c, lc, uc, lb, ub)


@gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_presolve_with_column_singletons_matches_the_plain_solve(precision):
    # A 300 x 750 random LP whose rows partly carry free slack columns: presolve substitutes
    # those rows away (each settles at the bound its slack's cost prefers) and the GPU
    # solves the rest; the postsolved pair is optimal for the original LP and has the
    # objective of the solve without presolve.
    lp = with_slack_columns(G.random_lp(m=300, n=750, seed=4), seed=4)
    a = c_api.solve(lp, tol=1e-8, precision=precision, time_limit=120.0, presolve=False)
    b = c_api.solve(lp, tol=1e-8, precision=precision, time_limit=120.0, presolve=True)
    assert a.status == b.status == 1
    assert abs(a.primal_objective - b.primal_objective) <= 1e-6 * (1 + abs(a.primal_objective))
    assert kkt_ok(lp, b, tol=1e-5)


@gpu
def test_presolve_on_skewed_lp_with_empty_rows():
    lp = G.skewed_lp(m=4000, n=8000, seed=3)  # many empty rows
    a = c_api.solve(lp, tol=1e-8, time_limit=120.0, presolve=False)
    b = c_api.solve(lp, tol=1e-8, time_limit=120.0, presolve=True)
    assert a.status == b.status == 1
    assert abs(b.primal_objective - lp.obj_opt) <= 1e-6 * (1 + abs(lp.obj_opt))
    assert abs(a.primal_objective - b.primal_objective) <= 1e-6 * (1 + abs(lp.obj_opt))
    assert b.x.shape == (lp.n,) and b.y.shape == (lp.m,)


def _random_reducible(seed):
    """Random LPs made only of singleton rows (several per variable, both coefficient
    signs, equality / <= / >= / ranged), empty rows, fixed and free columns: presolve
    solves them completely, so no GPU is needed."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 9))
    lb = rng.uniform(-10, 0, n)
    ub = lb + rng.uniform(0, 20, n)
    fixed = rng.random(n) < 0.2
    ub[fixed] = lb[fixed]
    xs = lb + rng.random(n) * (ub - lb)  # a point the rows keep feasible, unless shifted below
    rows, lc, uc = [], [], []
    for _ in range(int(rng.integers(1, 3 * n))):
        j = int(rng.integers(0, n))
        a = float(rng.choice([-1, 1]) * rng.uniform(0.5, 3.0))
        kind = rng.integers(0, 4)
        v = a * xs[j] + (float(rng.uniform(1, 3)) if rng.random() < 0.1 else 0.0)
        d1, d2 = float(rng.uniform(0, 2)), float(rng.uniform(0, 2))
        lo, hi = {0: (v, v), 1: (-INF, v + d1), 2: (v - d1, INF), 3: (v - d1, v + d2)}[int(kind)]
        r = np.zeros(n)
        r[j] = a
        rows.append(r)
        lc.append(lo)
        uc.append(hi)
    for _ in range(int(rng.integers(0, 3))):  # empty rows that hold
        rows.append(np.zeros(n))
        lc.append(-float(rng.uniform(0, 1)))
        uc.append(float(rng.uniform(0, 1)))
    c = rng.normal(size=n)
    return lp_from(np.array(rows), c, lc, uc, lb, ub, maximize=bool(rng.random() < 0.3))


def test_presolve_random_reducible_lps_against_highs():
    from scipy.optimize import linprog

    solved = infeasible = 0
    for seed in range(300):
        lp = _random_reducible(seed)
        A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
        sgn = -1.0 if lp.maximize else 1.0
        fin_u, fin_l = np.isfinite(lp.con_ub), np.isfinite(lp.con_lb)
        A_ub = np.vstack([A[fin_u], -A[fin_l]]) if (fin_u.any() or fin_l.any()) else None
        b_ub = np.concatenate([lp.con_ub[fin_u], -lp.con_lb[fin_l]]) if A_ub is not None else None
        ref = linprog(sgn * lp.c, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(lp.var_lb, lp.var_ub)), method="highs")
        r = c_api.solve(lp, tol=1e-9, presolve=True)
        assert r.error == 0, (seed, r.error_string)
        if ref.status == 2:
            assert r.status == 2, seed
            infeasible += 1
            continue
        assert ref.status == 0 and r.status == 1, (seed, ref.status, r.status)
        assert abs(r.primal_objective - sgn * ref.fun) <= 1e-9 * (1 + abs(ref.fun)), seed
        assert kkt_ok(lp, r, tol=1e-9), seed
        solved += 1
    assert solved > 150 and infeasible > 10


def _random_activity_case(seed):
    """Small LPs with two- and three-variable rows and box bounds, about half of them
    made infeasible by a row whose bound lies beyond its activity range."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 7))
    lb = rng.uniform(-5, 0, n)
    ub = lb + rng.uniform(0.5, 5, n)
    rows, lc, uc = [], [], []
    for _ in range(int(rng.integers(1, 6))):
        r = np.zeros(n)
        cols = rng.choice(n, size=min(n, int(rng.integers(2, 4))), replace=False)
        r[cols] = rng.choice([-1, 1], size=len(cols)) * rng.uniform(0.5, 2.0, size=len(cols))
        amin = np.sum(np.where(r > 0, r * lb, r * ub))
        amax = np.sum(np.where(r > 0, r * ub, r * lb))
        if rng.random() < 0.5:  # reachable: a band inside [amin, amax]
            t = rng.uniform(amin, amax)
            lo, hi = t - rng.uniform(0, 1), t + rng.uniform(0, 1)
        elif rng.random() < 0.5:  # beyond the top of the activity range
            lo, hi = amax + rng.uniform(1e-3, 2), INF
        else:  # below the bottom
            lo, hi = -INF, amin - rng.uniform(1e-3, 2)
        rows.append(r)
        lc.append(lo)
        uc.append(hi)
    return lp_from(np.array(rows), rng.normal(size=n), lc, uc, lb, ub)


def test_presolve_infeasibility_is_sound_against_highs():
    # whenever presolve reports infeasibility (no GPU involved) HiGHS agrees
    from scipy.optimize import linprog

    detected = 0
    for seed in range(300):
        lp = _random_activity_case(seed)
        r = c_api.solve(lp, tol=1e-9, presolve=True)
        if r.status != 2:
            continue
        A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
        fin_u, fin_l = np.isfinite(lp.con_ub), np.isfinite(lp.con_lb)
        A_ub = np.vstack([A[fin_u], -A[fin_l]])
        b_ub = np.concatenate([lp.con_ub[fin_u], -lp.con_lb[fin_l]])
        ref = linprog(lp.c, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(lp.var_lb, lp.var_ub)), method="highs")
        assert ref.status == 2, seed
        detected += 1
    assert detected > 100


def test_redundant_row_and_dual_fixing_hand_worked():
    # min x0 - x1 + 2 x2, x in [0, 10]^4:
    #   r0: x0 + x1 + x2 <= 50   redundant: its activity is at most 30
    #   r1: x0 - x2 <= 4         x0 only loosens it downwards, cost 1 > 0: x0 = 0 (dual fixing);
    #                            then -x2 <= 4 is a singleton row, x2 empty at cost 2: x2 = 0
    #   r2: x1 + x3 >= 1         x1 only loosens it upwards, cost -1 < 0: x1 = 10; x3 (cost 0) too
    # Optimum -10 at (0, 10, 0, 10); every row slack, so y = 0 and the reduced costs are c.
    A = np.array([[1.0, 1.0, 1.0, 0.0], [1.0, 0.0, -1.0, 0.0], [0.0, 1.0, 0.0, 1.0]])
    lp = lp_from(A, [1.0, -1.0, 2.0, 0.0], [-INF, -INF, 1.0], [50.0, 4.0, INF], [0.0] * 4, [10.0] * 4)
    r = c_api.solve(lp, tol=1e-9, presolve=True)
    assert r.error == 0 and r.status == 1
    np.testing.assert_array_equal(r.x, [0.0, 10.0, 0.0, 10.0])
    np.testing.assert_array_equal(r.y, [0.0, 0.0, 0.0])
    np.testing.assert_array_equal(r.reduced_cost, lp.c)
    assert r.primal_objective == -10.0
    assert kkt_ok(lp, r)


def _one_sided_case(seed):
    """Small LPs of one-sided rows (<= or >=) over box bounds, with costs of either
    sign: dual fixing and redundant rows often settle them completely."""
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(2, 7)), int(rng.integers(1, 5))
    lb = rng.uniform(-3, 0, n)
    ub = lb + rng.uniform(0.5, 4, n)
    A = np.where(rng.random((m, n)) < 0.5, rng.choice([-1.0, 1.0], (m, n)) * rng.uniform(0.5, 2, (m, n)), 0.0)
    amin = np.sum(np.where(A > 0, A * lb, A * ub), axis=1)
    amax = np.sum(np.where(A > 0, A * ub, A * lb), axis=1)
    t = rng.uniform(amin - 1, amax + 1)
    le = rng.random(m) < 0.5
    lc = np.where(le, -INF, t)
    uc = np.where(le, t, INF)
    return lp_from(A, rng.normal(size=n), lc, uc, lb, ub, maximize=bool(rng.random() < 0.3))


def test_presolve_dual_fixing_and_redundant_rows_against_highs():
    # LPs that presolve settles without the GPU are checked against HiGHS: status,
    # objective, and optimality of the postsolved (x, y, reduced costs) in numpy.
    from scipy.optimize import linprog

    settled = 0
    for seed in range(400):
        lp = _one_sided_case(seed)
        r = c_api.solve(lp, tol=1e-9, presolve=True)
        if r.error != 0:  # left a reduced problem for the GPU
            continue
        A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
        sgn = -1.0 if lp.maximize else 1.0
        fin_u, fin_l = np.isfinite(lp.con_ub), np.isfinite(lp.con_lb)
        A_ub = np.vstack([A[fin_u], -A[fin_l]])
        b_ub = np.concatenate([lp.con_ub[fin_u], -lp.con_lb[fin_l]])
        ref = linprog(sgn * lp.c, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(lp.var_lb, lp.var_ub)), method="highs")
        if ref.status == 2:
            assert r.status == 2, seed
        else:
            assert ref.status == 0 and r.status == 1, (seed, ref.status, r.status)
            assert abs(r.primal_objective - sgn * ref.fun) <= 1e-9 * (1 + abs(ref.fun)), seed
            assert kkt_ok(lp, r, tol=1e-9), seed
        settled += 1
    assert settled > 120


def test_free_column_singleton_hand_worked():
    # min x0 - 3 x1 + s, x0, x1 in [0, 10], s free:
    #   r0: x0 + x1 - s = 2   s appears only here and is free: s = x0 + x1 - 2
    # Substituted, the objective is 2 x0 - 2 x1 - 2 (cost 1 of s moved onto x0 and x1 as
    # -1 * a_i / a = +1 each, offset 1 * 2 / -1 = -2), so x0 = 0, x1 = 10, s = 8, objective -22.
    # The row's dual is c_s / a = -1, and z = c - A^T y = (2, -2, 0).
    A = np.array([[1.0, 1.0, -1.0]])
    lp = lp_from(A, [1.0, -3.0, 1.0], [2.0], [2.0], [0.0, 0.0, -INF], [10.0, 10.0, INF])
    r = c_api.solve(lp, tol=1e-9, presolve=True)
    assert r.error == 0 and r.status == 1
    np.testing.assert_array_equal(r.x, [0.0, 10.0, 8.0])
    np.testing.assert_array_equal(r.y, [-1.0])
    np.testing.assert_array_equal(r.reduced_cost, [2.0, -2.0, 0.0])
    assert r.primal_objective == -22.0
    assert kkt_ok(lp, r)


def test_implied_free_singleton_in_an_inequality_row_hand_worked():
    # max 2 x0 + x1 - s with x0, x1 in [0, 1], s in [-5, 5] (implied free: from
    # r0, s = x0 + x1 - t with t in [-1, 1] stays in [-1, 3]):
    #   r0: -1 <= x0 + x1 - s <= 1
    # Maximising -s pushes the row activity up to 1 (c_s / a = (-(-1)) / -1 < 0 in the
    # minimisation sense takes the upper bound): s = x0 + x1 - 1 and the objective is
    # x0 + 0 x1 + 1, so x0 = 1; x1 (cost 0) goes to its projection of 0 and s = 0.
    A = np.array([[1.0, 1.0, -1.0]])
    lp = lp_from(A, [2.0, 1.0, -1.0], [-1.0], [1.0], [0.0, 0.0, -5.0], [1.0, 1.0, 5.0], maximize=True)
    r = c_api.solve(lp, tol=1e-9, presolve=True)
    assert r.error == 0 and r.status == 1
    np.testing.assert_array_equal(r.x, [1.0, 0.0, 0.0])
    assert r.primal_objective == 2.0
    np.testing.assert_array_equal(r.y, [1.0])  # c_s / a = -1 / -1
    assert kkt_ok(lp, r)


def _singleton_column_case(seed):
    """Rows that each own a (implied) free column singleton -- the slack-like columns of
    modelled LPs -- over box-bounded shared columns (some fixed), equality or ranged rows,
    costs of either sign: presolve substitutes every row away and the shared columns end
    empty."""
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(1, 6)), int(rng.integers(1, 5))
    lb = rng.uniform(-3, 0, n)
    ub = lb + rng.uniform(0.5, 4, n)
    fixed = rng.random(n) < 0.25  # substituted before the rows are solved for their slacks
    ub[fixed] = lb[fixed]
    A = np.where(rng.random((m, n)) < 0.6, rng.normal(size=(m, n)), 0.0)
    S = np.diag(rng.choice([-1.0, 1.0], m) * rng.uniform(0.5, 2.0, m))  # one private column per row
    lc, uc, slb, sub = [], [], [], []
    for i in range(m):
        t = rng.uniform(-2, 2)
        if rng.random() < 0.5:
            lc.append(t)
            uc.append(t)
        else:
            lc.append(t - rng.uniform(0.1, 2))
            uc.append(t + rng.uniform(0.1, 2))
        if rng.random() < 0.5:  # free
            slb.append(-INF)
            sub.append(INF)
        else:  # wide enough to be implied by the row
            slb.append(-1e3)
            sub.append(1e3)
    full = np.hstack([A, S])
    c = rng.normal(size=n + m)
    c[n:][rng.random(m) < 0.2] = 0.0
    return lp_from(full, c, lc, uc, np.concatenate([lb, slb]), np.concatenate([ub, sub]),
                   maximize=bool(rng.random() < 0.3))


def test_free_column_singletons_against_highs():
    # LPs that presolve settles with column-singleton substitutions (no GPU): status,
    # objective and optimality of the postsolved (x, y, reduced costs) against HiGHS.
    from scipy.optimize import linprog

    settled = 0
    for seed in range(400):
        lp = _singleton_column_case(seed)
        r = c_api.solve(lp, tol=1e-9, presolve=True)
        if r.error != 0:  # left a reduced problem for the GPU
            continue
        A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
        sgn = -1.0 if lp.maximize else 1.0
        eq = lp.con_lb == lp.con_ub
        fin_u, fin_l = np.isfinite(lp.con_ub) & ~eq, np.isfinite(lp.con_lb) & ~eq
        A_ub = np.vstack([A[fin_u], -A[fin_l]])
        b_ub = np.concatenate([lp.con_ub[fin_u], -lp.con_lb[fin_l]])
        ref = linprog(sgn * lp.c, A_ub=A_ub if len(b_ub) else None, b_ub=b_ub if len(b_ub) else None,
                      A_eq=A[eq] if eq.any() else None, b_eq=lp.con_lb[eq] if eq.any() else None,
                      bounds=list(zip(lp.var_lb, lp.var_ub)), method="highs")
        if ref.status == 2:
            assert r.status == 2, seed
        else:
            assert ref.status == 0 and r.status == 1, (seed, ref.status, r.status)
            assert abs(r.primal_objective - sgn * ref.fun) <= 1e-9 * (1 + abs(ref.fun)), seed
            assert kkt_ok(lp, r, tol=1e-9), seed
        settled += 1
    assert settled > 300


def test_singleton_row_after_a_cost_transfer_hand_worked():
    # min -2 x0 + 0.5 x1 + s, x0, x1 in [0, 10], s free:
    #   r0: x0 + x1 - s = 2    s is a free column singleton: s = x0 + x1 - 2, its cost 1
    #                          moves onto x0 and x1 (costs -1 and 1.5, offset -2)
    #   r1: x0 + x1 >= 13      then x0 (cost -1, only loosens r1 upwards) is fixed at 10
    #                          by dual fixing, r1 becomes the singleton x1 >= 3, and x1
    #                          (cost 1.5) takes that bound: x = (10, 3, 11), objective -7.5.
    # Duals: r1 holds x1 at the bound it supplied, so y_1 is x1's reduced cost in the problem
    # as it stood then, 1.5 (its cost after the transfer), and y_0 = c_s / a = -1; z = (-2.5, 0, 0).
    A = np.array([[1.0, 1.0, -1.0], [1.0, 1.0, 0.0]])
    lp = lp_from(A, [-2.0, 0.5, 1.0], [2.0, 13.0], [2.0, INF], [0.0, 0.0, -INF], [10.0, 10.0, INF])
    r = c_api.solve(lp, tol=1e-9, presolve=True)
    assert r.error == 0 and r.status == 1
    np.testing.assert_allclose(r.x, [10.0, 3.0, 11.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(r.y, [-1.0, 1.5], rtol=0, atol=1e-12)
    np.testing.assert_allclose(r.reduced_cost, [-2.5, 0.0, 0.0], rtol=0, atol=1e-12)
    assert abs(r.primal_objective - (-7.5)) < 1e-12
    assert kkt_ok(lp, r)


def test_column_singleton_row_with_a_fixed_column_hand_worked():
    # The hand-worked free-column-singleton LP with a fixed column x3 = 3 added to its row
    # (and 3 to the row's bounds): x3 is substituted first, so the row reads x0 + x1 - s = 2
    # when s is solved from it; postsolve must not count x3 twice.  Same optimum as before:
    # x = (0, 10, 8, 3), objective -22 + 0 * x3, y = -1.
    A = np.array([[1.0, 1.0, -1.0, 1.0]])
    lp = lp_from(A, [1.0, -3.0, 1.0, 0.0], [5.0], [5.0], [0.0, 0.0, -INF, 3.0], [10.0, 10.0, INF, 3.0])
    r = c_api.solve(lp, tol=1e-9, presolve=True)
    assert r.error == 0 and r.status == 1
    np.testing.assert_array_equal(r.x, [0.0, 10.0, 8.0, 3.0])
    np.testing.assert_array_equal(r.y, [-1.0])
    assert r.primal_objective == -22.0
    assert kkt_ok(lp, r)


@pytest.mark.parametrize("fname,status", [("good-mps-fixed-var.mps", 2), ("good-mps-free-var.mps", 3),
                                          ("good-mps-lower-bound-inf-var.mps", 3)])
def test_reference_parser_fixtures_settled_by_presolve(fname, status):
    # Fixtures of the reference parser corpus that HiGHS finds infeasible (2) or unbounded
    # (3): presolve reaches the same verdict with no GPU.
    import os

    lp = c_api.read_lp(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mps", fname))
    r = c_api.solve(lp, tol=1e-8, presolve=True, time_limit=5.0)
    assert r.error == 0 and r.status == status
