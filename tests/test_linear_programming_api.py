"""The reference's Python LP interface over the C ABI (cuopt_amd.linear_programming).

CPU part: data model to C problem and back, MPS parsing with names, settings by
name; GPU part: Solve / BatchSolve semantics against known optima (afiro
-464.7531, test_lp_solver.py:92-111; good-max 17, pdlp_test.cu:988-1026) and the
closed-form optimum of the generated LPs."""

import os

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd import generators as G
from cuopt_amd import linear_programming as LPM

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _model(lp, ranged=False):
    dm = LPM.DataModel()
    dm.set_csr_constraint_matrix(lp.values, lp.col_indices, lp.row_offsets)
    dm.set_objective_coefficients(lp.c)
    dm.set_variable_lower_bounds(lp.var_lb)
    dm.set_variable_upper_bounds(lp.var_ub)
    dm.set_maximize(lp.maximize)
    dm.set_objective_offset(lp.offset)
    if ranged:
        dm.set_constraint_lower_bounds(lp.con_lb)
        dm.set_constraint_upper_bounds(lp.con_ub)
    else:
        s, rhs = lp.row_sense()
        dm.set_row_types(np.array(list(s.tobytes().decode())))
        dm.set_constraint_bounds(rhs)
    return dm


def _roundtrip(dm):
    h = LPM._to_problem(dm)
    try:
        return c_api.problem_to_lp(h)
    finally:
        c_api.lib().cuOptDestroyProblem(__import__("ctypes").byref(h))


@pytest.mark.parametrize("ranged", [False, True])
def test_data_model_reaches_the_c_problem(ranged):
    lp = G.skewed_lp(m=300, n=700, seed=2)
    back = _roundtrip(_model(lp, ranged))
    assert np.array_equal(back.row_offsets, lp.row_offsets) and np.array_equal(back.col_indices, lp.col_indices)
    assert np.array_equal(back.values, lp.values) and np.array_equal(back.c, lp.c)
    assert np.array_equal(back.con_lb, lp.con_lb) and np.array_equal(back.con_ub, lp.con_ub)
    assert np.array_equal(back.var_ub, lp.var_ub)


def test_default_variable_bounds_are_nonnegative_reals():
    # problem_helpers.cuh:54-79: unset bounds are [0, +inf)
    dm = LPM.DataModel()
    dm.set_csr_constraint_matrix(np.array([3.0, 4.0, 2.7, 10.1]), np.array([0, 1, 0, 1]), np.array([0, 2, 4]))
    dm.set_constraint_bounds(np.array([5.4, 4.9]))
    dm.set_objective_coefficients(np.array([0.2, 0.1]))
    dm.set_row_types(np.array(["L", "L"]))
    back = _roundtrip(dm)
    assert np.array_equal(back.var_lb, [0.0, 0.0]) and np.all(np.isinf(back.var_ub))


def test_inconsistent_models_raise():
    dm = LPM.DataModel()
    with pytest.raises(ValueError):
        LPM._to_problem(dm)  # no matrix
    dm.set_csr_constraint_matrix(np.array([1.0]), np.array([0]), np.array([0, 1]))
    dm.set_objective_coefficients(np.array([1.0]))
    with pytest.raises(ValueError):
        LPM._to_problem(dm)  # no row types / bounds
    dm.set_row_types(np.array(["E"]))
    dm.set_constraint_bounds(np.array([1.0]))
    dm.set_quadratic_objective_matrix(np.array([1.0]), np.array([0]), np.array([0, 1]))
    with pytest.raises(ValueError, match="barrier"):
        LPM._to_problem(dm)
    with pytest.raises(ValueError):
        dm.set_csr_constraint_matrix(np.array(["a"]), np.array([0]), np.array([0, 1]))


def _validity_steps():
    """The model of test_lp_solver.py:309-368, completed step by step: every model
    before the row types are set is incomplete."""
    dm = LPM.DataModel()
    yield dm  # nothing set
    dm.set_csr_constraint_matrix(np.array([1.0]), np.array([0], np.int32), np.array([0, 1], np.int32))
    yield dm  # no b
    dm.set_constraint_bounds(np.array([1.0]))
    yield dm  # no c
    dm.set_objective_coefficients(np.array([1.0]))
    dm.set_maximize(True)
    yield dm  # no row types
    dm.set_maximize(False)
    yield dm  # still no row types


def test_incomplete_models_give_validation_error_solutions():
    # test_lp_solver.py:309-346: Solve reports incomplete data as a Solution with
    # ErrorStatus.ValidationError instead of raising (no device is touched)
    for dm in _validity_steps():
        sol = LPM.Solve(dm)
        assert sol.get_error_status() == LPM.ErrorStatus.ValidationError
        assert sol.get_termination_status() == LPM.LPTerminationStatus.NoTermination
        assert sol.get_error_message()


def _named_model():
    dm = LPM.DataModel()
    dm.set_csr_constraint_matrix(np.array([1.0, 2.0, 3.0]), np.array([0, 1, 1], np.int32),
                                 np.array([0, 2, 3], np.int32))
    dm.set_objective_coefficients(np.array([1.0, -1.0]))
    dm.set_row_types(np.array(["L", "G"]))
    dm.set_constraint_bounds(np.array([4.0, 1.0]))
    dm.set_variable_types(np.array(["I", "C"]))
    dm.set_variable_names(["x", "y"])
    dm.set_row_names(["cap", "need"])
    dm.set_objective_name("profit")
    dm.set_problem_name("tiny")
    return dm


def test_names_reach_the_written_mps(tmp_path):
    # the data model's names reach the C problem (cuOptAmdSet*Names) and the
    # written MPS, which reads back with them; an integer column without bounds
    # is written with an explicit bound so that it does not read back as binary
    f = tmp_path / "named.mps"
    _named_model().writeMPS(str(f))
    back = LPM.ParseMps(str(f))
    assert list(back.get_variable_names()) == ["x", "y"]
    assert list(back.get_row_names()) == ["cap", "need"]
    assert back.get_objective_name() == "profit" and back.get_problem_name() == "tiny"
    assert list(back.get_variable_types()) == ["I", "C"]
    assert back.get_variable_upper_bounds().tolist() == [np.inf, np.inf]
    assert rows_equal(back, _named_model())


def rows_equal(a, b):
    return (np.array_equal(a.get_constraint_matrix_offsets(), b.get_constraint_matrix_offsets())
            and np.array_equal(a.get_constraint_matrix_indices(), b.get_constraint_matrix_indices())
            and np.array_equal(a.get_constraint_matrix_values(), b.get_constraint_matrix_values()))


def test_unreadable_names_fall_back_to_generated_ones(tmp_path):
    dm = _named_model()
    dm.set_variable_names(["x 1", "y"])  # a blank: the columns get C0, C1
    dm.set_row_names(["cap", "cap"])     # not unique: the rows get R0, R1
    f = tmp_path / "gen.mps"
    dm.writeMPS(str(f))
    back = LPM.ParseMps(str(f))
    assert list(back.get_variable_names()) == ["C0", "C1"]
    assert list(back.get_row_names()) == ["R0", "R1"]
    assert rows_equal(back, dm)


def test_name_count_must_match():
    dm = _named_model()
    dm.set_variable_types(np.array(["C", "C"]))
    dm.set_variable_names(["x"])
    with pytest.raises(ValueError):
        LPM._to_problem(dm)
    sol = LPM.Solve(dm)
    assert sol.get_error_status() == LPM.ErrorStatus.ValidationError


@pytest.mark.parametrize("fname", ["afiro_original.mps", "good-mps-fixed-ranges.mps", "lp_model_with_var_bounds.mps"])
def test_parse_mps_matches_the_c_problem_and_keeps_names(fname):
    # the data model carries what cuOptReadProblem's getters report
    path = os.path.join(GOLDEN, fname)
    dm = LPM.ParseMps(path)
    ref = c_api.read_lp(path)
    assert np.array_equal(dm.get_constraint_matrix_offsets(), ref.row_offsets)
    assert np.array_equal(dm.get_constraint_matrix_values(), ref.values)
    assert np.array_equal(dm.get_objective_coefficients(), ref.c)
    assert np.array_equal(dm.get_constraint_lower_bounds(), ref.con_lb)
    assert np.array_equal(dm.get_constraint_upper_bounds(), ref.con_ub)
    text = open(path).read().split()
    assert len(dm.get_variable_names()) == ref.n and all(v in text for v in dm.get_variable_names())
    assert len(dm.get_row_names()) == ref.m and all(r in text for r in dm.get_row_names())


def test_parse_mps_afiro_names():
    dm = LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps"))
    assert dm.get_problem_name() == "AFIRO" and dm.get_objective_name() == "COST"
    assert dm.get_row_names()[:3] == ["R09", "R10", "X05"]
    assert dm.get_variable_names()[0] == "X01"
    assert set(dm.get_variable_types()) == {"C"}
    assert "".join(dm.get_row_types()[:3]) == "ELL" or len(dm.get_row_types()) == 27


def test_solver_settings_by_name():
    s = LPM.SolverSettings()
    assert s.get_parameter("absolute_primal_tolerance") == 1e-4
    assert s.get_parameter("log_to_console") in (True, False)
    with pytest.raises(ValueError):
        s.set_parameter("no_such_parameter", 1)
    with pytest.raises(ValueError):
        s.get_parameter("no_such_parameter")
    s.set_optimality_tolerance(1e-7)
    d = s.toDict()  # get_parameter also records what it read, as the reference's does
    tols = [k for k in d if k.endswith("_tolerance")]
    assert len(tols) == 6 and all(d[k] == 1e-7 for k in tols) and d["log_to_console"] in (True, False)
    s.set_parameter("method", LPM.SolverMethod.PDLP)
    s.set_parameter("pdlp_solver_mode", LPM.PDLPSolverMode.Stable3)
    s.set_parameter("log_to_console", False)
    h = s._handle()
    v = __import__("ctypes").c_double()
    assert h.get_string("method") == "1"
    assert c_api.lib().cuOptGetFloatParameter(h.handle, b"relative_gap_tolerance", __import__("ctypes").byref(v)) == 0
    assert v.value == 1e-7
    s.set_parameter("iteration_limit", -3)
    with pytest.raises(ValueError):
        s._handle()
    # PDLP warm-start data: Stable3 returns NoTermination (pdlp_test.cu:880-900), batch raises
    s.set_parameter("iteration_limit", 100)
    s.set_pdlp_warm_start_data(object())
    sol = LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps")), s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.NoTermination
    with pytest.raises(ValueError):
        LPM.BatchSolve([LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps"))], s)


def test_enums_follow_the_c_constants():
    # constants.h:83-93 and the names c_api reports for them
    for st in LPM.LPTerminationStatus:
        if st.value in c_api.STATUS:
            assert c_api.STATUS[st.value] == st.name
    assert LPM.LPTerminationStatus.Optimal == 1 and LPM.LPTerminationStatus.NumericalError == 6
    assert LPM.LPTerminationStatus.PrimalFeasible == 7
    assert LPM.ErrorStatus.RuntimeError == c_api.CUOPT_RUNTIME_ERROR
    assert str(LPM.SolverMethod.Barrier) == "3" and str(LPM.PDLPSolverMode.Stable3) == "4"


def test_solve_reports_the_missing_device():
    if c_api.device_count() > 0:
        pytest.skip("a GPU is present")
    lp = G.random_lp(m=20, n=50, seed=5)
    sol = LPM.Solve(_model(lp))
    assert sol.get_error_status() == LPM.ErrorStatus.RuntimeError and "HIP" in sol.get_error_message()


# ---- GPU ------------------------------------------------------------------------

gpu = pytest.mark.gpu


@gpu
def test_solve_afiro_and_good_max():
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-8)
    s.set_parameter("method", LPM.SolverMethod.PDLP)
    s.set_parameter("time_limit", 120.0)
    sol = LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps")), s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert sol.get_termination_reason() == "Optimal" and sol.get_error_status() == LPM.ErrorStatus.Success
    assert abs(sol.get_primal_objective() - (-464.7531)) < 1e-3
    assert set(sol.get_vars()) >= {"X01", "X02"} and len(sol.get_vars()) == 32
    st = sol.get_lp_stats()
    assert st["nb_iterations"] > 0 and st["primal_residual"] >= 0.0
    assert sol.get_solved_by_pdlp() and sol.get_problem_category() == LPM.ProblemCategory.LP
    with pytest.raises(AttributeError):
        sol.get_milp_stats()
    sol = LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "good-max.mps")), s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert abs(sol.get_primal_objective() - 17.0) < 1e-6


@gpu
def test_solve_row_types_and_ranged_forms_agree_and_scaling_factor():
    lp = G.random_lp(m=150, n=375, seed=9)
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-8)
    s.set_parameter("time_limit", 120.0)
    a = LPM.Solve(_model(lp, ranged=False), s)
    dm = _model(lp, ranged=True)
    dm.set_objective_scaling_factor(2.0)
    b = LPM.Solve(dm, s)
    assert a.get_termination_status() == b.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert abs(a.get_primal_objective() - lp.obj_opt) < 1e-6 * (1 + abs(lp.obj_opt))
    assert abs(b.get_primal_objective() - 2.0 * a.get_primal_objective()) < 1e-9 * (1 + abs(lp.obj_opt))
    assert a.get_vars() == {}  # no names given


@gpu
def test_batch_solve_returns_solutions_and_batch_time():
    lps = [G.random_lp(m=m, n=int(2.5 * m), seed=20 + m) for m in (60, 90, 140)]
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-6)
    s.set_parameter("time_limit", 120.0)
    sols, t = LPM.BatchSolve([_model(lp) for lp in lps], s)
    assert len(sols) == 3 and t > 0.0
    for lp, sol in zip(lps, sols):
        assert sol.get_termination_reason() == "Optimal"
        assert abs(sol.get_primal_objective() - lp.obj_opt) < 1e-4 * (1 + abs(lp.obj_opt))
        assert sol.get_solve_time() == t


@gpu
def test_completed_model_solves():
    # test_lp_solver.py:348-368: with row types (and then ranged bounds) the model solves
    *_, dm = _validity_steps()
    dm.set_row_types(np.array(["E"]))
    sol = LPM.Solve(dm)
    assert sol.get_error_status() == LPM.ErrorStatus.Success
    assert sol.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert abs(sol.get_primal_objective() - 1.0) < 1e-6
    dm.set_constraint_lower_bounds(np.array([1.0]))
    dm.set_constraint_upper_bounds(np.array([1.0]))
    assert LPM.Solve(dm).get_error_status() == LPM.ErrorStatus.Success


@gpu
def test_user_problem_file_and_solution_file(tmp_path):
    # test_lp_solver.py:691-722: user_problem_file writes the problem before the
    # solve (solve.cu:1281-1284), solution_file the .sol after it (solve.cu:1327-1330)
    afiro = LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps"))
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-8)
    s.set_parameter("time_limit", 120.0)
    out = tmp_path / "afiro_out.mps"
    s.set_parameter("user_problem_file", str(out))
    LPM.Solve(afiro, s)
    assert out.is_file()
    again = LPM.ParseMps(str(out))
    np.testing.assert_array_equal(again.get_objective_coefficients(), afiro.get_objective_coefficients())
    s.set_parameter("user_problem_file", "")
    sol_file = tmp_path / "afiro.sol"
    s.set_parameter("solution_file", str(sol_file))
    sol = LPM.Solve(afiro, s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert abs(sol.get_primal_objective() - (-464.7531)) < 1e-3
    lines = sol_file.read_text().splitlines()
    assert lines[0] == "# Status: Optimal"
    assert abs(float(lines[1].split(":")[1]) - sol.get_primal_objective()) < 1e-9
    values = dict(line.split() for line in lines[2:])
    assert len(values) == 32 and abs(float(values["X01"]) - 80.0) < 1e-3


@gpu
def test_batch_solve_reports_an_incomplete_model_in_place():
    lps = [G.random_lp(m=m, n=int(2.5 * m), seed=30 + m) for m in (60, 90)]
    s = LPM.SolverSettings()
    s.set_parameter("time_limit", 120.0)
    sols, t = LPM.BatchSolve([_model(lps[0]), LPM.DataModel(), _model(lps[1])], s)
    assert len(sols) == 3 and t > 0.0
    assert sols[1].get_error_status() == LPM.ErrorStatus.ValidationError
    for lp, sol in zip(lps, (sols[0], sols[2])):
        assert sol.get_termination_reason() == "Optimal"
        assert abs(sol.get_primal_objective() - lp.obj_opt) < 1e-3 * (1 + abs(lp.obj_opt))


@gpu
@pytest.mark.parametrize("presolve", [0, 1])
def test_tiny_lp_of_the_reference_solver_test(presolve):
    # test_lp_solver.py:46-79: min x s.t. x <= 1 twice, x >= 0.  The reference runs
    # it in Stable2 ("Stable3 infinite-loops on this sample trivial problem"); here
    # Stable3 reaches the optimum x = 0.
    dm = LPM.DataModel()
    dm.set_csr_constraint_matrix(np.array([1.0, 1.0]), np.array([0, 0]), np.array([0, 1, 2]))
    dm.set_constraint_bounds(np.array([1.0, 1.0]))
    dm.set_objective_coefficients(np.array([1.0]))
    dm.set_row_types(np.array(["L", "L"]))
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-2)
    s.set_parameter("method", LPM.SolverMethod.PDLP)
    s.set_parameter("presolve", presolve)
    s.set_parameter("time_limit", 60.0)
    sol = LPM.Solve(dm, s)
    assert sol.get_termination_reason() == "Optimal"
    assert sol.get_primal_solution()[0] == pytest.approx(0.0)
    assert sol.get_primal_objective() == pytest.approx(0.0) and sol.get_dual_objective() == pytest.approx(0.0)
    st = sol.get_lp_stats()
    assert st["primal_residual"] == pytest.approx(0.0) and st["dual_residual"] == pytest.approx(0.0)
    assert st["gap"] == pytest.approx(0.0)
    assert sol.get_solved_by_pdlp()


@gpu
def test_parser_and_solver():
    # test_lp_solver.py:82-89
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-2)
    s.set_parameter("time_limit", 60.0)
    sol = LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "mps", "good-mps-1.mps")), s)
    assert sol.get_termination_reason() == "Optimal"


@gpu
def test_very_low_tolerance_afiro():
    # test_lp_solver.py:92-112 (afiro to 1e-12)
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-12)
    s.set_parameter("time_limit", 120.0)
    sol = LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps")), s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert sol.get_primal_objective() == pytest.approx(-464.7531)


@gpu
def test_iteration_limit_returns_the_iterate():
    # test_lp_solver.py:115-133 on afiro (savsched1 is a downloaded dataset): the
    # iteration limit wins over a large time limit, and the returned point is the
    # iterate, not zeros
    s = LPM.SolverSettings()
    s.set_optimality_tolerance(1e-12)
    s.set_parameter("iteration_limit", 1)
    s.set_parameter("time_limit", 99999999)
    sol = LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps")), s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.IterationLimit
    assert sol.get_primal_objective() != 0.0
    assert np.any(sol.get_primal_solution())


@gpu
def test_log_file_gets_the_status_line(tmp_path):
    # log_file (solve.cu:1136): the final status line goes to the file as well
    log = tmp_path / "solve.log"
    s = LPM.SolverSettings()
    s.set_parameter("time_limit", 60.0)
    s.set_parameter("log_file", str(log))
    LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps")), s)
    LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "good-max.mps")), s)
    lines = log.read_text().splitlines()
    assert len(lines) == 2 and all(line.startswith("Status: Optimal") for line in lines)


def test_reference_solver_settings_sequence():
    # test_lp_solver.py:254-308, step by step (the out-of-range solver mode is caught
    # when Solve applies the settings, before any device work)
    s = LPM.SolverSettings()
    names = ["absolute_dual_tolerance", "relative_dual_tolerance", "absolute_primal_tolerance",
             "relative_primal_tolerance", "absolute_gap_tolerance", "relative_gap_tolerance",
             "primal_infeasible_tolerance", "dual_infeasible_tolerance"]
    for name in names:
        s.set_parameter(name, 1e-5)
    assert all(s.get_parameter(name) == 1e-5 for name in names)
    assert s.get_parameter("time_limit") == float("inf")
    s.set_parameter("iteration_limit", 10)
    s.set_parameter("time_limit", 10.2)
    assert s.get_parameter("iteration_limit") == 10 and s.get_parameter("time_limit") == 10.2
    s.set_parameter("infeasibility_detection", False)
    assert not s.get_parameter("infeasibility_detection")
    assert s.get_parameter("pdlp_solver_mode") == int(LPM.PDLPSolverMode.Stable3)
    with pytest.raises(ValueError):
        s.set_parameter("pdlp_solver_mode", 10)
        LPM.Solve(LPM.ParseMps(os.path.join(GOLDEN, "mps", "good-mps-1.mps")), s)
    s.set_parameter("pdlp_solver_mode", LPM.PDLPSolverMode.Methodical1)
    assert s.get_parameter("pdlp_solver_mode") == int(LPM.PDLPSolverMode.Methodical1)


@gpu
def test_dual_postsolve_off_gives_an_empty_dual_solution():
    # pdlp_test.cu:903-930: with presolve on and dual_postsolve off the solution has no duals
    dm = LPM.ParseMps(os.path.join(GOLDEN, "afiro_original.mps"))
    s = LPM.SolverSettings()
    s.set_parameter("time_limit", 60.0)
    sol = LPM.Solve(dm, s)
    assert len(sol.get_dual_solution()) == 27
    s.set_parameter("dual_postsolve", False)
    sol = LPM.Solve(dm, s)
    assert sol.get_termination_status() == LPM.LPTerminationStatus.Optimal
    assert len(sol.get_dual_solution()) == 0 and len(sol.get_primal_solution()) == 32
