"""C-ABI boundary tests that need no GPU: exports, parameter registry, problem
handles, argument validation and the host-side early exits of cuOptSolve.

Expectations follow the reference C API tests
(cpp/tests/linear_programming/c_api_tests/c_api_test.c, c_api_tests.cpp:21-261)
and the registry ranges of cpp/src/math_optimization/solver_settings.cu:61-132.
"""

import ctypes
import os
import re

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd import generators as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = []
    for h in ("cuopt_c.h", "cuopt_amd_ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:cuopt_int_t|void|int8_t)\s+(cuOpt\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = c_api.lib()
    names = declared_functions()
    assert len(names) >= 45
    for n in names:
        assert hasattr(L, n), n


def test_type_sizes_and_version():
    L = c_api.lib()
    assert L.cuOptGetFloatSize() == 8  # c_api_tests.cpp:23
    assert L.cuOptGetIntSize() == 4
    a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert L.cuOptGetVersion(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert L.cuOptGetVersion(None, ctypes.byref(b), ctypes.byref(c)) == c_api.CUOPT_INVALID_ARGUMENT


def test_parameter_defaults_and_ranges():
    s = c_api.SolverSettings()
    L = c_api.lib()
    v = ctypes.c_double()
    assert L.cuOptGetFloatParameter(s.handle, b"absolute_primal_tolerance", ctypes.byref(v)) == 0
    assert v.value == 1e-4
    assert s.set("relative_gap_tolerance", 1e-8) == 0
    assert L.cuOptGetFloatParameter(s.handle, b"relative_gap_tolerance", ctypes.byref(v)) == 0 and v.value == 1e-8
    assert s.set("relative_gap_tolerance", 0.5) == c_api.CUOPT_INVALID_ARGUMENT  # range [0, 1e-1]
    assert s.set("no_such_parameter", 3) == c_api.CUOPT_INVALID_ARGUMENT
    i = ctypes.c_int32()
    assert L.cuOptGetIntegerParameter(s.handle, b"method", ctypes.byref(i)) == 0 and i.value == 0
    assert L.cuOptGetIntegerParameter(s.handle, b"pdlp_solver_mode", ctypes.byref(i)) == 0 and i.value == 4
    assert s.set("pdlp_solver_mode", 7) == c_api.CUOPT_INVALID_ARGUMENT
    assert s.set("iteration_limit", -1) == c_api.CUOPT_INVALID_ARGUMENT
    # a boolean through the integer setter (cuopt_c.cpp:682-690)
    assert s.set("infeasibility_detection", 1) == 0
    assert L.cuOptGetIntegerParameter(s.handle, b"infeasibility_detection", ctypes.byref(i)) == 0 and i.value == 1
    # string forms
    assert L.cuOptSetParameter(s.handle, b"time_limit", b"12.5") == 0
    assert s.get_string("time_limit").startswith("12.5")
    assert L.cuOptSetParameter(s.handle, b"log_to_console", b"false") == 0
    assert s.get_string("log_to_console") == "false"
    assert L.cuOptSetParameter(s.handle, b"log_to_console", b"maybe") == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptSetParameter(s.handle, b"iteration_limit", b"12x") == c_api.CUOPT_INVALID_ARGUMENT
    # MIP-only names are accepted so that existing callers keep working
    assert s.set("mip_relative_gap", 1e-3) == 0
    assert s.set("pdlp_precision", 1) == 0 and s.set("pdlp_precision", 2) == c_api.CUOPT_INVALID_ARGUMENT


def test_null_arguments_are_rejected():
    L = c_api.lib()
    assert L.cuOptCreateSolverSettings(None) == c_api.CUOPT_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    assert L.cuOptCreateProblem(1, 1, 1, 0.0, None, None, None, None, None, None, None, None, None,
                                ctypes.byref(h)) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptSolve(None, None, None) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptSetParameter(None, b"method", b"1") == c_api.CUOPT_INVALID_ARGUMENT


def test_problem_round_trip_getters():
    lp = G.random_lp(m=30, n=70, seed=1)
    p = c_api.Problem(lp)
    L = c_api.lib()
    i = ctypes.c_int32()
    assert L.cuOptGetNumConstraints(p.handle, ctypes.byref(i)) == 0 and i.value == 30
    assert L.cuOptGetNumVariables(p.handle, ctypes.byref(i)) == 0 and i.value == 70
    assert L.cuOptGetNumNonZeros(p.handle, ctypes.byref(i)) == 0 and i.value == lp.nnz
    ptr, idx, val = np.zeros(31, np.int32), np.zeros(lp.nnz, np.int32), np.zeros(lp.nnz)
    assert L.cuOptGetConstraintMatrix(p.handle, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data) == 0
    assert np.array_equal(ptr, lp.row_offsets) and np.array_equal(idx, lp.col_indices)
    assert np.array_equal(val, lp.values)
    c = np.zeros(70)
    assert L.cuOptGetObjectiveCoefficients(p.handle, c.ctypes.data) == 0 and np.array_equal(c, lp.c)
    sense = np.zeros(30, np.uint8)
    assert L.cuOptGetConstraintSense(p.handle, sense.ctypes.data) == 0
    assert bytes(sense) == b"E" * 30
    ub = np.zeros(70)
    assert L.cuOptGetVariableUpperBounds(p.handle, ub.ctypes.data) == 0 and np.all(ub == 10.0)
    mip = ctypes.c_int32()
    assert L.cuOptIsMIP(p.handle, ctypes.byref(mip)) == 0 and mip.value == 0


def test_malformed_csr_is_rejected():
    lp = G.random_lp(m=10, n=20, seed=2)
    bad = lp.col_indices.copy()
    bad[3] = 20  # column out of range
    lp2 = G.LP(**{**lp.__dict__, "col_indices": bad})
    with pytest.raises(c_api.CuOptError) as e:
        c_api.Problem(lp2)
    assert e.value.code == c_api.CUOPT_INVALID_ARGUMENT
    ptr = lp.row_offsets.copy()
    ptr[5] = ptr[6] + 1  # decreasing offsets
    lp3 = G.LP(**{**lp.__dict__, "row_offsets": ptr})
    with pytest.raises(c_api.CuOptError):
        c_api.Problem(lp3)


def test_crossing_bounds_report_infeasible_without_gpu_work():
    lp = G.random_lp(m=10, n=20, seed=3)
    lp.var_lb = lp.var_lb.copy()
    lp.var_lb[4] = 11.0  # above the upper bound 10 (has_crossing_bounds, solve.cu:1176-1179)
    r = c_api.solve(lp)
    assert r.error == 0 and r.status == 2


def test_no_constraints():
    lp = c_api.read_lp(os.path.join(GOLDEN, "lp-model-no-constraints.mps"))
    r = c_api.solve(lp, method=c_api.CUOPT_METHOD_PDLP)
    assert r.status == 6  # PDLP cannot run without constraints (solve.cu:563-568)
    r = c_api.solve(lp, method=c_api.CUOPT_METHOD_CONCURRENT)
    assert r.status == 1 and abs(r.primal_objective - 1.0) < 1e-12  # pdlp_test.cu:1028-1043


def test_mip_and_other_methods_are_rejected():
    lp = G.random_lp(m=10, n=20, seed=4)
    r = c_api.solve(lp, method=c_api.CUOPT_METHOD_DUAL_SIMPLEX)
    assert r.error == c_api.CUOPT_INVALID_ARGUMENT and "PDLP" in r.error_string


def test_solve_without_gpu_fails_cleanly():
    if c_api.device_count() > 0:
        pytest.skip("a GPU is present")
    lp = G.random_lp(m=20, n=50, seed=5)
    r = c_api.solve(lp, iteration_limit=10)
    assert r.error == c_api.CUOPT_RUNTIME_ERROR
    assert "HIP" in r.error_string


def test_batch_without_gpu_fails_cleanly_and_keeps_host_results():
    # crossing bounds and constraint-free LPs are answered on the host; the rest
    # report the HIP error of the missing device, every handle filled and in order
    if c_api.device_count() > 0:
        pytest.skip("a GPU is present")
    good = G.random_lp(m=20, n=50, seed=5)
    crossing = G.random_lp(m=20, n=50, seed=6)
    crossing.var_lb = crossing.var_lb.copy()
    crossing.var_lb[0] = 11.0
    rs = c_api.batch_solve([good, crossing], iteration_limit=10)
    assert rs[0].error == c_api.CUOPT_RUNTIME_ERROR and "HIP" in rs[0].error_string
    assert rs[1].status == 2 and len(rs[1].x) == crossing.n


def test_batch_and_unrolled_argument_checks():
    L = c_api.lib()
    s = c_api.SolverSettings()
    assert L.cuOptAmdBatchSolve(-1, None, s.handle, None) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptAmdBatchSolve(1, None, s.handle, None) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptAmdBatchSolve(0, None, s.handle, None) == c_api.CUOPT_SUCCESS
    h = ctypes.c_void_p()
    L.cuOptAmdUnrolledCreate.argtypes = [ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 5 + [
        ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    ptr = np.array([0, 1], np.int32)
    idx = np.array([3], np.int32)  # column out of range for n = 2
    val, lb, ub = np.ones(1), np.zeros(2), np.ones(2)
    rc = L.cuOptAmdUnrolledCreate(1, 2, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data, lb.ctypes.data,
                                  ub.ctypes.data, 1e-3, 1, ctypes.byref(h))
    assert rc == c_api.CUOPT_INVALID_ARGUMENT and not h.value
    idx[0] = 1
    rc = L.cuOptAmdUnrolledCreate(1, 2, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data, lb.ctypes.data,
                                  ub.ctypes.data, -1.0, 1, ctypes.byref(h))
    assert rc == c_api.CUOPT_INVALID_ARGUMENT  # negative ridge


# ---- MPS files through cuOptReadProblem / cuOptWriteProblem ---------------------
GOLDEN_MPS = ["afiro_original.mps", "good-max.mps", "max_offset.mps", "lp-model-no-constraints.mps",
              "good-mps-fixed-ranges.mps", "lp_model_with_var_bounds.mps"]


def assert_same_lp(a, b):
    assert (a.m, a.n, a.maximize) == (b.m, b.n, b.maximize)
    assert a.offset == b.offset
    assert np.array_equal(a.row_offsets, b.row_offsets) and np.array_equal(a.col_indices, b.col_indices)
    for f in ("values", "c", "con_lb", "con_ub", "var_lb", "var_ub"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("fname", GOLDEN_MPS)
def test_read_mps_matches_python_reader(fname):
    path = os.path.join(GOLDEN, fname)
    assert_same_lp(c_api.Problem.read_mps(path).lp, c_api.read_lp(path))


@pytest.mark.parametrize("fname", GOLDEN_MPS)
def test_write_then_read_mps_round_trips(fname, tmp_path):
    p = c_api.Problem.read_mps(os.path.join(GOLDEN, fname))
    out = tmp_path / "w.mps"
    assert p.write_mps(out) == c_api.CUOPT_SUCCESS
    assert_same_lp(c_api.Problem.read_mps(out).lp, p.lp)


def test_read_mps_known_values():
    # afiro: 27 rows, 32 columns, 83 nonzeros; good-max maximises with OBJSENSE
    afiro = c_api.Problem.read_mps(os.path.join(GOLDEN, "afiro_original.mps")).lp
    assert (afiro.m, afiro.n, afiro.nnz) == (27, 32, 83)
    assert c_api.Problem.read_mps(os.path.join(GOLDEN, "good-max.mps")).lp.maximize


def test_read_mps_errors(tmp_path):
    L = c_api.lib()
    h = ctypes.c_void_p(1)
    assert L.cuOptReadProblem(str(tmp_path / "missing.mps").encode(), ctypes.byref(h)) == c_api.CUOPT_MPS_FILE_ERROR
    assert h.value is None
    bad = tmp_path / "bad.mps"
    bad.write_text("NAME x\nROWS\n N OBJ\n L R1\nCOLUMNS\n    X1 R9 1.0\nENDATA\n")
    assert L.cuOptReadProblem(str(bad).encode(), ctypes.byref(h)) == c_api.CUOPT_MPS_PARSE_ERROR
    bad.write_text("NAME x\nROWS\n N OBJ\nCOLUMNS\n    X1 OBJ abc\nENDATA\n")
    assert L.cuOptReadProblem(str(bad).encode(), ctypes.byref(h)) == c_api.CUOPT_MPS_PARSE_ERROR
    p = c_api.Problem.read_mps(os.path.join(GOLDEN, "afiro_original.mps"))
    assert p.write_mps(tmp_path / "x.mps", fmt=7) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptWriteProblem(p.handle, b"", 0) == c_api.CUOPT_INVALID_ARGUMENT


def _qp_args(lp, q_ptr, q_idx, q_val):
    s, rhs = lp.row_sense()
    keep = [np.ascontiguousarray(a) for a in (lp.c, q_ptr, q_idx, q_val, lp.row_offsets.astype(np.int32),
                                              lp.col_indices.astype(np.int32), lp.values, np.frombuffer(s.tobytes(), np.uint8).copy(),
                                              rhs, lp.var_lb, lp.var_ub)]
    return keep, [k.ctypes.data for k in keep]


def test_quadratic_create_refuses_qp_and_accepts_empty_q():
    # The reference sends QPs to barrier (solve.cu:1142-1152); with no quadratic
    # nonzeros the problem is the LP itself, with continuous variables.
    L = c_api.lib()
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    L.cuOptCreateQuadraticProblem.argtypes = [i32, i32, i32, f64] + [vp] * 11 + [ctypes.POINTER(vp)]
    L.cuOptCreateQuadraticRangedProblem.argtypes = [i32, i32, i32, f64] + [vp] * 11 + [ctypes.POINTER(vp)]
    lp = G.random_lp(m=12, n=30, seed=7)
    h = ctypes.c_void_p()
    empty = (np.zeros(31, np.int32), np.zeros(1, np.int32), np.zeros(1))
    keep, ptrs = _qp_args(lp, *empty)
    assert L.cuOptCreateQuadraticProblem(12, 30, 0, ctypes.c_double(0.0), *ptrs, ctypes.byref(h)) == 0 and h.value
    nz, mip = ctypes.c_int32(), ctypes.c_int32()
    assert L.cuOptGetNumNonZeros(h, ctypes.byref(nz)) == 0 and nz.value == lp.nnz
    assert L.cuOptIsMIP(h, ctypes.byref(mip)) == 0 and mip.value == 0
    L.cuOptDestroyProblem(ctypes.byref(h))
    q_ptr = np.zeros(31, np.int32)
    q_ptr[1:] = 1  # one diagonal entry in row 0
    keep, ptrs = _qp_args(lp, q_ptr, np.zeros(1, np.int32), np.ones(1))
    h = ctypes.c_void_p()
    assert L.cuOptCreateQuadraticProblem(12, 30, 0, ctypes.c_double(0.0), *ptrs, ctypes.byref(h)) == \
        c_api.CUOPT_INVALID_ARGUMENT
    assert not h.value
    # ranged form: same rule
    lc, uc = np.ascontiguousarray(lp.con_lb), np.ascontiguousarray(lp.con_ub)
    args = [lp.c, empty[0], empty[1], empty[2], lp.row_offsets.astype(np.int32), lp.col_indices.astype(np.int32),
            lp.values, lc, uc, lp.var_lb, lp.var_ub]
    args = [np.ascontiguousarray(a) for a in args]
    h = ctypes.c_void_p()
    assert L.cuOptCreateQuadraticRangedProblem(12, 30, 0, ctypes.c_double(0.0), *[a.ctypes.data for a in args],
                                               ctypes.byref(h)) == 0 and h.value
    L.cuOptDestroyProblem(ctypes.byref(h))
    args[1] = q_ptr
    h = ctypes.c_void_p()
    assert L.cuOptCreateQuadraticRangedProblem(12, 30, 0, ctypes.c_double(0.0), *[a.ctypes.data for a in args],
                                               ctypes.byref(h)) == c_api.CUOPT_INVALID_ARGUMENT


def test_mip_only_settings_and_getters():
    # As cuopt_c.cpp:758-836: callbacks and MIP starts are kept with the
    # settings (argument checks only); an LP solution has no MIP gap or bound
    # (cuopt_c.cpp:1050-1082).
    L = c_api.lib()
    vp = ctypes.c_void_p
    L.cuOptAddMIPStart.argtypes = [vp, vp, ctypes.c_int32]
    L.cuOptGetMIPGap.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.cuOptGetSolutionBound.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.cuOptSolve.argtypes = [vp, vp, ctypes.POINTER(vp)]
    s = c_api.SolverSettings()
    x = np.ones(5)
    assert L.cuOptAddMIPStart(s.handle, x.ctypes.data, 5) == 0
    assert L.cuOptAddMIPStart(s.handle, x.ctypes.data, 0) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptAddMIPStart(None, x.ctypes.data, 5) == c_api.CUOPT_INVALID_ARGUMENT
    cb_t = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
    cb = cb_t(lambda a, b, c, d: None)
    L.cuOptSetMIPGetSolutionCallback.argtypes = [ctypes.c_void_p, cb_t, ctypes.c_void_p]
    L.cuOptSetMIPSetSolutionCallback.argtypes = [ctypes.c_void_p, cb_t, ctypes.c_void_p]
    assert L.cuOptSetMIPGetSolutionCallback(s.handle, cb, None) == 0
    assert L.cuOptSetMIPSetSolutionCallback(s.handle, cb, None) == 0
    assert L.cuOptSetMIPGetSolutionCallback(None, cb, None) == c_api.CUOPT_INVALID_ARGUMENT
    # a host-answered LP solution (crossing bounds)
    lp = G.random_lp(m=10, n=20, seed=3)
    lp.var_lb = lp.var_lb.copy()
    lp.var_lb[4] = 11.0
    p = c_api.Problem(lp)
    sol = ctypes.c_void_p()
    assert L.cuOptSolve(p.handle, c_api.make_settings().handle, ctypes.byref(sol)) == 0
    v = ctypes.c_double()
    assert L.cuOptGetMIPGap(sol, ctypes.byref(v)) == c_api.CUOPT_INVALID_ARGUMENT
    assert L.cuOptGetSolutionBound(sol, ctypes.byref(v)) == c_api.CUOPT_INVALID_ARGUMENT
    L.cuOptDestroySolution(ctypes.byref(sol))


@pytest.mark.parametrize("name", ["per_constraint_residual", "first_primal_feasible", "save_best_primal_so_far"])
def test_batch_rejects_single_lp_termination_options(name):
    # the reference's batch mode rejects them (convergence_information.cu:393-395,
    # pdlp.cu:931-933); no device work is done for a rejected problem
    lps = [G.random_lp(m=60, n=150, seed=44)]
    (r,) = c_api.batch_solve(lps, tol=1e-4, time_limit=60.0, **{name: True})
    assert r.error == c_api.CUOPT_VALIDATION_ERROR
