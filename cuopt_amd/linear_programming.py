"""The reference's Python LP interface over this library's C ABI.

Mirrors ``cuopt.linear_programming`` for the LP path: ``DataModel``
(python/cuopt/cuopt/linear_programming/data_model/data_model.py),
``SolverSettings`` with ``SolverMethod`` / ``PDLPSolverMode``
(solver_settings/solver_settings.py), ``Solve`` / ``BatchSolve``
(solver/solver.py:12-105 and the Cython wrapper solver_wrapper.pyx:494-560),
``Solution`` (solution/solution.py) and ``ParseMps``
(cuopt_mps_parser/parser.py:10).  Names, argument meaning, defaults and the
termination / error enums are the reference's:

- unset variable bounds are [0, +inf) (problem_helpers.cuh:54-79);
- constraints come either from row types ('E', 'L', 'G') with the bounds
  vector b, or from constraint lower / upper bounds (these win when both are
  given);
- ``Solve`` returns a Solution whose error status / message carry solver
  errors; argument errors raise ``ValueError``, as the reference's setters do;
- ``BatchSolve`` returns ``(solutions, solve_time)`` for the whole batch and
  solves it in one persistent launch (cuOptAmdBatchSolve).

Not available on this engine (the reference routes them elsewhere): MIP
solves, quadratic objectives (barrier) and PDLP warm-start data; each raises
or reports the reason.  No CPU fallback: the library must be built.
"""

from __future__ import annotations

import ctypes
from enum import IntEnum, auto

import numpy as np

from . import c_api
from .c_api import _a, _check, _p, lib

__all__ = ["DataModel", "SolverSettings", "SolverMethod", "PDLPSolverMode", "Solve", "BatchSolve", "Solution",
           "LPTerminationStatus", "ErrorStatus", "ProblemCategory", "ParseMps"]


class LPTerminationStatus(IntEnum):
    """pdlp_termination_status_t (pdlp/solver_solution.hpp:27-37) with the C constants' values."""
    NoTermination = 0
    NumericalError = 6
    Optimal = 1
    PrimalInfeasible = 2
    DualInfeasible = 3
    IterationLimit = 4
    TimeLimit = 5
    PrimalFeasible = 7


class ErrorStatus(IntEnum):
    """error_type_t (error.hpp:20-25)."""
    Success = 0
    ValidationError = 4
    OutOfMemoryError = 5
    RuntimeError = 6


class ProblemCategory(IntEnum):
    LP = 0
    MIP = 1
    IP = 2


class SolverMethod(IntEnum):
    Concurrent = 0
    PDLP = auto()
    DualSimplex = auto()
    Barrier = auto()

    def __str__(self):
        return "%d" % self.value


class PDLPSolverMode(IntEnum):
    Stable1 = 0
    Stable2 = auto()
    Methodical1 = auto()
    Fast1 = auto()
    Stable3 = auto()

    def __str__(self):
        return "%d" % self.value


def _f64(x, name):
    a = np.asarray(x)
    if a.size and not np.issubdtype(a.dtype, np.number):
        raise ValueError(f"{name} must be numeric")
    return np.ascontiguousarray(a, dtype=np.float64).ravel()


def _i32(x, name):
    a = np.asarray(x)
    if a.size and not np.issubdtype(a.dtype, np.integer):
        raise ValueError(f"{name} must be an integer array")
    return np.ascontiguousarray(a, dtype=np.int32).ravel()


def _chars(x):
    a = np.asarray(x)
    if a.dtype.kind in ("S", "U"):
        return np.array([str(v.decode() if isinstance(v, bytes) else v)[:1] for v in a.ravel()])
    return np.array([chr(int(v)) for v in a.ravel()])


class DataModel:
    """A linear program: minimize (or maximize) c.x over l <= x <= u with
    constraints given by row types and bounds b, or by lower / upper bounds.
    The objective reported is objective_scaling_factor * (c.x + objective_offset)."""

    def __init__(self):
        self.maximize = False
        self.A_values = np.zeros(0)
        self.A_indices = np.zeros(0, np.int32)
        self.A_offsets = np.zeros(0, np.int32)
        self.b = np.zeros(0)
        self.c = np.zeros(0)
        self.objective_scaling_factor = 1.0
        self.objective_offset = 0.0
        self.Q_values = np.zeros(0)
        self.Q_indices = np.zeros(0, np.int32)
        self.Q_offsets = np.zeros(0, np.int32)
        self.variable_lower_bounds = np.zeros(0)
        self.variable_upper_bounds = np.zeros(0)
        self.constraint_lower_bounds = np.zeros(0)
        self.constraint_upper_bounds = np.zeros(0)
        self.row_types = np.zeros(0, dtype="<U1")
        self.variable_types = np.zeros(0, dtype="<U1")
        self.variable_names = []
        self.row_names = []
        self.objective_name = ""
        self.problem_name = ""
        self.initial_primal_solution = np.zeros(0)
        self.initial_dual_solution = np.zeros(0)

    # ---- setters (data_model.py:154-486)
    def set_maximize(self, maximize):
        self.maximize = bool(maximize)

    def set_csr_constraint_matrix(self, A_values, A_indices, A_offsets):
        self.A_values = _f64(A_values, "A_values")
        self.A_indices = _i32(A_indices, "A_indices")
        self.A_offsets = _i32(A_offsets, "A_offsets")

    def set_constraint_bounds(self, b):
        self.b = _f64(b, "b")

    def set_objective_coefficients(self, c):
        self.c = _f64(c, "c")

    def set_objective_scaling_factor(self, objective_scaling_factor):
        self.objective_scaling_factor = float(objective_scaling_factor)

    def set_objective_offset(self, objective_offset):
        self.objective_offset = float(objective_offset)

    def set_quadratic_objective_matrix(self, Q_values, Q_indices, Q_offsets):
        self.Q_values = _f64(Q_values, "Q_values")
        self.Q_indices = _i32(Q_indices, "Q_indices")
        self.Q_offsets = _i32(Q_offsets, "Q_offsets")

    def set_variable_lower_bounds(self, variable_lower_bounds):
        self.variable_lower_bounds = _f64(variable_lower_bounds, "variable_lower_bounds")

    def set_variable_upper_bounds(self, variable_upper_bounds):
        self.variable_upper_bounds = _f64(variable_upper_bounds, "variable_upper_bounds")

    def set_constraint_lower_bounds(self, constraint_lower_bounds):
        self.constraint_lower_bounds = _f64(constraint_lower_bounds, "constraint_lower_bounds")

    def set_constraint_upper_bounds(self, constraint_upper_bounds):
        self.constraint_upper_bounds = _f64(constraint_upper_bounds, "constraint_upper_bounds")

    def set_variable_types(self, variable_types):
        self.variable_types = _chars(variable_types)

    def set_row_types(self, row_types):
        self.row_types = _chars(row_types)

    def set_variable_names(self, variables_names):
        self.variable_names = [str(v) for v in variables_names]

    def set_row_names(self, row_names):
        self.row_names = [str(v) for v in row_names]

    def set_objective_name(self, objective_name):
        self.objective_name = str(objective_name)

    def set_problem_name(self, problem_name):
        self.problem_name = str(problem_name)

    def set_initial_primal_solution(self, initial_primal_solution):
        self.initial_primal_solution = _f64(initial_primal_solution, "initial_primal_solution")

    def set_initial_dual_solution(self, initial_dual_solution):
        self.initial_dual_solution = _f64(initial_dual_solution, "initial_dual_solution")

    # ---- getters (data_model.py:488-681)
    def get_sense(self):
        return self.maximize

    def get_constraint_matrix_values(self):
        return self.A_values

    def get_constraint_matrix_indices(self):
        return self.A_indices

    def get_constraint_matrix_offsets(self):
        return self.A_offsets

    def get_constraint_bounds(self):
        return self.b

    def get_objective_coefficients(self):
        return self.c

    def get_objective_scaling_factor(self):
        return self.objective_scaling_factor

    def get_objective_offset(self):
        return self.objective_offset

    def get_quadratic_objective_values(self):
        return self.Q_values

    def get_quadratic_objective_indices(self):
        return self.Q_indices

    def get_quadratic_objective_offsets(self):
        return self.Q_offsets

    def get_variable_lower_bounds(self):
        return self.variable_lower_bounds

    def get_variable_upper_bounds(self):
        return self.variable_upper_bounds

    def get_constraint_lower_bounds(self):
        return self.constraint_lower_bounds

    def get_constraint_upper_bounds(self):
        return self.constraint_upper_bounds

    def get_row_types(self):
        return self.row_types

    def get_ascii_row_types(self):
        return np.array([ord(t) for t in self.row_types], dtype=np.int8)

    def get_initial_primal_solution(self):
        return self.initial_primal_solution

    def get_initial_dual_solution(self):
        return self.initial_dual_solution

    def get_variable_types(self):
        return self.variable_types

    def get_variable_names(self):
        return self.variable_names

    def get_row_names(self):
        return self.row_names

    def get_objective_name(self):
        return self.objective_name

    def get_problem_name(self):
        return self.problem_name

    def writeMPS(self, user_problem_file):
        """cuOptWriteProblem of this model (free MPS with generated row / column names)."""
        p = _to_problem(self)
        try:
            _check(lib().cuOptWriteProblem(p, str(user_problem_file).encode(), 0), "cuOptWriteProblem")
        finally:
            lib().cuOptDestroyProblem(ctypes.byref(p))

    # ---- internal
    def _sizes(self):
        n = len(self.c)
        m = len(self.A_offsets) - 1 if len(self.A_offsets) else 0
        return m, n

    def _is_mip(self):
        return any(t == "I" for t in self.variable_types)


class _InvalidModel(ValueError):
    """Incomplete or inconsistent problem data: Solve reports it as a Solution with
    ErrorStatus.ValidationError, as the reference does (test_lp_solver.py:309-368)."""


def _to_problem(dm: DataModel) -> ctypes.c_void_p:
    """A C problem handle for dm (the caller destroys it); raises _InvalidModel on
    incomplete or inconsistent data and ValueError on a quadratic objective."""
    m, n = dm._sizes()
    if len(dm.A_offsets) == 0:
        raise _InvalidModel("set_csr_constraint_matrix must be called before solving")
    nnz = int(dm.A_offsets[-1])
    if len(dm.A_values) != nnz or len(dm.A_indices) != nnz:
        raise _InvalidModel("A_values and A_indices must hold A_offsets[-1] entries")
    lb = dm.variable_lower_bounds if len(dm.variable_lower_bounds) else np.zeros(n)
    ub = dm.variable_upper_bounds if len(dm.variable_upper_bounds) else np.full(n, np.inf)
    for name, v, k in (("variable_lower_bounds", lb, n), ("variable_upper_bounds", ub, n)):
        if len(v) != k:
            raise _InvalidModel(f"{name} must have {k} entries")
    vt = dm.variable_types if len(dm.variable_types) else np.array(["C"] * n)
    if len(vt) != n:
        raise _InvalidModel(f"variable_types must have {n} entries")
    vt_b = np.frombuffer("".join(vt).encode(), np.uint8).copy() if n else np.zeros(1, np.uint8)
    if len(dm.Q_offsets) and int(dm.Q_offsets[-1]) > 0:
        raise ValueError("quadratic objectives are solved by the reference's barrier method, which this LP "
                         "engine does not run")
    L = lib()
    h = ctypes.c_void_p()
    sense = c_api.CUOPT_MAXIMIZE if dm.maximize else c_api.CUOPT_MINIMIZE
    c = _a(dm.c, np.float64) if n else np.zeros(1)
    ptr, idx, val = _a(dm.A_offsets, np.int32), _a(dm.A_indices, np.int32), _a(dm.A_values, np.float64)
    idx = idx if nnz else np.zeros(1, np.int32)
    val = val if nnz else np.zeros(1)
    lb, ub = _a(lb, np.float64), _a(ub, np.float64)
    if n == 0:
        lb, ub = np.zeros(1), np.zeros(1)
    if len(dm.constraint_lower_bounds) and len(dm.constraint_upper_bounds):
        lc, uc = _a(dm.constraint_lower_bounds, np.float64), _a(dm.constraint_upper_bounds, np.float64)
        if len(lc) != m or len(uc) != m:
            raise _InvalidModel(f"constraint bounds must have {m} entries")
        rc = L.cuOptCreateRangedProblem(m, n, sense, dm.objective_offset, _p(c), _p(ptr), _p(idx), _p(val),
                                        _p(lc if m else np.zeros(1)), _p(uc if m else np.zeros(1)), _p(lb), _p(ub),
                                        _p(vt_b), ctypes.byref(h))
    else:
        rt = dm.row_types
        if len(rt) != m or len(dm.b) != m:
            raise _InvalidModel("give row types and constraint bounds b (or constraint lower and upper bounds) "
                                f"for the {m} constraints")
        s = np.frombuffer("".join(rt).encode(), np.uint8).copy() if m else np.zeros(1, np.uint8)
        b = _a(dm.b, np.float64) if m else np.zeros(1)
        rc = L.cuOptCreateProblem(m, n, sense, dm.objective_offset, _p(c), _p(ptr), _p(idx), _p(val), _p(s), _p(b),
                                  _p(lb), _p(ub), _p(vt_b), ctypes.byref(h))
    if rc != c_api.CUOPT_SUCCESS:
        raise _InvalidModel(f"invalid problem data (cuOpt error {rc})")
    try:
        _set_names(L, h, dm, m, n)
    except _InvalidModel:
        L.cuOptDestroyProblem(ctypes.byref(h))
        raise
    return h


def _set_names(L, h, dm: DataModel, m: int, n: int):
    """Hands the data model's names to the C problem (they name the solution_file
    variables and the rows and columns of user_problem_file)."""
    for setter, names, k, what in ((L.cuOptAmdSetVariableNames, dm.variable_names, n, "variable"),
                                   (L.cuOptAmdSetConstraintNames, dm.row_names, m, "row")):
        if len(names) == 0:
            continue
        if len(names) != k:
            raise _InvalidModel(f"{what} names must have {k} entries")
        arr = (ctypes.c_char_p * k)(*[str(v).encode() for v in names])
        setter.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32]
        _check(setter(h, arr, k), setter.__name__)
    for setter, name in ((L.cuOptAmdSetObjectiveName, dm.objective_name), (L.cuOptAmdSetProblemName, dm.problem_name)):
        if name:
            setter.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
            _check(setter(h, str(name).encode()), setter.__name__)


class SolverSettings:
    """Solver parameters by name (the constants of constants.h:23-76), applied at solve time."""

    def __init__(self):
        self.settings_dict = {}
        self.pdlp_warm_start_data = None
        self.mip_callbacks = []

    @staticmethod
    def _valid(name):
        s = c_api.SolverSettings()
        buf = ctypes.create_string_buffer(64)
        return lib().cuOptGetParameter(s.handle, str(name).encode(), 64, buf) == c_api.CUOPT_SUCCESS

    def to_base_type(self, value):
        if value.lower() == "true":
            return True
        if value.lower() == "false":
            return False
        try:
            return int(value)
        except ValueError:
            try:
                return float(value)
            except ValueError:
                return value

    def get_parameter(self, name):
        if not self._valid(name):
            raise ValueError("Invalid parameter. Please check documentation")
        if name in self.settings_dict:
            v = self.settings_dict[name]
            return self.to_base_type(v) if isinstance(v, str) else v
        value = self.to_base_type(c_api.SolverSettings().get_string(name))
        self.settings_dict[name] = value
        return value

    def set_parameter(self, name, value):
        if not self._valid(name):
            raise ValueError("Invalid parameter. Please check documentation")
        self.settings_dict[name] = value

    def set_optimality_tolerance(self, eps_optimal):
        """All six LP tolerances (absolute / relative primal, dual and gap) to eps_optimal."""
        for p in ("absolute_primal_tolerance", "relative_primal_tolerance", "absolute_dual_tolerance",
                  "relative_dual_tolerance", "absolute_gap_tolerance", "relative_gap_tolerance"):
            self.settings_dict[p] = eps_optimal

    def set_pdlp_warm_start_data(self, pdlp_warm_start_data):
        """Stored as in the reference.  Every solver mode runs Stable3 here, where the
        reference does not support PDLP warm starts: Solve then returns a Solution with
        LPTerminationStatus.NoTermination (pdlp_test.cu:880-900), and BatchSolve raises
        (solver_wrapper.pyx:528).  Initial primal / dual solutions are the supported form."""
        self.pdlp_warm_start_data = pdlp_warm_start_data

    def get_pdlp_warm_start_data(self):
        return self.pdlp_warm_start_data

    def set_mip_callback(self, callback, user_data):
        self.mip_callbacks.append((callback, user_data))

    def get_mip_callbacks(self):
        return self.mip_callbacks

    def toDict(self):
        return dict(self.settings_dict)

    def _handle(self) -> c_api.SolverSettings:
        s = c_api.SolverSettings()
        for name, value in self.settings_dict.items():
            if isinstance(value, bool):
                text = "true" if value else "false"
            elif isinstance(value, IntEnum):
                text = str(int(value))
            else:
                text = str(value)
            rc = lib().cuOptSetParameter(s.handle, name.encode(), text.encode())
            if rc != c_api.CUOPT_SUCCESS:
                raise ValueError(f"invalid value {value!r} for parameter {name}")
        return s


class Solution:
    """LP solver output (solution/solution.py)."""

    def __init__(self, problem_category, vars, solve_time=0.0, primal_solution=None, dual_solution=None,
                 reduced_cost=None, termination_status=0, error_status=0, error_message="", primal_residual=0.0,
                 dual_residual=0.0, primal_objective=0.0, dual_objective=0.0, gap=0.0, nb_iterations=0,
                 solved_by_pdlp=None):
        self.problem_category = ProblemCategory(problem_category)
        self.vars = vars
        self.solve_time = solve_time
        self.primal_solution = primal_solution
        self.dual_solution = dual_solution
        self.reduced_cost = reduced_cost
        self.termination_status = LPTerminationStatus(termination_status)
        self.error_status = ErrorStatus(error_status) if error_status in ErrorStatus._value2member_map_ \
            else error_status
        self.error_message = error_message
        self.primal_objective = primal_objective
        self.dual_objective = dual_objective
        self.solved_by_pdlp = solved_by_pdlp
        self.lp_stats = {"primal_residual": primal_residual, "dual_residual": dual_residual, "gap": gap,
                         "nb_iterations": nb_iterations}

    def raise_if_milp_solution(self, function_name):
        if self.problem_category in (ProblemCategory.MIP, ProblemCategory.IP):
            raise AttributeError(f"Attribute {function_name} is not supported for milp solution")

    def raise_if_lp_solution(self, function_name):
        if self.problem_category == ProblemCategory.LP:
            raise AttributeError(f"Attribute {function_name} is not supported for lp solution")

    def get_primal_solution(self):
        return self.primal_solution

    def get_dual_solution(self):
        self.raise_if_milp_solution("get_dual_solution")
        return self.dual_solution

    def get_primal_objective(self):
        return self.primal_objective

    def get_dual_objective(self):
        self.raise_if_milp_solution("get_dual_objective")
        return self.dual_objective

    def get_termination_status(self):
        return self.termination_status

    def get_termination_reason(self):
        return self.termination_status.name

    def get_error_status(self):
        return self.error_status

    def get_error_message(self):
        return self.error_message

    def get_solve_time(self):
        return self.solve_time

    def get_solved_by_pdlp(self):
        return self.solved_by_pdlp

    def get_vars(self):
        return self.vars

    def get_lp_stats(self):
        self.raise_if_milp_solution("get_lp_stats")
        return self.lp_stats

    def get_reduced_cost(self):
        return self.reduced_cost

    def get_pdlp_warm_start_data(self):
        return None

    def get_milp_stats(self):
        self.raise_if_lp_solution("get_milp_stats")

    def get_problem_category(self):
        return self.problem_category


def _solution(dm: DataModel, r: c_api.Result) -> Solution:
    x = np.asarray(r.x, dtype=np.float64)
    obj = dm.objective_scaling_factor * r.primal_objective
    dobj = dm.objective_scaling_factor * r.dual_objective
    names = dm.variable_names
    return Solution(ProblemCategory.LP, dict(zip(names, x)) if names else {}, solve_time=r.solve_time,
                    primal_solution=x, dual_solution=np.asarray(r.y, dtype=np.float64),
                    reduced_cost=np.asarray(r.reduced_cost, dtype=np.float64), termination_status=r.status,
                    error_status=r.error, error_message=r.error_string,
                    primal_residual=r.info.get("l2_primal_residual", 0.0),
                    dual_residual=r.info.get("l2_dual_residual", 0.0), primal_objective=obj, dual_objective=dobj,
                    gap=abs(obj - dobj), nb_iterations=int(r.iterations), solved_by_pdlp=True)


def _invalid_solution(message: str) -> Solution:
    return Solution(ProblemCategory.LP, {}, primal_solution=np.zeros(0), dual_solution=np.zeros(0),
                    reduced_cost=np.zeros(0), termination_status=LPTerminationStatus.NoTermination,
                    error_status=ErrorStatus.ValidationError, error_message=message)


def _prepare(dm: DataModel, settings: SolverSettings | None):
    if settings is None:
        settings = SolverSettings()
    if dm._is_mip():
        raise ValueError("MIP problems are not handled by this LP engine")
    return settings


def Solve(data_model, solver_settings=None):
    """Solves the LP of data_model and returns a Solution (solver.py:12-105)."""
    settings = _prepare(data_model, solver_settings)
    if settings.get_pdlp_warm_start_data() is not None:  # Stable3: no PDLP warm start (pdlp_test.cu:880-900)
        return Solution(ProblemCategory.LP, {}, primal_solution=np.zeros(0), dual_solution=np.zeros(0),
                        reduced_cost=np.zeros(0), termination_status=LPTerminationStatus.NoTermination,
                        error_message="PDLP warm-start data is not supported in Stable3 mode")
    s = settings._handle()
    m, n = data_model._sizes()
    if len(data_model.initial_primal_solution):
        s.set_initial_primal(data_model.initial_primal_solution)
    if len(data_model.initial_dual_solution):
        s.set_initial_dual(data_model.initial_dual_solution)
    try:
        p = _to_problem(data_model)
    except _InvalidModel as e:
        return _invalid_solution(str(e))
    sol = ctypes.c_void_p()
    L = lib()
    try:
        L.cuOptSolve(p, s.handle, ctypes.byref(sol))
        if not sol.value:
            raise RuntimeError("cuOptSolve returned no solution")
        shim = _LpShape(m, n)
        r = c_api._take_result(sol, shim)
    finally:
        L.cuOptDestroyProblem(ctypes.byref(p))
    sol = _solution(data_model, r)
    # presolve without dual postsolve gives no duals (pdlp_test.cu:903-930)
    dual_post = settings.get_parameter("dual_postsolve")
    if isinstance(dual_post, str):
        dual_post = dual_post.strip().lower() in ("1", "true")
    if int(settings.get_parameter("presolve")) != 0 and not dual_post:
        sol.dual_solution, sol.reduced_cost = np.zeros(0), np.zeros(0)
    return sol


def BatchSolve(data_model_list, solver_settings=None):
    """Solves every LP of data_model_list in one persistent launch; returns
    (solutions, solve_time) with the engine time of the whole batch (solver.py:108-150)."""
    settings = solver_settings if solver_settings is not None else SolverSettings()
    if settings.get_pdlp_warm_start_data() is not None:
        raise ValueError("Cannot use warmstart data with Batch Solve")  # solver_wrapper.pyx:528
    for dm in data_model_list:
        _prepare(dm, settings)
    s = settings._handle()
    L = lib()
    handles, invalid = [], {}  # models with incomplete data get a ValidationError solution
    try:
        for i, dm in enumerate(data_model_list):
            try:
                handles.append((i, _to_problem(dm)))
            except _InvalidModel as e:
                invalid[i] = _invalid_solution(str(e))
        k = len(handles)
        hs = (ctypes.c_void_p * max(k, 1))(*[h.value for _, h in handles])
        sols = (ctypes.c_void_p * max(k, 1))()
        if k:
            L.cuOptAmdBatchSolve(k, hs, s.handle, sols)
        out = dict(invalid)
        if k:  # one bulk read of every solution (cuOptAmdGetSolutions), which destroys them all
            shapes = [_Shaped(_LpShape(*data_model_list[i]._sizes())) for i, _ in handles]
            for (i, _), r in zip(handles, c_api._take_batch(sols, shapes)):
                out[i] = _solution(data_model_list[i], r)
    finally:
        for _, h in handles:
            L.cuOptDestroyProblem(ctypes.byref(h))
    solved = [out[i] for i in range(len(data_model_list))]
    times = [out[i].get_solve_time() for i, _ in handles]
    return solved, (times[0] if times else 0.0)


class _LpShape:
    """The sizes c_api._take_result reads."""

    def __init__(self, m, n):
        self.m, self.n = m, n


class _Shaped:
    """A problem's sizes in the form c_api._take_batch reads (``.lp.m``, ``.lp.n``)."""

    def __init__(self, lp):
        self.lp = lp


def _name(getter, handle, *idx):
    buf = ctypes.create_string_buffer(256)
    _check(getter(handle, *idx, buf, 256))
    return buf.value.decode(errors="replace")


def ParseMps(mps_file_path, fixed_mps_format=False):
    """Reads an MPS file (cuOptReadProblem) into a DataModel with its names
    (cuopt_mps_parser/parser.py:10).  The reader follows the free-format rules that
    cuOptReadProblem uses (cuopt_c.cpp:103): fields are separated by blanks, so
    fixed-format names with embedded spaces are rejected."""
    del fixed_mps_format
    L = lib()
    for f in ("cuOptAmdGetVariableName", "cuOptAmdGetConstraintName"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]
    for f in ("cuOptAmdGetObjectiveName", "cuOptAmdGetProblemName"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32]
    h = ctypes.c_void_p()
    rc = L.cuOptReadProblem(str(mps_file_path).encode(), ctypes.byref(h))
    if rc != c_api.CUOPT_SUCCESS:
        raise ValueError(f"cannot read {mps_file_path} (cuOpt error {rc})")
    try:
        lp = c_api.problem_to_lp(h)
        dm = DataModel()
        dm.set_maximize(lp.maximize)
        dm.set_csr_constraint_matrix(lp.values, lp.col_indices, lp.row_offsets)
        dm.set_objective_coefficients(lp.c)
        dm.set_objective_offset(lp.offset)
        dm.set_variable_lower_bounds(lp.var_lb)
        dm.set_variable_upper_bounds(lp.var_ub)
        dm.set_constraint_lower_bounds(lp.con_lb)
        dm.set_constraint_upper_bounds(lp.con_ub)
        if lp.m:
            s, rhs = np.zeros(lp.m, np.uint8), np.zeros(lp.m)
            if L.cuOptGetConstraintSense(h, _p(s)) == c_api.CUOPT_SUCCESS:
                _check(L.cuOptGetConstraintRightHandSide(h, _p(rhs)))
                dm.set_row_types(np.array(list(s.tobytes().decode())))
                dm.set_constraint_bounds(rhs)
        vt = np.zeros(max(lp.n, 1), np.uint8)
        _check(L.cuOptGetVariableTypes(h, _p(vt)))
        dm.set_variable_types(np.array(list(vt[: lp.n].tobytes().decode())))
        dm.set_variable_names([_name(L.cuOptAmdGetVariableName, h, j) for j in range(lp.n)])
        dm.set_row_names([_name(L.cuOptAmdGetConstraintName, h, i) for i in range(lp.m)])
        dm.set_objective_name(_name(L.cuOptAmdGetObjectiveName, h))
        dm.set_problem_name(_name(L.cuOptAmdGetProblemName, h))
        return dm
    finally:
        L.cuOptDestroyProblem(ctypes.byref(h))
