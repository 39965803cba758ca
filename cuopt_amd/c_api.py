"""ctypes binding of libcuopt_amd.so: the cuOpt C API (LP subset) plus extensions.

Mirrors the reference C API one to one (``cuOptCreateProblem``,
``cuOptSetParameter``, ``cuOptSolve``, getters).  Loading fails loudly when the
HIP library has not been built; nothing here computes a solution itself.
"""

from __future__ import annotations

import ctypes
import dataclasses
import gc
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuopt_amd.so")

# constants (include/cuopt_c.h)
CUOPT_SUCCESS, CUOPT_INVALID_ARGUMENT, CUOPT_MPS_FILE_ERROR, CUOPT_MPS_PARSE_ERROR = 0, 1, 2, 3
CUOPT_VALIDATION_ERROR, CUOPT_OUT_OF_MEMORY, CUOPT_RUNTIME_ERROR = 4, 5, 6
CUOPT_MINIMIZE, CUOPT_MAXIMIZE = 1, -1
CUOPT_METHOD_CONCURRENT, CUOPT_METHOD_PDLP, CUOPT_METHOD_DUAL_SIMPLEX, CUOPT_METHOD_BARRIER = 0, 1, 2, 3
CUOPT_PRESOLVE_DEFAULT, CUOPT_PRESOLVE_OFF, CUOPT_PRESOLVE_PAPILO, CUOPT_PRESOLVE_PSLP = -1, 0, 1, 2
# termination statuses (cuopt_c.h; the reference's spelling)
CUOPT_TERIMINATION_STATUS_NO_TERMINATION, CUOPT_TERIMINATION_STATUS_OPTIMAL = 0, 1
CUOPT_TERIMINATION_STATUS_INFEASIBLE, CUOPT_TERIMINATION_STATUS_UNBOUNDED = 2, 3
CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT, CUOPT_TERIMINATION_STATUS_TIME_LIMIT = 4, 5
CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR, CUOPT_TERIMINATION_STATUS_PRIMAL_FEASIBLE = 6, 7
STATUS = {0: "NoTermination", 1: "Optimal", 2: "PrimalInfeasible", 3: "DualInfeasible", 4: "IterationLimit",
          5: "TimeLimit", 6: "NumericalError", 7: "PrimalFeasible", 9: "ConcurrentLimit"}
INFO_NAMES = [
    "steps", "l2_primal_residual", "l2_relative_primal_residual", "l2_dual_residual",
    "l2_relative_dual_residual", "primal_objective", "dual_objective", "gap", "relative_gap",
    "solve_time", "setup_time", "loop_time", "eta0", "omega0", "omega", "restarts", "power_iterations",
    "sigma_max_sq", "beta", "gamma", "precision", "checks", "structure_time", "scaling_time", "power_time",
    "refine_rounds", "tile_a", "lanes_bias_a", "tile_at", "lanes_bias_at", "tune_time", "merge_a", "merge_at",
    "lp_time", "lp_start",
]
TRACE_WIDTH = 10
PRECISION = {"fp64": 0, "fp32_mixed": 1, "mixed": 1, "fp32": 1}

_lib = None


class LibraryNotBuilt(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        # PyTorch ships its own HIP runtime under the same soname as the system one;
        # the first to load serves the whole process.  Loading torch first keeps one
        # runtime shared by torch tensors and this library (autograd path).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        i32, f64, vp = ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
        pi32, pf64, pch = ctypes.POINTER(i32), ctypes.POINTER(f64), ctypes.c_char_p
        L.cuOptCreateProblem.argtypes = [i32, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(vp)]
        L.cuOptCreateRangedProblem.argtypes = [i32, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                               ctypes.POINTER(vp)]
        L.cuOptReadProblem.argtypes = [pch, ctypes.POINTER(vp)]
        L.cuOptWriteProblem.argtypes = [vp, pch, i32]
        for f in ("cuOptDestroyProblem", "cuOptDestroySolverSettings", "cuOptDestroySolution"):
            getattr(L, f).argtypes = [ctypes.POINTER(vp)]
            getattr(L, f).restype = None
        L.cuOptCreateSolverSettings.argtypes = [ctypes.POINTER(vp)]
        L.cuOptSetParameter.argtypes = [vp, pch, pch]
        L.cuOptGetParameter.argtypes = [vp, pch, i32, ctypes.c_char_p]
        L.cuOptSetIntegerParameter.argtypes = [vp, pch, i32]
        L.cuOptGetIntegerParameter.argtypes = [vp, pch, pi32]
        L.cuOptSetFloatParameter.argtypes = [vp, pch, f64]
        L.cuOptGetFloatParameter.argtypes = [vp, pch, pf64]
        L.cuOptSetInitialPrimalSolution.argtypes = [vp, vp, i32]
        L.cuOptSetInitialDualSolution.argtypes = [vp, vp, i32]
        L.cuOptSolve.argtypes = [vp, vp, ctypes.POINTER(vp)]
        L.cuOptIsMIP.argtypes = [vp, pi32]
        for f in ("cuOptGetTerminationStatus", "cuOptGetErrorStatus"):
            getattr(L, f).argtypes = [vp, pi32]
        L.cuOptGetErrorString.argtypes = [vp, ctypes.c_char_p, i32]
        for f in ("cuOptGetPrimalSolution", "cuOptGetDualSolution", "cuOptGetReducedCosts"):
            getattr(L, f).argtypes = [vp, vp]
        for f in ("cuOptGetObjectiveValue", "cuOptGetSolveTime", "cuOptGetDualObjectiveValue"):
            getattr(L, f).argtypes = [vp, pf64]
        for f in ("cuOptGetNumConstraints", "cuOptGetNumVariables", "cuOptGetObjectiveSense",
                  "cuOptGetNumNonZeros"):
            getattr(L, f).argtypes = [vp, pi32]
        L.cuOptGetObjectiveOffset.argtypes = [vp, pf64]
        for f in ("cuOptGetObjectiveCoefficients", "cuOptGetConstraintRightHandSide",
                  "cuOptGetConstraintLowerBounds", "cuOptGetConstraintUpperBounds",
                  "cuOptGetVariableLowerBounds", "cuOptGetVariableUpperBounds", "cuOptGetConstraintSense",
                  "cuOptGetVariableTypes"):
            getattr(L, f).argtypes = [vp, vp]
        L.cuOptGetConstraintMatrix.argtypes = [vp, vp, vp, vp]
        L.cuOptGetFloatSize.restype = ctypes.c_int8
        L.cuOptGetIntSize.restype = ctypes.c_int8
        L.cuOptGetVersion.argtypes = [pi32, pi32, pi32]
        L.cuOptAmdGetTerminationInfo.argtypes = [vp, vp, i32]
        L.cuOptAmdGetTrace.argtypes = [vp, vp, i32, pi32]
        L.cuOptAmdSpmv.argtypes = [i32, i32, vp, vp, vp, vp, vp, i32]
        L.cuOptAmdSpmvNoAlloc.argtypes = [i32, i32, vp, vp, vp, vp, vp, i32]
        L.cuOptAmdSpmvScratchBytes.argtypes = [ctypes.POINTER(ctypes.c_size_t)]
        L.cuOptAmdDeviceCount.restype = i32
        L.cuOptAmdBenchmarkSteps.argtypes = [vp, vp, i32, i32, vp, i32]
        L.cuOptAmdBatchSolve.argtypes = [i32, vp, vp, vp]
        L.cuOptAmdGetSolutions.argtypes = [i32, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _a(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def _p(a):
    """Address of a contiguous numpy array, for c_void_p arguments (the caller keeps a alive)."""
    return a.ctypes.data


class CuOptError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"cuOpt error {code}: {msg}")
        self.code = code


def _check(rc, what=""):
    if rc != CUOPT_SUCCESS:
        raise CuOptError(rc, what)


class Problem:
    """Owns a cuOptOptimizationProblem handle."""

    def __init__(self, lp, ranged: bool | None = None):
        L = lib()
        self.lp = lp
        self.handle = ctypes.c_void_p()
        self._keep = []
        ptr, idx, val = _a(lp.row_offsets, np.int32), _a(lp.col_indices, np.int32), _a(lp.values, np.float64)
        c, lb, ub = _a(lp.c, np.float64), _a(lp.var_lb, np.float64), _a(lp.var_ub, np.float64)
        vt = np.frombuffer(b"C" * lp.n, dtype=np.uint8).copy() if lp.n else np.zeros(1, np.uint8)
        sense = CUOPT_MAXIMIZE if lp.maximize else CUOPT_MINIMIZE
        sr = None
        if ranged is None:
            try:
                sr = lp.row_sense()
                ranged = False
            except ValueError:
                ranged = True
        if ranged:
            lc, uc = _a(lp.con_lb, np.float64), _a(lp.con_ub, np.float64)
            rc = L.cuOptCreateRangedProblem(lp.m, lp.n, sense, lp.offset, _p(c), _p(ptr), _p(idx), _p(val), _p(lc),
                                            _p(uc), _p(lb), _p(ub), _p(vt), ctypes.byref(self.handle))
        else:
            s, rhs = sr if sr is not None else lp.row_sense()
            s = np.frombuffer(s.tobytes(), dtype=np.uint8).copy() if lp.m else np.zeros(1, np.uint8)
            rhs = _a(rhs, np.float64) if lp.m else np.zeros(1)
            rc = L.cuOptCreateProblem(lp.m, lp.n, sense, lp.offset, _p(c), _p(ptr), _p(idx), _p(val), _p(s), _p(rhs),
                                      _p(lb), _p(ub), _p(vt), ctypes.byref(self.handle))
        _check(rc, "cuOptCreateProblem")

    @classmethod
    def read_mps(cls, path: str) -> "Problem":
        """cuOptReadProblem; the returned problem's ``lp`` is rebuilt from the C getters."""
        self = cls.__new__(cls)
        self.handle = ctypes.c_void_p()
        self._keep = []
        _check(lib().cuOptReadProblem(str(path).encode(), ctypes.byref(self.handle)), f"cuOptReadProblem {path}")
        self.lp = problem_to_lp(self.handle)
        return self

    def write_mps(self, path: str, fmt: int = 0) -> int:
        """cuOptWriteProblem; returns the C status code."""
        return lib().cuOptWriteProblem(self.handle, str(path).encode(), fmt)

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            lib().cuOptDestroyProblem(ctypes.byref(self.handle))


def problem_to_lp(handle):
    """An LP from the problem getters (constraint bounds from whichever form the problem holds)."""
    from .generators import LP

    L = lib()
    i32 = ctypes.c_int32
    m, n, nnz, sense = i32(), i32(), i32(), i32()
    L.cuOptGetNumConstraints(handle, ctypes.byref(m))
    L.cuOptGetNumVariables(handle, ctypes.byref(n))
    L.cuOptGetNumNonZeros(handle, ctypes.byref(nnz))
    L.cuOptGetObjectiveSense(handle, ctypes.byref(sense))
    m, n, nnz = m.value, n.value, nnz.value
    off = ctypes.c_double()
    L.cuOptGetObjectiveOffset(handle, ctypes.byref(off))
    ptr, idx, val = np.zeros(m + 1, np.int32), np.zeros(max(nnz, 1), np.int32), np.zeros(max(nnz, 1))
    _check(L.cuOptGetConstraintMatrix(handle, _p(ptr), _p(idx), _p(val)))
    c, lb, ub = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
    L.cuOptGetObjectiveCoefficients(handle, _p(c))
    L.cuOptGetVariableLowerBounds(handle, _p(lb))
    L.cuOptGetVariableUpperBounds(handle, _p(ub))
    lo, hi = np.zeros(max(m, 1)), np.zeros(max(m, 1))
    if L.cuOptGetConstraintLowerBounds(handle, _p(lo)) == CUOPT_SUCCESS:
        L.cuOptGetConstraintUpperBounds(handle, _p(hi))
    else:
        s, rhs = np.zeros(max(m, 1), np.uint8), np.zeros(max(m, 1))
        _check(L.cuOptGetConstraintSense(handle, _p(s)))
        _check(L.cuOptGetConstraintRightHandSide(handle, _p(rhs)))
        t = s.tobytes().decode()
        lo = np.array([-np.inf if t[i] == "L" else rhs[i] for i in range(m)])
        hi = np.array([np.inf if t[i] == "G" else rhs[i] for i in range(m)])
    return LP(m=m, n=n, row_offsets=ptr, col_indices=idx[:nnz], values=val[:nnz], c=c[:n], con_lb=lo[:m],
              con_ub=hi[:m], var_lb=lb[:n], var_ub=ub[:n], maximize=sense.value == CUOPT_MAXIMIZE,
              offset=off.value, name="mps")


class SolverSettings:
    """Owns a cuOptSolverSettings handle."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        _check(lib().cuOptCreateSolverSettings(ctypes.byref(self.handle)))
        self._keep = []

    def set(self, name: str, value) -> int:
        L = lib()
        if isinstance(value, bool):
            return L.cuOptSetIntegerParameter(self.handle, name.encode(), int(value))
        if isinstance(value, (int, np.integer)):
            return L.cuOptSetIntegerParameter(self.handle, name.encode(), int(value))
        if isinstance(value, float):
            return L.cuOptSetFloatParameter(self.handle, name.encode(), value)
        return L.cuOptSetParameter(self.handle, name.encode(), str(value).encode())

    def get_string(self, name: str) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(lib().cuOptGetParameter(self.handle, name.encode(), 256, buf))
        return buf.value.decode()

    def set_initial_primal(self, x):
        x = _a(x, np.float64)
        self._keep.append(x)
        _check(lib().cuOptSetInitialPrimalSolution(self.handle, _p(x), len(x)))

    def set_initial_dual(self, y):
        y = _a(y, np.float64)
        self._keep.append(y)
        _check(lib().cuOptSetInitialDualSolution(self.handle, _p(y), len(y)))

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            lib().cuOptDestroySolverSettings(ctypes.byref(self.handle))


@dataclasses.dataclass
class Result:
    status: int
    error: int
    error_string: str
    x: np.ndarray
    y: np.ndarray
    reduced_cost: np.ndarray
    primal_objective: float
    dual_objective: float
    solve_time: float
    info: dict
    trace: np.ndarray

    @property
    def iterations(self) -> int:
        return int(self.info["steps"]) if self.info["steps"] == self.info["steps"] else -1


def solve_problem(problem: Problem, settings: SolverSettings) -> Result:
    L = lib()
    sol = ctypes.c_void_p()
    err = L.cuOptSolve(problem.handle, settings.handle, ctypes.byref(sol))
    if not sol.value:
        raise CuOptError(err, "cuOptSolve returned no solution")
    return _take_result(sol, problem.lp)


_NO_TRACE = np.zeros((0, TRACE_WIDTH))


def _take_result(sol, lp) -> Result:
    """Reads every getter of a solution handle, then destroys it.  x, y, the reduced costs
    and the termination info are views of one host buffer, so that one array address
    serves the four getters that fill them."""
    L = lib()
    try:
        st, es = ctypes.c_int32(), ctypes.c_int32()
        L.cuOptGetTerminationStatus(sol, ctypes.byref(st))
        L.cuOptGetErrorStatus(sol, ctypes.byref(es))
        msg = ""
        if es.value != 0:  # the message of a successful solve is empty
            buf = ctypes.create_string_buffer(1024)
            L.cuOptGetErrorString(sol, buf, 1024)
            msg = buf.value.decode()
        n, m, ni = lp.n, lp.m, len(INFO_NAMES)
        out = np.zeros(2 * n + m + ni)  # a failed solve may leave arrays unwritten
        base = out.__array_interface__["data"][0]
        L.cuOptGetPrimalSolution(sol, base)
        L.cuOptGetDualSolution(sol, base + 8 * n)
        L.cuOptGetReducedCosts(sol, base + 8 * (n + m))
        L.cuOptAmdGetTerminationInfo(sol, base + 8 * (2 * n + m), ni)
        po, do, tm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        L.cuOptGetObjectiveValue(sol, ctypes.byref(po))
        L.cuOptGetDualObjectiveValue(sol, ctypes.byref(do))
        L.cuOptGetSolveTime(sol, ctypes.byref(tm))
        nt = ctypes.c_int32()
        L.cuOptAmdGetTrace(sol, None, 0, ctypes.byref(nt))
        tr = _NO_TRACE
        if nt.value > 0:
            tr = np.zeros((nt.value, TRACE_WIDTH))
            L.cuOptAmdGetTrace(sol, _p(tr), nt.value, ctypes.byref(nt))
            tr = tr[: nt.value]
        info = out[2 * n + m:]
        return Result(st.value, es.value, msg, out[:n], out[n:n + m], out[n + m:2 * n + m], po.value, do.value,
                      tm.value, dict(zip(INFO_NAMES, info.tolist())), tr)
    finally:
        L.cuOptDestroySolution(ctypes.byref(sol))


def make_settings(tol: float = 1e-4, precision: str = "fp64", iteration_limit: int | None = None,
                  time_limit: float | None = None, method: int = CUOPT_METHOD_PDLP, log: bool = False,
                  refinement: bool | None = None, presolve: bool = False) -> SolverSettings:
    """Settings with all six tolerances at ``tol``, presolve off (unless asked for) and the given limits."""
    s = SolverSettings()
    for name in ("absolute_primal_tolerance", "relative_primal_tolerance", "absolute_dual_tolerance",
                 "relative_dual_tolerance", "absolute_gap_tolerance", "relative_gap_tolerance"):
        _check(s.set(name, float(tol)), name)
    _check(s.set("method", method))
    _check(s.set("presolve", CUOPT_PRESOLVE_PSLP if presolve else 0))
    _check(s.set("pdlp_precision", PRECISION[precision]))
    _check(s.set("log_to_console", bool(log)))
    if iteration_limit is not None:
        _check(s.set("iteration_limit", int(iteration_limit)))
    if time_limit is not None:
        _check(s.set("time_limit", float(time_limit)))
    if refinement is not None:
        _check(s.set("pdlp_refinement", int(bool(refinement))))
    return s


def read_lp(path: str):
    """An MPS file read by cuOptReadProblem, as a generators.LP rebuilt from the C getters."""
    return Problem.read_mps(path).lp


def solve(lp, tol: float = 1e-4, precision: str = "fp64", iteration_limit: int | None = None,
          time_limit: float | None = None, x_init=None, y_init=None, method: int = CUOPT_METHOD_PDLP,
          log: bool = False, refinement: bool | None = None, presolve: bool = False,
          params: dict | None = None) -> Result:
    """Solve ``lp`` (a generators.LP, or a :class:`Problem` such as one read by
    ``Problem.read_mps``) through the C ABI with all six tolerances set to ``tol``;
    ``params`` sets further parameters by name."""
    prob = lp if isinstance(lp, Problem) else Problem(lp)
    s = make_settings(tol, precision, iteration_limit, time_limit, method, log, refinement, presolve)
    for name, v in (params or {}).items():
        _check(s.set(name, v), name)
    if x_init is not None:
        s.set_initial_primal(x_init)
    if y_init is not None:
        s.set_initial_dual(y_init)
    return solve_problem(prob, s)


def batch_solve(lps, tol: float = 1e-4, precision: str = "fp64", iteration_limit: int | None = None,
                time_limit: float | None = None, **params) -> list[Result]:
    """cuOptAmdBatchSolve over ``lps`` (LP objects or already built ``Problem`` handles): one
    persistent launch for the whole list.  Further keyword arguments are settings by parameter
    name."""
    probs = [lp if isinstance(lp, Problem) else Problem(lp) for lp in lps]
    s = make_settings(tol, precision, iteration_limit, time_limit)
    for name, v in params.items():
        _check(s.set(name, v), name)
    hs = (ctypes.c_void_p * max(len(probs), 1))(*[p.handle.value for p in probs])
    sols = (ctypes.c_void_p * max(len(probs), 1))()
    L = lib()
    L.cuOptAmdBatchSolve(len(probs), hs, s.handle, sols)
    return _take_batch(sols, probs)


def _take_batch(sols, probs) -> list[Result]:
    """Reads the solutions of a batch with one cuOptAmdGetSolutions call into one host
    buffer (every result's x, y and reduced costs are views of it), then destroys them.
    Batch solutions carry no per-check trace."""
    L = lib()
    count = len(probs)
    try:
        for i in range(count):
            if not sols[i]:
                raise CuOptError(CUOPT_RUNTIME_ERROR, "cuOptAmdBatchSolve returned no solution")
        ni = len(INFO_NAMES)
        ns = np.fromiter((p.lp.n for p in probs), np.int64, count)
        ms = np.fromiter((p.lp.m for p in probs), np.int64, count)
        offsets = np.zeros(count + 1, np.int64)
        np.cumsum(2 * ns + ms + ni, out=offsets[1:])
        store = np.zeros(max(int(offsets[-1]), 1))
        status, err = np.zeros(count, np.int32), np.zeros(count, np.int32)
        objs = np.zeros((count, 3))
        _check(L.cuOptAmdGetSolutions(count, sols, _p(store), _p(offsets), _p(status), _p(err), _p(objs)),
               "cuOptAmdGetSolutions")
        st, er, ob = status.tolist(), err.tolist(), objs.tolist()
        out = []
        # Thousands of small results: the cyclic collector would rescan them all as they
        # accumulate (none of these objects forms a cycle).
        collect = gc.isenabled()
        gc.disable()
        try:
            for i, (o, n, m) in enumerate(zip(offsets[:-1].tolist(), ns.tolist(), ms.tolist())):
                msg = ""
                if er[i] != 0:
                    buf = ctypes.create_string_buffer(1024)
                    L.cuOptGetErrorString(sols[i], buf, 1024)
                    msg = buf.value.decode()
                e = o + 2 * n + m
                out.append(Result(st[i], er[i], msg, store[o:o + n], store[o + n:o + n + m], store[o + n + m:e],
                                  ob[i][0], ob[i][1], ob[i][2], dict(zip(INFO_NAMES, store[e:e + ni].tolist())),
                                  _NO_TRACE))
        finally:
            if collect:
                gc.enable()
        return out
    finally:
        for i in range(count):
            if sols[i]:
                h = ctypes.c_void_p(sols[i])
                L.cuOptDestroySolution(ctypes.byref(h))
                sols[i] = None


_BATCH_COST = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int32, ctypes.c_void_p)


def batch_size_search(max_batch: int, cost) -> int:
    """cuOptAmdBatchSizeSearch: the batch-size search over ``cost(b)`` (time per batch
    member at batch size b); see cuOptAmdUnrolledOptimalBatchSize for the device-timed form."""
    L = lib()
    if not getattr(L, "_bss_bound", False):
        L.cuOptAmdBatchSizeSearch.argtypes = [ctypes.c_int32, _BATCH_COST, ctypes.c_void_p,
                                              ctypes.POINTER(ctypes.c_int32)]
        L._bss_bound = True
    errors = []

    def cb(b, _ctx):
        try:
            return float(cost(int(b)))
        except Exception as e:  # an exception cannot cross the C boundary
            errors.append(e)
            return float("inf")

    fn = _BATCH_COST(cb)
    out = ctypes.c_int32()
    rc = L.cuOptAmdBatchSizeSearch(int(max_batch), fn, None, ctypes.byref(out))
    if errors:
        raise errors[0]
    _check(rc, "cuOptAmdBatchSizeSearch")
    return out.value


def spmv(row_offsets, col_indices, values, x, ncols: int, precision: str = "fp64") -> np.ndarray:
    """y = A x on the GPU with the solver's tiled SpMV kernel."""
    ptr, idx, val = _a(row_offsets, np.int32), _a(col_indices, np.int32), _a(values, np.float64)
    x = _a(x, np.float64)
    m = len(ptr) - 1
    y = np.zeros(max(m, 1))
    _check(lib().cuOptAmdSpmv(m, ncols, _p(ptr), _p(idx), _p(val), _p(x), _p(y), PRECISION[precision]),
           "cuOptAmdSpmv")
    return y[:m]


def spmv_noalloc(row_offsets, col_indices, values, x, ncols: int, precision: str = "fp64") -> np.ndarray:
    """y = A x through cuOptAmdSpmvNoAlloc: the same kernels as :func:`spmv`, with every
    device buffer in the library's fixed per-GPU array (no device allocation)."""
    ptr, idx, val = _a(row_offsets, np.int32), _a(col_indices, np.int32), _a(values, np.float64)
    x = _a(x, np.float64)
    m = len(ptr) - 1
    y = np.zeros(max(m, 1))
    _check(lib().cuOptAmdSpmvNoAlloc(m, ncols, _p(ptr), _p(idx), _p(val), _p(x), _p(y), PRECISION[precision]),
           "cuOptAmdSpmvNoAlloc")
    return y[:m]


def spmv_scratch_bytes() -> int:
    b = ctypes.c_size_t(0)
    _check(lib().cuOptAmdSpmvScratchBytes(ctypes.byref(b)), "cuOptAmdSpmvScratchBytes")
    return int(b.value)


BENCH_NAMES = ["window_ms", "primal_kernel_ms", "dual_kernel_ms", "iterations", "setup_s", "checks", "status",
               "precision", "event_iterations"]


def benchmark_steps(lp, warmup: int, steps: int, precision: str = "fp32_mixed", tol: float = 0.0,
                    params: dict | None = None) -> dict:
    """Times exactly ``steps`` PDHG iterations after ``warmup`` (cuOptAmdBenchmarkSteps)."""
    prob = Problem(lp)
    s = SolverSettings()
    for name, v in (params or {}).items():
        _check(s.set(name, v), name)
    for name in ("absolute_primal_tolerance", "relative_primal_tolerance", "absolute_dual_tolerance",
                 "relative_dual_tolerance", "absolute_gap_tolerance", "relative_gap_tolerance"):
        _check(s.set(name, float(tol)), name)
    _check(s.set("pdlp_precision", PRECISION[precision]))
    _check(s.set("log_to_console", False))
    out = np.zeros(len(BENCH_NAMES))
    _check(lib().cuOptAmdBenchmarkSteps(prob.handle, s.handle, int(warmup), int(steps), _p(out), len(out)),
           "cuOptAmdBenchmarkSteps")
    return dict(zip(BENCH_NAMES, out.tolist()))


def device_count() -> int:
    return int(lib().cuOptAmdDeviceCount())
