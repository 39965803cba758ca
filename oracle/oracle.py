"""ctypes front end of the CPU FP64 oracle (``pdlp_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package never
imports this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

# Slots shared with pdlp_oracle.cpp.
P_ABS_PRIMAL, P_REL_PRIMAL, P_ABS_DUAL, P_REL_DUAL, P_ABS_GAP, P_REL_GAP = range(6)
P_ITER_LIMIT, P_TIME_LIMIT, P_THREADS, P_MODE, P_PER_CONSTRAINT, P_NPARAMS = 6, 7, 8, 9, 10, 11
INFO_NAMES = [
    "status", "iterations", "l2_primal_residual", "l2_relative_primal_residual",
    "l2_dual_residual", "l2_relative_dual_residual", "primal_objective", "dual_objective",
    "gap", "relative_gap", "solve_time", "eta0", "omega0", "beta", "gamma", "omega",
    "restarts", "power_iterations", "sigma_max_sq", "b_norm", "c_norm", "setup_time",
]
TRACE_NAMES = ["k", "l2_primal", "l2_dual", "pobj", "dobj", "gap", "restarted", "omega", "fpe", "k_r"]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "build/liboracle.so"], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
        assert _lib.oracle_num_params() == P_NPARAMS
        assert _lib.oracle_num_info() == len(INFO_NAMES)
    return _lib


def _p(a, dtype):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def spmv(ptr, idx, val, x, ncols):
    ptr, idx, val, x = _p(ptr, np.int32), _p(idx, np.int32), _p(val, np.float64), _p(x, np.float64)
    y = np.zeros(len(ptr) - 1)
    lib().oracle_spmv(ctypes.c_int(len(ptr) - 1), ctypes.c_int(ncols), _ptr(ptr), _ptr(idx), _ptr(val), _ptr(x), _ptr(y))
    return y


def transpose(ptr, idx, val, ncols):
    ptr, idx, val = _p(ptr, np.int32), _p(idx, np.int32), _p(val, np.float64)
    nnz = int(ptr[-1])
    tp = np.zeros(ncols + 1, np.int32)
    ti = np.zeros(nnz, np.int32)
    tv = np.zeros(nnz)
    lib().oracle_transpose(ctypes.c_int(len(ptr) - 1), ctypes.c_int(ncols), _ptr(ptr), _ptr(idx), _ptr(val),
                           _ptr(tp), _ptr(ti), _ptr(tv))
    return tp, ti, tv


def setup(lp, mode: int = 0, L=None) -> dict:
    """Scaling vectors and initial step/weight (mode 0 = Stable3, 1 = Methodical1)."""
    a = _arrays(lp)
    Dr, Dc, out = np.zeros(lp.m), np.zeros(lp.n), np.zeros(8)
    sv = np.zeros(lp.nnz)
    rc = _fn_lib(L).oracle_setup(ctypes.c_int(lp.m), ctypes.c_int(lp.n), *[_ptr(v) for v in a],
                            ctypes.c_int(int(lp.maximize)), ctypes.c_int(mode), _ptr(Dr), _ptr(Dc), _ptr(out), _ptr(sv))
    if rc != 0:
        raise ValueError("crossing bounds")
    keys = ["beta", "gamma", "eta", "omega", "sigma_max_sq", "power_iterations", "b_norm", "c_norm"]
    d = dict(zip(keys, out.tolist()))
    d.update(Dr=Dr, Dc=Dc, scaled_values=sv)
    return d


def kkt(lp, x, y, slack, rel_primal: float = 0.0, rel_dual: float = 0.0, L=None) -> dict:
    """KKT record of (x, y, slack) on the unscaled problem; linf_primal / linf_dual are
    the per_constraint_residual maxima for the given relative tolerances."""
    ptr, idx, val = _p(lp.row_offsets, np.int32), _p(lp.col_indices, np.int32), _p(lp.values, np.float64)
    c, lc, uc = _p(lp.c, np.float64), _p(lp.con_lb, np.float64), _p(lp.con_ub, np.float64)
    x, y, slack = _p(x, np.float64), _p(y, np.float64), _p(slack, np.float64)
    out = np.zeros(10)
    _fn_lib(L).oracle_kkt(ctypes.c_int(lp.m), ctypes.c_int(lp.n), _ptr(ptr), _ptr(idx), _ptr(val), _ptr(c), _ptr(lc),
                     _ptr(uc), ctypes.c_int(int(lp.maximize)), ctypes.c_double(lp.offset), _ptr(x), _ptr(y),
                     _ptr(slack), ctypes.c_double(rel_primal), ctypes.c_double(rel_dual), _ptr(out))
    return dict(zip(["l2_primal", "l2_dual", "pobj", "dobj", "gap", "absobj", "b_norm", "c_norm", "linf_primal",
                     "linf_dual"], out.tolist()))


def _arrays(lp):
    return (
        _p(lp.row_offsets, np.int32), _p(lp.col_indices, np.int32), _p(lp.values, np.float64),
        _p(lp.c, np.float64), _p(lp.con_lb, np.float64), _p(lp.con_ub, np.float64),
        _p(lp.var_lb, np.float64), _p(lp.var_ub, np.float64),
    )


def solve(lp, tol: float = 1e-4, iteration_limit: int = 2**31 - 1, time_limit: float = float("inf"),
          threads: int = 1, x_init=None, y_init=None, trace_cap: int = 0, per_constraint_residual: bool = False,
          tolerances=None) -> dict:
    """Stable3 PDLP solve; returns dict with x, y, reduced_cost, info fields and optional trace.
    ``tolerances`` (optional): abs_primal, rel_primal, abs_dual, rel_dual, abs_gap, rel_gap in
    place of ``tol`` for all six."""
    a = _arrays(lp)
    params = np.zeros(P_NPARAMS)
    params[:6] = tol if tolerances is None else tolerances
    params[P_PER_CONSTRAINT] = 1.0 if per_constraint_residual else 0.0
    params[P_ITER_LIMIT] = iteration_limit
    params[P_TIME_LIMIT] = time_limit
    params[P_THREADS] = threads
    x, y, rc = np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n)
    info = np.zeros(len(INFO_NAMES))
    xi, yi = _p(x_init, np.float64), _p(y_init, np.float64)
    trace = np.zeros((max(trace_cap, 1), len(TRACE_NAMES)))
    tlen = ctypes.c_int(0)
    st = lib().oracle_solve(ctypes.c_int(lp.m), ctypes.c_int(lp.n), *[_ptr(v) for v in a],
                            ctypes.c_int(int(lp.maximize)), ctypes.c_double(lp.offset), _ptr(params),
                            _ptr(xi), _ptr(yi), _ptr(x), _ptr(y), _ptr(rc), _ptr(info),
                            _ptr(trace) if trace_cap else None, ctypes.c_int(trace_cap), ctypes.byref(tlen))
    d = dict(zip(INFO_NAMES, info.tolist()))
    d["status"] = int(st)
    d["iterations"] = int(d["iterations"])
    d.update(x=x, y=y, reduced_cost=rc)
    if trace_cap:
        d["trace"] = trace[: tlen.value].copy()
    return d


def _fn_lib(L):
    return lib() if L is None else L


def pid_update(state: dict, primal_distance: float, dual_distance: float, rel_dual: float, rel_primal: float,
               step_size: float, L=None) -> dict:
    """One PID primal-weight update; ``state`` holds omega, best_omega, error_sum,
    last_error, best_gap, tau, sigma.  Returns the new state."""
    keys = ["omega", "best_omega", "error_sum", "last_error", "best_gap", "tau", "sigma"]
    s = np.array([state[k] for k in keys], dtype=np.float64)
    f = _fn_lib(L).oracle_pid_update
    f.restype = None
    f(ctypes.c_double(primal_distance), ctypes.c_double(dual_distance), ctypes.c_double(rel_dual),
      ctypes.c_double(rel_primal), ctypes.c_double(step_size), _ptr(s))
    return dict(zip(keys, s.tolist()))


def restart_decision(k: int, k_r: int, fpe: float, fpe0: float, last_trial: float, L=None) -> bool:
    """The Stable3 restart decision at a major check."""
    f = _fn_lib(L).oracle_restart_decision
    f.restype = ctypes.c_int
    return bool(f(ctypes.c_longlong(k), ctypes.c_longlong(k_r), ctypes.c_double(fpe), ctypes.c_double(fpe0),
                  ctypes.c_double(last_trial)))


def fpe(ptr, idx, val, ncols: int, x, xbar, y, ybar, omega: float, eta: float, L=None) -> float:
    """Fixed-point error of the step (x, y) -> (xbar, ybar) under the CSR matrix."""
    ptr, idx, val = _p(ptr, np.int32), _p(idx, np.int32), _p(val, np.float64)
    x, xbar, y, ybar = (_p(v, np.float64) for v in (x, xbar, y, ybar))
    f = _fn_lib(L).oracle_fpe
    f.restype = ctypes.c_double
    return float(f(ctypes.c_int(len(ptr) - 1), ctypes.c_int(ncols), _ptr(ptr), _ptr(idx), _ptr(val), _ptr(x),
                   _ptr(xbar), _ptr(y), _ptr(ybar), ctypes.c_double(omega), ctypes.c_double(eta)))
