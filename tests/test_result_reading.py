"""CPU: cuOptAmdGetSolutions (the bulk reader a batch uses) against the one-call-per-getter
reads of the same solution handles.  The handles come from cuOptSolve on problems that
presolve settles on the host (so no GPU is needed) and from a MIP problem, which this LP
engine rejects with an error solution whose arrays were never written."""

import ctypes

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd import generators as G

L = None


def setup_module():
    global L
    try:
        L = c_api.lib()
    except c_api.LibraryNotBuilt as e:
        pytest.skip(str(e))


def settled_lp(seed: int, m: int, n: int) -> G.LP:
    """Singleton equality rows a_i x_i = b_i on the first m columns; the other columns
    appear in no row, so presolve fixes every variable and recovers y and z."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, size=m) * rng.choice([-1.0, 1.0], size=m)
    x = rng.uniform(0.0, 4.0, size=m)
    return G.LP(m=m, n=n, row_offsets=np.arange(m + 1, dtype=np.int32), col_indices=np.arange(m, dtype=np.int32),
                values=a, c=rng.normal(size=n), con_lb=a * x, con_ub=a * x, var_lb=np.zeros(n),
                var_ub=np.full(n, 5.0))


def solve_handle(prob_handle, settings):
    sol = ctypes.c_void_p()
    L.cuOptSolve(prob_handle, settings.handle, ctypes.byref(sol))
    assert sol.value
    return sol


def mip_handle(lp):
    """A problem with one integer variable (cuOptCreateProblem with variable type 'I')."""
    sense, rhs = lp.row_sense()
    vt = np.frombuffer(b"I" + b"C" * (lp.n - 1), dtype=np.uint8).copy()
    h = ctypes.c_void_p()
    arrays = (lp.c, lp.row_offsets, lp.col_indices, lp.values, sense, rhs, lp.var_lb, lp.var_ub, vt)
    assert L.cuOptCreateProblem(lp.m, lp.n, c_api.CUOPT_MINIMIZE, 0.0, *[a.ctypes.data for a in arrays],
                                ctypes.byref(h)) == 0
    return h, arrays


TIMES = {"solve_time", "setup_time", "loop_time", "structure_time", "scaling_time", "power_time", "tune_time",
         "lp_time", "lp_start"}


def same(a, b):
    return a == b or (a != a and b != b)


def test_bulk_read_matches_per_getter_reads():
    settings = c_api.make_settings(1e-4, "fp64", None, None, presolve=True)
    lps = [settled_lp(s, m, n) for s, (m, n) in enumerate([(3, 7), (1, 1), (40, 90), (0, 5), (12, 12)])]
    probs = [c_api.Problem(lp) for lp in lps]
    mip_lp = settled_lp(9, 4, 6)
    mh, keep = mip_handle(mip_lp)
    handles = [p.handle for p in probs] + [mh]
    lps_all = lps + [mip_lp]
    count, ni = len(handles), len(c_api.INFO_NAMES)
    # 1. The bulk call and the per-getter reads of the same handles agree exactly.
    sols = (ctypes.c_void_p * count)(*[solve_handle(h, settings).value for h in handles])
    sizes = [2 * lp.n + lp.m + ni for lp in lps_all]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    store = np.full(offsets[-1], np.pi)
    status, err, objs = np.zeros(count, np.int32), np.zeros(count, np.int32), np.zeros((count, 3))
    assert L.cuOptAmdGetSolutions(count, sols, store.ctypes.data, offsets.ctypes.data, status.ctypes.data,
                                  err.ctypes.data, objs.ctypes.data) == 0
    for i, lp in enumerate(lps_all):
        r = c_api._take_result(ctypes.c_void_p(sols[i]), lp)  # reads every getter, then destroys
        o, n, m = int(offsets[i]), lp.n, lp.m
        assert (status[i], err[i]) == (r.status, r.error)
        assert np.array_equal(store[o:o + n], r.x) and np.array_equal(store[o + n:o + n + m], r.y)
        assert np.array_equal(store[o + n + m:o + 2 * n + m], r.reduced_cost)
        info = store[o + 2 * n + m:o + 2 * n + m + ni]
        assert all(same(v, r.info[k]) for k, v in zip(c_api.INFO_NAMES, info))
        assert same(objs[i, 0], r.primal_objective) and same(objs[i, 1], r.dual_objective)
        assert objs[i, 2] == r.solve_time
    assert status[:-1].tolist() == [c_api.CUOPT_TERIMINATION_STATUS_OPTIMAL] * len(lps)
    assert err[-1] == c_api.CUOPT_INVALID_ARGUMENT
    o = int(offsets[-2])  # arrays of the failed solve read as zeros
    assert not store[o:o + 2 * mip_lp.n + mip_lp.m].any()

    # 2. _take_batch (what batch_solve returns) over fresh handles of the same problems.
    sols = (ctypes.c_void_p * count)(*[solve_handle(h, settings).value for h in handles])

    class P:  # what _take_batch reads of a problem
        def __init__(self, lp):
            self.lp = lp

    bulk = c_api._take_batch(sols, [P(lp) for lp in lps_all])
    assert all(s is None for s in sols)  # every handle destroyed
    one = [c_api._take_result(solve_handle(h, settings), lp) for h, lp in zip(handles, lps_all)]
    for r1, r2 in zip(one, bulk):
        assert (r1.status, r1.error, r1.error_string) == (r2.status, r2.error, r2.error_string)
        for a, b in ((r1.x, r2.x), (r1.y, r2.y), (r1.reduced_cost, r2.reduced_cost)):
            assert a.shape == b.shape and np.array_equal(a, b)
        assert same(r1.primal_objective, r2.primal_objective) and same(r1.dual_objective, r2.dual_objective)
        assert all(same(r1.info[k], r2.info[k]) for k in r1.info if k not in TIMES)
        assert r2.trace.shape == (0, c_api.TRACE_WIDTH)
    assert bulk[-1].error == c_api.CUOPT_INVALID_ARGUMENT and bulk[-1].error_string
    for lp, r in zip(lps, bulk):  # the settled point itself
        ax = np.zeros(lp.m)
        np.add.at(ax, np.repeat(np.arange(lp.m), np.diff(lp.row_offsets)), lp.values * r.x[lp.col_indices])
        assert np.allclose(ax, lp.con_lb)
    L.cuOptDestroyProblem(ctypes.byref(mh))
    del keep


def test_bulk_read_rejects_wrong_sizes_before_writing():
    settings = c_api.make_settings(1e-4, "fp64", None, None, presolve=True)
    lps = [settled_lp(1, 3, 5), settled_lp(2, 2, 4)]
    probs = [c_api.Problem(lp) for lp in lps]
    sols = (ctypes.c_void_p * 2)(*[solve_handle(p.handle, settings).value for p in probs])
    ni = len(c_api.INFO_NAMES)
    sizes = [2 * lp.n + lp.m + ni for lp in lps]
    store = np.full(sum(sizes) + 1, 7.0)
    status, err, objs = np.full(2, -5, np.int32), np.full(2, -5, np.int32), np.full(6, 7.0)
    try:
        bad = np.array([0, sizes[0], sum(sizes) + 1], np.int64)  # second solution one slot too long
        rc = L.cuOptAmdGetSolutions(2, sols, store.ctypes.data, bad.ctypes.data, status.ctypes.data,
                                    err.ctypes.data, objs.ctypes.data)
        assert rc == c_api.CUOPT_INVALID_ARGUMENT
        assert (store == 7.0).all() and (status == -5).all() and (err == -5).all() and (objs == 7.0).all()
        assert L.cuOptAmdGetSolutions(-1, sols, store.ctypes.data, bad.ctypes.data, status.ctypes.data,
                                      err.ctypes.data, objs.ctypes.data) == c_api.CUOPT_INVALID_ARGUMENT
        assert L.cuOptAmdGetSolutions(2, sols, None, bad.ctypes.data, status.ctypes.data, err.ctypes.data,
                                      objs.ctypes.data) == c_api.CUOPT_INVALID_ARGUMENT
        good = np.array([0, sizes[0], sum(sizes)], np.int64)
        assert L.cuOptAmdGetSolutions(2, sols, store.ctypes.data, good.ctypes.data, status.ctypes.data,
                                      err.ctypes.data, objs.ctypes.data) == 0
        assert (status == c_api.CUOPT_TERIMINATION_STATUS_OPTIMAL).all() and (err == 0).all()
        assert store[-1] == 7.0  # nothing past the last offset
    finally:
        for i in range(2):
            h = ctypes.c_void_p(sols[i])
            L.cuOptDestroySolution(ctypes.byref(h))
