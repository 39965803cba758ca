"""Writes highs_optima.json: FP64 optima of small generated LPs from SciPy's HiGHS.

Run in the build container (SciPy present); the committed JSON holds the numbers
the oracle and the GPU solver are checked against.  Each entry records the
generator call that rebuilds the instance.
"""

import json
import os
import sys

import numpy as np
import scipy
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cuopt_amd import c_api  # noqa: E402
from cuopt_amd import generators as G  # noqa: E402

CASES = {
    "transport_small": ("transportation_lp", dict(sources=200, sinks=500, arcs=50, seed=3)),
    "transport_medium": ("transportation_lp", dict(sources=2000, sinks=5000, arcs=50, seed=4)),
    # the full 20k x 50k transportation LP (n = 1M arcs, m = 70k rows)
    "transport_full": ("transportation_lp", dict(sources=20000, sinks=50000, arcs=50, seed=0)),
    "skewed_small": ("skewed_lp", dict(m=2000, n=4000, seed=5)),
    "random_small": ("random_lp", dict(m=200, n=500, seed=6)),
}
MPS = ["lp_model_with_var_bounds.mps", "afiro_original.mps",
       "good-max.mps", "max_offset.mps"]


def highs(lp):
    A = lp.to_scipy()
    eq = lp.con_lb == lp.con_ub
    up = np.isfinite(lp.con_ub) & ~eq
    lo = np.isfinite(lp.con_lb) & ~eq
    A_ub = __import__("scipy").sparse.vstack([A[up], -A[lo]]).tocsr()
    b_ub = np.concatenate([lp.con_ub[up], -lp.con_lb[lo]])
    sign = -1.0 if lp.maximize else 1.0
    r = linprog(sign * lp.c, A_ub=A_ub if A_ub.shape[0] else None, b_ub=b_ub if A_ub.shape[0] else None,
                A_eq=A[eq] if eq.any() else None, b_eq=lp.con_lb[eq] if eq.any() else None,
                bounds=list(zip(lp.var_lb, lp.var_ub)), method="highs")
    assert r.status == 0, r.message
    return sign * r.fun + lp.offset


def main():
    out = {"scipy": scipy.__version__, "cases": {}}
    for name, (fn, kw) in CASES.items():
        lp = getattr(G, fn)(**kw)
        out["cases"][name] = {"generator": fn, "kwargs": kw, "objective": highs(lp)}
        print(name, out["cases"][name]["objective"])
    for f in MPS:
        lp = c_api.read_lp(os.path.join(HERE, f))
        out["cases"][f] = {"mps": f, "objective": highs(lp)}
        print(f, out["cases"][f]["objective"])
    with open(os.path.join(HERE, "highs_optima.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
