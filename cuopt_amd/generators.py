"""Seeded synthetic LP generators for the benchmark and parity workloads.

Every instance is returned as an :class:`LP` holding host arrays in the form the
cuOpt C API takes them (CSR int32/float64, constraint bounds, variable bounds).

Workloads
---------
* ``random_lp``   -- m x n, k nonzeros per column at distinct uniform rows,
  N(0,1) values, bounds [0, u], all equality rows.  An optimal pair (x0, y0)
  is built in: b = A x0 and c = A^T y0 + s0 with s0 >= 0 supported only where
  x0 = 0, so obj* = c.x0 = b.y0 (complementary slackness by construction).
* ``skewed_lp``   -- m=500 000 x n=1 000 000 by default; each column has
  1+Poisson(4) rows drawn from a truncated Zipf(1.1) over ranks 1..m (rank r is
  row r-1, duplicates inside a column merged), half equality rows and half
  ``<=`` rows whose right-hand side carries a U(0,1) slack.  y0 is zero on the
  ``<=`` rows so (x0, y0) stays optimal and obj* = c.x0 = b.y0.
* ``transportation_lp`` -- S sources x T sinks, each source with ``arcs`` arcs
  to distinct uniform sinks; supply rows ``<=``, demand rows ``=``; integer costs
  U(1,100); total supply = 1.05 x total demand.  No closed-form optimum.
"""

from __future__ import annotations

import dataclasses

import numpy as np

INF = np.inf


@dataclasses.dataclass
class LP:
    """Host-side LP in cuOpt C-API form (minimise c.x + offset)."""

    m: int
    n: int
    row_offsets: np.ndarray  # int32, m+1
    col_indices: np.ndarray  # int32, nnz
    values: np.ndarray  # float64, nnz
    c: np.ndarray  # float64, n
    con_lb: np.ndarray  # float64, m
    con_ub: np.ndarray  # float64, m
    var_lb: np.ndarray  # float64, n
    var_ub: np.ndarray  # float64, n
    maximize: bool = False
    offset: float = 0.0
    # Known optimal point and objective when the generator builds one in.
    x_opt: np.ndarray | None = None
    y_opt: np.ndarray | None = None
    s_opt: np.ndarray | None = None
    obj_opt: float | None = None
    name: str = "lp"

    @property
    def nnz(self) -> int:
        return int(self.row_offsets[-1])

    def row_sense(self) -> tuple[np.ndarray, np.ndarray]:
        """(sense chars, rhs) for cuOptCreateProblem when every row is E/L/G."""
        sense = np.empty(self.m, dtype="S1")
        rhs = np.empty(self.m)
        eq = self.con_lb == self.con_ub
        le = ~eq & np.isinf(self.con_lb)
        ge = ~eq & ~le & np.isinf(self.con_ub)
        if not np.all(eq | le | ge):
            raise ValueError("ranged rows need cuOptCreateRangedProblem")
        sense[eq], rhs[eq] = b"E", self.con_ub[eq]
        sense[le], rhs[le] = b"L", self.con_ub[le]
        sense[ge], rhs[ge] = b"G", self.con_lb[ge]
        return sense, rhs

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.values, self.col_indices, self.row_offsets), shape=(self.m, self.n))


def _csr_from_coo(rows: np.ndarray, cols: np.ndarray, vals: np.ndarray, m: int, n: int):
    key = rows.astype(np.int64) * n + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=m)
    ptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    return ptr.astype(np.int32), cols.astype(np.int32), vals.astype(np.float64)


def _spmv(ptr, idx, val, x):
    m = len(ptr) - 1
    rows = np.repeat(np.arange(m), np.diff(ptr))
    return np.bincount(rows, weights=val * x[idx], minlength=m)


def _spmv_t(ptr, idx, val, y, n):
    m = len(ptr) - 1
    rows = np.repeat(np.arange(m), np.diff(ptr))
    return np.bincount(idx, weights=val * y[rows], minlength=n)


def _optimal_pair(rng, ptr, idx, val, m, n, frac=0.6, y_zero_rows=None):
    x0 = np.zeros(n)
    support = rng.permutation(n)[: int(round(frac * n))]
    x0[support] = rng.uniform(0.0, 1.0, size=support.size)
    y0 = rng.standard_normal(m)
    if y_zero_rows is not None:
        y0[y_zero_rows] = 0.0
    s0 = np.zeros(n)
    off = x0 == 0.0
    s0[off] = rng.uniform(0.0, 1.0, size=int(off.sum()))
    b = _spmv(ptr, idx, val, x0)
    c = _spmv_t(ptr, idx, val, y0, n) + s0
    return x0, y0, s0, b, c


def random_lp(m: int = 2000, n: int = 5000, nnz_per_col: int = 8, upper: float = 10.0, seed: int = 0) -> LP:
    """Random equality LP with a built-in optimum (the 2 000 x 5 000 workload by default).

    Each column gets ``nnz_per_col`` distinct uniform rows (drawn by redrawing
    duplicates, which gives the law of sampling without replacement) with N(0,1)
    values; x0 is U(0,1) on a random 60 % of columns and 0 elsewhere.
    """
    rng = np.random.default_rng(seed)
    k = min(nnz_per_col, m)
    rows = rng.integers(0, m, size=(n, k))
    while True:
        s = np.sort(rows, axis=1)
        dup = np.any(s[:, 1:] == s[:, :-1], axis=1)
        if not dup.any():
            break
        rows[dup] = rng.integers(0, m, size=(int(dup.sum()), k))
    cols = np.repeat(np.arange(n), k)
    vals = rng.standard_normal(n * k)
    ptr, idx, val = _csr_from_coo(rows.reshape(-1), cols, vals, m, n)
    x0, y0, s0, b, c = _optimal_pair(rng, ptr, idx, val, m, n)
    return LP(
        m=m, n=n, row_offsets=ptr, col_indices=idx, values=val, c=c,
        con_lb=b.copy(), con_ub=b.copy(), var_lb=np.zeros(n), var_ub=np.full(n, upper),
        x_opt=x0, y_opt=y0, s_opt=s0, obj_opt=float(c @ x0), name=f"random_{m}x{n}_s{seed}",
    )


def skewed_lp(m: int = 500_000, n: int = 1_000_000, zipf_a: float = 1.1, mean_extra: float = 4.0,
              seed: int = 0) -> LP:
    """Zipf-row LP with heavy rows, half E / half <= rows, built-in optimum."""
    rng = np.random.default_rng(seed)
    per_col = 1 + rng.poisson(mean_extra, size=n)
    total = int(per_col.sum())
    # Truncated Zipf over ranks 1..m by inverse CDF.
    w = np.arange(1, m + 1, dtype=np.float64) ** (-zipf_a)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    rows = np.searchsorted(cdf, rng.random(total), side="right")
    rows = np.minimum(rows, m - 1)
    cols = np.repeat(np.arange(n), per_col)
    key = np.unique(cols.astype(np.int64) * m + rows)  # merge duplicates inside a column
    cols = (key // m).astype(np.int64)
    rows = (key % m).astype(np.int64)
    vals = rng.standard_normal(key.size)
    ptr, idx, val = _csr_from_coo(rows, cols, vals, m, n)
    half = m // 2
    le_rows = np.arange(half, m)
    x0, y0, s0, ax, c = _optimal_pair(rng, ptr, idx, val, m, n, y_zero_rows=le_rows)
    b = ax.copy()
    b[le_rows] += rng.uniform(0.0, 1.0, size=le_rows.size)
    lb = np.empty(m)
    ub = b.copy()
    lb[:half] = b[:half]
    lb[half:] = -INF
    var_ub = rng.uniform(1.0, 100.0, size=n)
    return LP(
        m=m, n=n, row_offsets=ptr, col_indices=idx, values=val, c=c,
        con_lb=lb, con_ub=ub, var_lb=np.zeros(n), var_ub=var_ub,
        x_opt=x0, y_opt=y0, s_opt=s0, obj_opt=float(c @ x0), name=f"skewed_{m}x{n}_s{seed}",
    )


def banded_lp(m: int = 4_000_000, width: int = 16, seed: int = 0) -> LP:
    """Equality LP with n = 2 m columns and a band: row i has `width` nonzeros at columns
    (2 i + d) mod n for d in [-width/2, width/2), so every column has width/2 entries in
    neighbouring rows and both products gather from nearby addresses.  b = A x0 for x0 in
    [0, 1]^n, 0 <= x <= 10, random costs (feasible and bounded; no known optimum).  It
    measures the step kernels where the gathered vector stays in cache."""
    rng = np.random.default_rng(seed)
    n = 2 * m
    d = np.arange(-(width // 2), width - width // 2, dtype=np.int64)
    cols = (2 * np.arange(m, dtype=np.int64)[:, None] + d[None, :]) % n
    cols.sort(axis=1)  # rows at the ends wrap around
    idx = cols.astype(np.int32).ravel()
    val = rng.standard_normal(idx.size)
    ptr = (np.arange(m + 1, dtype=np.int64) * width).astype(np.int32)
    x0 = rng.uniform(0.0, 1.0, size=n)
    b = (val.reshape(m, width) * x0[cols]).sum(axis=1)
    return LP(m=m, n=n, row_offsets=ptr, col_indices=idx, values=val, c=rng.standard_normal(n),
              con_lb=b, con_ub=b.copy(), var_lb=np.zeros(n), var_ub=np.full(n, 10.0), name=f"banded_{m}x{n}")


def transportation_lp(sources: int = 20_000, sinks: int = 50_000, arcs: int = 50, seed: int = 0) -> LP:
    """Transportation LP: supply rows <=, demand rows =, arcs x >= 0 unbounded above."""
    rng = np.random.default_rng(seed)
    n = sources * arcs
    m = sources + sinks
    dest = rng.integers(0, sinks, size=(sources, arcs))
    while True:  # distinct sinks per source
        s = np.sort(dest, axis=1)
        dup = np.any(s[:, 1:] == s[:, :-1], axis=1)
        if not dup.any():
            break
        dest[dup] = rng.integers(0, sinks, size=(int(dup.sum()), arcs))
    dest = dest.reshape(-1)
    src = np.repeat(np.arange(sources), arcs)
    # Every sink needs an incoming arc for feasibility: reroute one arc per orphan sink.
    seen = np.zeros(sinks, dtype=bool)
    seen[dest] = True
    orphans = np.flatnonzero(~seen)
    for t in orphans:
        a = int(rng.integers(0, n))
        dest[a] = t
    cost = rng.integers(1, 101, size=n).astype(np.float64)
    supply = rng.uniform(10.0, 100.0, size=sources)
    demand = rng.uniform(1.0, 10.0, size=sinks)
    demand *= supply.sum() / (1.05 * demand.sum())
    arc = np.arange(n)
    rows = np.concatenate([src, sources + dest])
    cols = np.concatenate([arc, arc])
    vals = np.ones(2 * n)
    ptr, idx, val = _csr_from_coo(rows, cols, vals, m, n)
    lb = np.concatenate([np.full(sources, -INF), demand])
    ub = np.concatenate([supply, demand])
    return LP(
        m=m, n=n, row_offsets=ptr, col_indices=idx, values=val, c=cost,
        con_lb=lb, con_ub=ub, var_lb=np.zeros(n), var_ub=np.full(n, INF),
        name=f"transport_{sources}x{sinks}x{arcs}_s{seed}",
    )


def batch_lps(count: int = 4096, m_min: int = 50, m_max: int = 2000, seed: int = 0) -> list[LP]:
    """The batched workload: ``count`` random equality LPs (random_lp), one seed
    each, with m log-uniform in [m_min, m_max], n = floor(2.5 m), 8 nonzeros per column."""
    rng = np.random.default_rng(seed)
    ms = np.exp(rng.uniform(np.log(m_min), np.log(m_max), size=count)).astype(np.int64)
    return [random_lp(m=int(m), n=int(2.5 * m), nnz_per_col=8, seed=int(seed) * 1_000_003 + i)
            for i, m in enumerate(ms)]


def kkt_residuals(lp: LP, x: np.ndarray, y: np.ndarray, s: np.ndarray | None = None) -> dict:
    """Independent numpy evaluation of primal/dual feasibility of (x, y) for ``lp``.

    Reduced costs are taken as z = c - A^T y when ``s`` is None.  Reported:
    bound-projected primal residual, dual residual of the sign conditions on z
    (z_j >= 0 where x_j is at a finite lower bound only, etc.), objectives.
    """
    A = lp.to_scipy()
    ax = A @ x
    rp = ax - np.clip(ax, lp.con_lb, lp.con_ub)
    z = lp.c - A.T @ y if s is None else s
    return {
        "primal_residual": float(np.linalg.norm(rp)),
        "bound_violation": float(np.max(np.maximum(lp.var_lb - x, 0) + np.maximum(x - lp.var_ub, 0), initial=0.0)),
        "primal_objective": float(lp.c @ x + lp.offset),
        "z": z,
    }
