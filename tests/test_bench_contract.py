"""bench.py's host-side pieces, on the CPU: the byte counts its roofline figures divide
by (checked against operand-by-operand counts of small shapes and the figures quoted for
the unrolled workload), and the one-GPU contract (--gpus above 1 exits with a message)."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_pdhg_iteration_bytes_count_each_operand_once():
    m, n, nnz = 3, 5, 7
    for sv in (4, 8):
        b = bench.kernel_bytes(m, n, nnz, sv)
        # primal: A^T offsets (n+1 ints), nnz indices + values, gather y once, x c l u x0 in, x-bar x+ out
        assert b["primal"] == 4 * (n + 1) + nnz * 4 + nnz * sv + m * sv + 5 * n * sv + 2 * n * sv
        # dual: A offsets (m+1 ints), nnz indices + values, gather x-bar once, y l_c u_c y0 in, y+ out
        assert b["dual"] == 4 * (m + 1) + nnz * 4 + nnz * sv + n * sv + 4 * m * sv + m * sv
        assert b["iteration"] == b["primal"] + b["dual"]
    # the headline workload: 54.0 MB per FP32 dual step, 72.0 MB per primal step
    b = bench.kernel_bytes(500_000, 1_000_000, 4_751_177, 4)
    assert round(b["dual"] / 1e6, 1) == 54.0 and round(b["primal"] / 1e6, 1) == 72.0


def test_unrolled_step_bytes():
    # m = 20 000, n = 50 000, 400 000 nonzeros, 64 objectives, FP32: 60.1 MB primal + 26.5 MB dual
    total = bench.unrolled_step_bytes(20_000, 50_000, 400_000, 64)
    assert abs(total - (60.1e6 + 26.5e6)) < 0.1e6
    # one objective, tiny shape, operand by operand
    m, n, nnz = 2, 3, 4
    primal = 4 * (n + 1) + nnz * 8 + m * 4 + 2 * n * 4 + 2 * n * 4 + 2 * n * 4
    dual = 4 * (m + 1) + nnz * 8 + n * 4 + m * 4 + 2 * m * 4 + m * 4
    assert bench.unrolled_step_bytes(m, n, nnz, 1) == primal + dual


def test_loop_roofline_arithmetic():
    r = bench.loop_roofline(10, 20, 30, 8, iterations=1000, loop_s=0.5)
    per = bench.kernel_bytes(10, 20, 30, 8)["iteration"]
    assert r["bytes_per_iteration"] == per
    assert abs(r["achieved"] - per * 1000 / 0.5 / 1e9) < 0.05
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-4


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)


def test_banded_lp_structure():
    # The banded workload: row i holds columns (2 i + d) mod n, d in [-8, 8), sorted;
    # every column has 8 entries; b = A x0 with 0 <= x0 <= 1 inside the bounds.
    import numpy as np
    import scipy.sparse as sp

    from cuopt_amd import generators as G

    lp = G.banded_lp(m=50, width=16, seed=3)
    A = sp.csr_matrix((lp.values, lp.col_indices, lp.row_offsets), shape=(lp.m, lp.n))
    assert lp.n == 100 and lp.nnz == 800 and A.has_sorted_indices
    assert (np.diff(lp.row_offsets) == 16).all() and (np.diff(A.tocsc().indptr) == 8).all()
    i = 20
    assert sorted(lp.col_indices[16 * i:16 * i + 16].tolist()) == list(range(2 * i - 8, 2 * i + 8))
    assert sorted(lp.col_indices[:16].tolist()) == sorted([(d % 100) for d in range(-8, 8)])
    assert np.array_equal(lp.con_lb, lp.con_ub)
    # some x in [0, 1]^n meets A x = b (the generator's x0), so the LP is feasible
    from scipy.optimize import linprog

    r = linprog(np.zeros(lp.n), A_eq=A, b_eq=lp.con_lb, bounds=(0, 1), method="highs")
    assert r.status == 0
