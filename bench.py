#!/usr/bin/env python3
"""Benchmark of the PDLP LP path on one MI355X.

Workload: the skewed LP with m = 500 000 rows and
n = 1 000 000 columns, 1 + Poisson(4) nonzeros per column at Zipf(1.1) rows
(4.75 M nonzeros, a 474 k-nonzero row, 207 k empty rows), half equality and
half <= rows, built-in optimum (x0, y0).  Synthetic, seeded (seed 0).

A "step" is one PDHG iteration of the Stable3 loop (primal step kernel over
A^T, dual step kernel over A, termination checks at the reference cadence).
`value` is FP32-mixed PDHG iterations per second measured over exactly
--steps iterations after --warmup untimed ones, inputs resident in HBM.
Also reported: time to 1e-4 relative KKT for FP64 and FP32-mixed full solves,
their objective error against the closed-form optimum, the roofline of the
dominant kernel (HIP events over the timed window) and the CPU oracle baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "time-to-1e-4 rel-KKT & PDHG iters/s, FP32-mixed vs FP64, 1 MI355X; % HBM peak"


def kernel_bytes(m: int, n: int, nnz: int, sv: int) -> dict:
    """Algorithmic bytes per launch of the two PDHG kernels (each operand touched once).

    primal step (rows of A^T): values+indices nnz*(sv+4), row offsets 4(n+1), gather y: m*sv,
      read x, c, l, u, x0 (5 n sv), write x-bar and the next x (2 n sv)
    dual step (rows of A): nnz*(sv+4), 4(m+1), gather x-bar: n*sv, read y, l_c, u_c, y0 (4 m sv),
      write the next y (m sv)
    Together: 2 nnz (sv+4) + 4 (m+n+2) + sv (8n + 6m).
    """
    primal = nnz * (sv + 4) + 4 * (n + 1) + m * sv + 7 * n * sv
    dual = nnz * (sv + 4) + 4 * (m + 1) + n * sv + 5 * m * sv
    return {"primal": primal, "dual": dual, "iteration": primal + dual}


def unrolled_step_bytes(m: int, n: int, nnz: int, batch: int, sv: int = 4) -> int:
    """Algorithmic bytes of one unrolled PDHG step over `batch` objectives (batch-interleaved
    vectors, each operand touched once):
      primal (rows of A^T): offsets 4(n+1), indices and values nnz(4+sv), gather y (m B sv),
        read x and c (2 n B sv), l and u (2 n sv), write x+ and x-bar (2 n B sv)
      dual (rows of A): offsets 4(m+1), nnz(4+sv), gather x-bar (n B sv), read y (m B sv),
        l_c and u_c (2 m sv), write y+ (m B sv)
    """
    primal = 4 * (n + 1) + nnz * (4 + sv) + m * batch * sv + 2 * n * batch * sv + 2 * n * sv + 2 * n * batch * sv
    dual = 4 * (m + 1) + nnz * (4 + sv) + n * batch * sv + m * batch * sv + 2 * m * sv + m * batch * sv
    return primal + dual


def loop_roofline(m: int, n: int, nnz: int, sv: int, iterations: int, loop_s: float) -> dict:
    """PDHG loop roofline of a single-LP solve: B_iter (kernel_bytes) times the iterations over
    the loop time (termination checks and recentrings are in the time, not in the bytes)."""
    moved = kernel_bytes(m, n, nnz, sv)["iteration"] * iterations
    achieved = moved / max(loop_s, 1e-12) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_iteration": kernel_bytes(m, n, nnz, sv)["iteration"],
            "loop_s": round(loop_s, 4)}


def copy_calibration(mib: int) -> None:
    """One device fill (writes `mib` MiB) and one device-to-device copy (reads and
    writes `mib` MiB), each a single runtime blit kernel: the reference points that
    scale the FETCH_SIZE / WRITE_SIZE counters of the solver kernels."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
    ok = hip.hipMemset(a, 1, nbytes) == 0 and hip.hipDeviceSynchronize() == 0
    ok = ok and hip.hipMemcpy(b, a, nbytes, 3) == 0 and hip.hipDeviceSynchronize() == 0  # device to device
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("copy calibration failed")
    print(f"[bench] copy calibration: fill of {mib} MiB, then copy of {mib} MiB", file=sys.stderr, flush=True)


def stream_bandwidth(mib: int = 1024, reps: int = 10) -> dict:
    """Achievable HBM bandwidth, STREAM style: torch's copy (b = a) and triad
    (c = a + s b) kernels over `mib` MiB arrays, timed with HIP events; the best
    of `reps` runs, counting the bytes each kernel must move (copy 2x, triad 3x)."""
    import torch

    n = (mib << 20) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.ones(n, device="cuda", dtype=torch.float32)
    c = torch.empty(n, device="cuda", dtype=torch.float32)
    out = {}
    for name, fn, moved in (("copy", lambda: c.copy_(a), 2), ("triad", lambda: torch.add(a, b, alpha=0.5, out=c), 3)):
        fn()
        best = None
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        out[name + "_GBps"] = round(moved * n * 4 / (best * 1e-3) / 1e9, 1)
    del a, b, c
    torch.cuda.empty_cache()
    out["array_MiB"] = mib
    print(f"[bench] stream: {out}", file=sys.stderr, flush=True)
    return out


def pmc_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 --pmc summary committed in
    profiles/pmc_traffic.json (FETCH_SIZE and WRITE_SIZE passes scaled by a plain
    copy kernel's counters), or None when no measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get("kernels", {}).get(kernel)
    return entry.get("hbm_bytes_per_launch") if entry else None


def highs_dual_simplex(lp, time_limit: float) -> dict:
    """HiGHS dual simplex (SciPy linprog method 'highs-ds', one thread, FP64) on the same
    host: the stand-in for the reference CPU dual simplex, which does not build here."""
    import numpy as np
    import scipy
    from scipy.optimize import linprog

    A = lp.to_scipy()
    eq = lp.con_lb == lp.con_ub
    up = np.isfinite(lp.con_ub) & ~eq
    lo = np.isfinite(lp.con_lb) & ~eq
    A_ub = scipy.sparse.vstack([A[up], -A[lo]]).tocsr()
    b_ub = np.concatenate([lp.con_ub[up], -lp.con_lb[lo]])
    t = time.time()
    res = linprog(lp.c, A_ub=A_ub if A_ub.shape[0] else None, b_ub=b_ub if A_ub.shape[0] else None,
                  A_eq=A[eq] if eq.any() else None, b_eq=lp.con_lb[eq] if eq.any() else None,
                  bounds=np.stack([lp.var_lb, lp.var_ub], axis=1), method="highs-ds",
                  options={"time_limit": float(time_limit)})
    out = {"solver": f"HiGHS dual simplex (SciPy {scipy.__version__} linprog highs-ds), 1 thread",
           "host_cpus": os.cpu_count(), "time_limit_s": time_limit, "time_s": round(time.time() - t, 3),
           "status": int(res.status), "message": str(res.message)[:80]}
    if res.status == 0:
        out["objective"] = float(res.fun)
    return out


def config0_extras(c_api, G, time_limit: float, simplex_time_limit: float) -> dict:
    """Plain FP64 PDLP on the small random LP (m = 2 000, n = 5 000, 8 nonzeros per column,
    0 <= x <= 10, equality rows, built-in optimum, seed 0), to 1e-4 and 1e-8, beside the
    HiGHS dual simplex on the same host.  FP32-mixed is timed too."""
    lp = G.random_lp(m=2000, n=5000, seed=0)
    obj = lp.obj_opt
    out = {"workload": "random LP m=2000 n=5000 nnz=%d, 8 nnz/col, 0<=x<=10, Ax=b, seed 0" % lp.nnz,
           "closed_form_objective": obj}
    for p in ("fp64", "fp32_mixed"):  # first use of the kernels in this process, untimed
        c_api.solve(lp, tol=1e-4, precision=p, time_limit=time_limit, iteration_limit=20)
    for tol in (1e-4, 1e-8):
        for p in ("fp64", "fp32_mixed"):
            r = c_api.solve(lp, tol=tol, precision=p, time_limit=time_limit)
            out[f"{p}@{tol:g}"] = {"status": c_api.STATUS.get(r.status, r.status), "iterations": r.iterations,
                                   "time_s": round(r.solve_time, 4), "setup_s": round(r.info["setup_time"], 4),
                                   "iters_per_s": round(r.iterations / max(r.info["loop_time"], 1e-12), 1),
                                   "obj_rel_err": abs(r.primal_objective - obj) / (1.0 + abs(obj))}
            print(f"[bench] config0 {p} tol {tol:g}: {out[f'{p}@{tol:g}']}", file=sys.stderr, flush=True)
    if simplex_time_limit > 0:
        try:
            cs = highs_dual_simplex(lp, simplex_time_limit)
            if "objective" in cs:
                cs["obj_rel_err"] = abs(cs["objective"] - obj) / (1.0 + abs(obj))
            out["cpu_simplex"] = cs
            print(f"[bench] config0 cpu simplex: {cs}", file=sys.stderr, flush=True)
        except Exception as e:  # scipy absent on the host
            out["cpu_simplex"] = {"error": repr(e)}
    return out


def transport_extras(c_api, G, time_limit: float, cpu_simplex: bool) -> dict:
    """Transportation LP (20 000 sources x 50 000 sinks, 50 arcs each: n = 1 M, m = 70 k):
    FP32-mixed with FP64 refinement against plain FP64, to 1e-4 and 1e-8, objective error
    against the HiGHS optimum committed in tests/golden/highs_optima.json."""
    with open(os.path.join(ROOT, "tests", "golden", "highs_optima.json")) as f:
        case = json.load(f)["cases"]["transport_full"]
    lp = G.transportation_lp(**case["kwargs"])
    obj = case["objective"]
    out = {"workload": "transportation LP 20000x50000, 50 arcs/source, seed 0 (n=%d m=%d nnz=%d)" % (lp.n, lp.m, lp.nnz),
           "oracle_objective": obj}
    for p in ("fp32_mixed", "fp64"):  # untimed: loads the kernels when this leg runs alone
        c_api.solve(lp, tol=1e-4, precision=p, iteration_limit=0, time_limit=time_limit)
    for tol in (1e-4, 1e-8):
        for p in ("fp32_mixed", "fp64"):
            r = c_api.solve(lp, tol=tol, precision=p, time_limit=time_limit)
            out[f"{p}@{tol:g}"] = {"status": c_api.STATUS.get(r.status, r.status), "iterations": r.iterations,
                                   "time_s": round(r.solve_time, 4), "recentrings": int(r.info["refine_rounds"]),
                                   "obj_rel_err": abs(r.primal_objective - obj) / (1.0 + abs(obj)),
                                   "roofline": loop_roofline(lp.m, lp.n, lp.nnz, 4 if p == "fp32_mixed" else 8,
                                                             r.iterations, r.info["loop_time"])}
            print(f"[bench] transport {p} tol {tol:g}: {out[f'{p}@{tol:g}']}", file=sys.stderr, flush=True)
    if cpu_simplex:
        try:
            import numpy as np
            from scipy.optimize import linprog

            A = lp.to_scipy()
            eq = lp.con_lb == lp.con_ub
            up = np.isfinite(lp.con_ub) & ~eq
            t = time.time()
            res = linprog(lp.c, A_ub=A[up], b_ub=lp.con_ub[up], A_eq=A[eq], b_eq=lp.con_lb[eq], bounds=(0, None),
                          method="highs-ds")
            out["cpu_simplex"] = {"solver": "HiGHS dual simplex (SciPy), 1 thread", "time_s": round(time.time() - t, 3),
                                  "status": int(res.status),
                                  "obj_rel_err": abs(float(res.fun) - obj) / (1.0 + abs(obj))}
            print(f"[bench] transport cpu simplex: {out['cpu_simplex']}", file=sys.stderr, flush=True)
        except Exception as e:  # scipy absent on the host
            out["cpu_simplex"] = {"error": repr(e)}
    return out


def batch_extras(c_api, G, count: int, sample: int, time_limit: float) -> dict:
    """4 096 random equality LPs, m log-uniform in [50, 2000], n = 2.5 m, 8 nonzeros per
    column, FP32-mixed to 1e-4: one persistent launch (cuOptAmdBatchSolve) against a
    sequential loop of single-LP solves over the first `sample` LPs."""
    t = time.time()
    lps = G.batch_lps(count=count, seed=0)
    gen = time.time() - t
    out = {"workload": f"{count} LPs m~logU[50,2000] n=2.5m 8 nnz/col (sum nnz {sum(lp.nnz for lp in lps)})",
           "generation_s": round(gen, 2)}
    # Problems are built once, outside the timed regions of the batch and of the loops below, as
    # the reference's batch call receives built problems: what is timed is the solve call and
    # the reading of every solution.
    t = time.time()
    probs = [c_api.Problem(lp) for lp in lps]
    out["problem_build_s"] = round(time.time() - t, 3)
    for prec in ("fp32_mixed", "fp64"):
        c_api.batch_solve(probs[:2], tol=1e-4, precision=prec, time_limit=time_limit)  # loads the kernel once
        t = time.time()
        rs = c_api.batch_solve(probs, tol=1e-4, precision=prec, time_limit=time_limit)
        wall = time.time() - t
        solved = sum(r.status == 1 for r in rs)
        errs = [abs(r.primal_objective - lp.obj_opt) / (1 + abs(lp.obj_opt)) for r, lp in zip(rs, lps)
                if r.status == 1]
        its = sorted(r.iterations for r in rs)
        # Roofline of the persistent launch: every LP's algorithmic bytes per PDHG iteration
        # (B_iter at its own sizes) times its iterations, over the launch's event time.
        sv = 4 if prec == "fp32_mixed" else 8
        moved = sum(kernel_bytes(lp.m, lp.n, lp.nnz, sv)["iteration"] * r.iterations for lp, r in zip(lps, rs))
        kern = rs[0].info["loop_time"]
        out[prec] = {"wall_s": round(wall, 4), "solve_time_s": round(rs[0].solve_time, 4), "solved": solved,
                     "lps_per_s": round(solved / wall, 1), "max_obj_rel_err": max(errs) if errs else None,
                     "iterations_p50_max": [its[len(its) // 2], its[-1]], "total_iterations": int(sum(its)),
                     "kernel_s": round(kern, 4),
                     "roofline": {"bound": "hbm", "achieved": round(moved / kern / 1e9, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(moved / kern / 1e9 / HBM_PEAK_GBS, 4),
                                  "algorithmic_bytes": moved}}
        print(f"[bench] batch {prec}: {out[prec]}", file=sys.stderr, flush=True)
    t = time.time()
    seq = [c_api.solve(p, tol=1e-4, precision="fp32_mixed", time_limit=time_limit) for p in probs[:sample]]
    seq_wall = time.time() - t
    out["sequential_sample"] = sample
    out["sequential_lps_per_s"] = round(sum(r.status == 1 for r in seq) / seq_wall, 2)
    out["speedup_vs_sequential"] = round(out["fp32_mixed"]["lps_per_s"] / max(out["sequential_lps_per_s"], 1e-9), 1)
    # The reference's batch loop (cython_solve.cu:286-352): at most four single-LP solves at a
    # time, each on its own stream (one solver per host thread; the C calls release the GIL).
    from concurrent.futures import ThreadPoolExecutor

    t = time.time()
    with ThreadPoolExecutor(max_workers=4) as ex:
        four = list(ex.map(lambda p: c_api.solve(p, tol=1e-4, precision="fp32_mixed", time_limit=time_limit),
                           probs[:sample]))
    four_wall = time.time() - t
    out["four_stream_lps_per_s"] = round(sum(r.status == 1 for r in four) / four_wall, 2)
    out["speedup_vs_four_stream"] = round(out["fp32_mixed"]["lps_per_s"] / max(out["four_stream_lps_per_s"], 1e-9), 1)
    print(f"[bench] batch: {out}", file=sys.stderr, flush=True)
    return out


def banded_extras(c_api, G, steps: int, warmup: int, tune: dict) -> dict:
    """The PDHG step kernels on a banded LP (m = 4 M rows, n = 8 M columns, 16 nonzeros per
    row, 64 M in all), where the gathered vector stays in cache: what the same kernels reach
    when the gathers do not bound them.  Roofline per launch: the algorithmic bytes of each
    step (kernel_bytes) over its HIP-event time.  With the SpMV tuner as the headline runs
    it (the plan it picks is reported: on this matrix the plans differ by up to 25 %)."""
    t = time.time()
    lp = G.banded_lp(m=4_000_000, width=16, seed=0)
    out = {"workload": f"banded LP m={lp.m} n={lp.n} nnz={lp.nnz} (16 per row, 8 per column)",
           "generation_s": round(time.time() - t, 2)}
    for p, sv in (("fp32_mixed", 4), ("fp64", 8)):
        plan = {}
        if tune:  # the tuner's choice, cached for the window below
            r = c_api.solve(lp, tol=0.0, precision=p, iteration_limit=0, time_limit=120.0, params=tune)
            plan = {"A": {"tile": int(r.info["tile_a"]), "lanes_shift": int(r.info["lanes_bias_a"])},
                    "A_T": {"tile": int(r.info["tile_at"]), "lanes_shift": int(r.info["lanes_bias_at"])}}
        b = c_api.benchmark_steps(lp, warmup, steps, precision=p, tol=0.0, params=tune)
        kb = kernel_bytes(lp.m, lp.n, lp.nnz, sv)
        res = {"iters_per_s": round(b["iterations"] / (b["window_ms"] * 1e-3), 1), "spmv_plan": plan or "fixed"}
        for step in ("primal", "dual"):
            us = b[f"{step}_kernel_ms"] * 1e3
            gbs = kb[step] / (us * 1e-6) / 1e9
            res[step] = {"mean_launch_us": round(us, 2), "algorithmic_bytes": kb[step],
                         "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": round(gbs / HBM_PEAK_GBS, 4)}}
        out[p] = res
        print(f"[bench] banded {p}: {res}", file=sys.stderr, flush=True)
    return out


def unrolled_extras(G, steps: int, batch: int) -> dict:
    """Differentiable unrolled PDHG: m = 20 000, n = 50 000 (8 nonzeros per column),
    ridge 1e-3, K steps in FP32 for a batch of objectives c + N(0, 0.1), VJP of
    ||x_K - target||^2 with checkpoints every 45 steps."""
    import numpy as np
    import torch

    from cuopt_amd.autograd import UnrolledPDHG

    lp = G.random_lp(m=20_000, n=50_000, seed=0)
    op = UnrolledPDHG(lp, ridge=1e-3, steps=steps, checkpoint_every=45, precision="fp32_mixed")
    rng = np.random.default_rng(0)
    c = torch.tensor(lp.c[None, :] + 0.1 * rng.normal(size=(batch, lp.n)), device="cuda", requires_grad=True)
    b = torch.tensor(lp.con_lb, device="cuda", requires_grad=True)
    target = torch.tensor(rng.uniform(0, 1, size=(batch, lp.n)), device="cuda")
    x, _ = op(c, b)  # warm-up: allocations and the captured segment graphs
    ((x - target) ** 2).sum().backward()
    c.grad = None
    torch.cuda.synchronize()
    # The library launches on torch's current stream, so these events bracket its work.
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t0 = time.time()
    ev[0].record()
    x, _ = op(c, b)
    loss = ((x - target) ** 2).sum()
    ev[1].record()
    torch.cuda.synchronize()
    t1 = time.time()
    loss.backward()
    ev[2].record()
    torch.cuda.synchronize()
    t2 = time.time()
    out = {"workload": f"m=20000 n=50000 nnz={lp.nnz} ridge 1e-3, K={steps}, batch {batch}, FP32, checkpoint 45",
           "forward_ms": round((t1 - t0) * 1e3, 2), "backward_ms": round((t2 - t1) * 1e3, 2),
           "forward_gpu_ms": round(ev[0].elapsed_time(ev[1]), 2), "backward_gpu_ms": round(ev[1].elapsed_time(ev[2]), 2),
           "forward_us_per_step": round((t1 - t0) * 1e6 / steps, 2), "eta": op.step_size}
    # Roofline of the forward pass: every step's algorithmic bytes over the HIP-event time.
    fwd_s = ev[0].elapsed_time(ev[1]) * 1e-3
    moved = unrolled_step_bytes(lp.m, lp.n, lp.nnz, batch) * steps
    achieved = moved / max(fwd_s, 1e-12) / 1e9
    out["forward_roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(achieved / HBM_PEAK_GBS, 4),
                               "bytes_per_step": unrolled_step_bytes(lp.m, lp.n, lp.nnz, batch)}
    print(f"[bench] unrolled: {out}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--no-solves", action="store_true", help="skip the time-to-tolerance solves")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--precisions", default="fp32_mixed,fp64")
    ap.add_argument("--solve-time-limit", type=float, default=300.0)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one plain device-to-device copy of this many MiB (PMC calibration)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the transportation, batched and unrolled workloads")
    ap.add_argument("--no-cpu-simplex", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-sequential-sample", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the headline workload without the SpMV autotuner (pdlp_spmv_autotune)")
    ap.add_argument("--simplex-time-limit", type=float, default=30.0,
                    help="time limit (s) of the HiGHS dual simplex beside the small random LP (0: skip)")
    ap.add_argument("--only", choices=("config0", "transport", "batch", "unrolled", "banded"),
                    help="run just this extra workload and print its result (for profiling)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np

    from cuopt_amd import c_api
    from cuopt_amd import generators as G

    if c_api.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
    extra_fns = {"config0": lambda: config0_extras(c_api, G, 120.0,
                                                   0.0 if args.no_cpu_simplex else args.simplex_time_limit),
                 "transport": lambda: transport_extras(c_api, G, 120.0, not args.no_cpu_simplex),
                 "batch": lambda: batch_extras(c_api, G, args.batch_count, args.batch_sequential_sample, 120.0),
                 "unrolled": lambda: unrolled_extras(G, 2000, 64),
                 "banded": lambda: banded_extras(c_api, G, 300, 30,
                                                 {} if args.no_autotune else {"pdlp_spmv_autotune": 1})}
    if args.only:
        print(json.dumps({"workload": args.only, "result": extra_fns[args.only]()}))
        return 0
    t = time.time()
    lp = G.skewed_lp(m=500_000, n=1_000_000, seed=0)
    gen_s = time.time() - t
    print(f"[bench] generated skewed LP m={lp.m} n={lp.n} nnz={lp.nnz} in {gen_s:.1f}s", file=sys.stderr, flush=True)

    precs = args.precisions.split(",")
    tune = {} if args.no_autotune else {"pdlp_spmv_autotune": 1}
    tuning = {}
    if tune:
        # The autotuner's first run on this matrix (timed in setup, then cached for the process).
        for p in precs:
            r = c_api.solve(lp, tol=0.0, precision=p, iteration_limit=0, time_limit=args.solve_time_limit,
                            params=tune)
            tuning[p] = {"tune_time_s": round(r.info["tune_time"], 4),
                         "A": {"tile": int(r.info["tile_a"]), "lanes_shift": int(r.info["lanes_bias_a"]),
                               "long_row_merge": int(r.info["merge_a"])},
                         "A_T": {"tile": int(r.info["tile_at"]), "lanes_shift": int(r.info["lanes_bias_at"]),
                                 "long_row_merge": int(r.info["merge_at"])}}
            print(f"[bench] spmv autotune {p}: {tuning[p]}", file=sys.stderr, flush=True)
    windows = {}
    for p in precs:
        b = c_api.benchmark_steps(lp, args.warmup, args.steps, precision=p, tol=0.0, params=tune)
        windows[p] = b
        print(f"[bench] {p}: {int(b['iterations'])} steps in {b['window_ms']:.1f} ms "
              f"(primal {b['primal_kernel_ms'] * 1e3:.1f} us, dual {b['dual_kernel_ms'] * 1e3:.1f} us, "
              f"{int(b['checks'])} checks, setup {b['setup_s']:.2f}s)", file=sys.stderr, flush=True)

    solves = {}
    if not args.no_solves:
        for p in precs:
            r = c_api.solve(lp, tol=args.tol, precision=p, time_limit=args.solve_time_limit, params=tune)
            err = abs(r.primal_objective - lp.obj_opt) / (1.0 + abs(lp.obj_opt))
            solves[p] = {
                "status": c_api.STATUS.get(r.status, r.status), "iterations": r.iterations,
                "time_to_tol_s": round(r.solve_time, 4), "loop_s": round(r.info["loop_time"], 4),
                "setup_s": round(r.info["setup_time"], 4),
                "setup_phases_s": {k: round(r.info[k + "_time"], 4) for k in ("structure", "scaling", "power")},
                "power_iterations": int(r.info["power_iterations"]),
                "rel_primal": r.info["l2_relative_primal_residual"], "rel_dual": r.info["l2_relative_dual_residual"],
                "rel_gap": r.info["relative_gap"], "objective": r.primal_objective,
                "obj_rel_err_vs_closed_form": err,
            }
            print(f"[bench] solve {p} tol {args.tol}: {solves[p]}", file=sys.stderr, flush=True)

    head = "fp32_mixed" if "fp32_mixed" in windows else precs[0]
    w = windows[head]
    sv = 4 if head == "fp32_mixed" else 8
    iters = w["iterations"]
    ms_per_step = w["window_ms"] / iters
    value = 1000.0 / ms_per_step
    kb = kernel_bytes(lp.m, lp.n, lp.nnz, sv)
    dom = "primal" if w["primal_kernel_ms"] >= w["dual_kernel_ms"] else "dual"
    dom_ms = w[f"{dom}_kernel_ms"]
    achieved = kb[dom] / (dom_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "kernel": f"{dom}_step ({head})", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(f"{dom}_step_{head}"),
        "algorithmic_bytes_per_launch": kb[dom], "mean_launch_us": round(dom_ms * 1e3, 2),
        "iteration_GBps": round(kb["iteration"] / ((w["primal_kernel_ms"] + w["dual_kernel_ms"]) * 1e-3) / 1e9, 1),
    }

    stream = None
    try:
        stream = stream_bandwidth()
        roofline["achievable_GBps"] = stream["copy_GBps"]
        roofline["frac_of_achievable"] = round(achieved / stream["copy_GBps"], 4)
    except Exception as e:  # never hides the headline line
        print(f"[bench] stream failed: {e!r}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = max(1, min(16, os.cpu_count() or 1))
        r = O.solve(lp, tol=0.0, iteration_limit=args.cpu_iters, threads=threads)
        loop = r["solve_time"] - r["setup_time"]
        cpu = {"value": round(r["iterations"] / loop, 3), "unit": "PDHG iters/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle Stable3 loop, {r['iterations']} iterations of the same LP "
                         f"(setup {r['setup_time']:.1f}s untimed), {threads} OpenMP threads"}
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    extras = {}
    if not args.no_extras:
        for name, fn in extra_fns.items():
            try:
                extras[name] = fn()
            except Exception as e:  # an extra workload never hides the headline line
                extras[name] = {"error": repr(e)}
                print(f"[bench] {name} failed: {e!r}", file=sys.stderr, flush=True)

    other = {p: {"iters_per_s": round(1000.0 * b["iterations"] / b["window_ms"], 1),
                 "primal_kernel_us": round(b["primal_kernel_ms"] * 1e3, 2),
                 "dual_kernel_us": round(b["dual_kernel_ms"] * 1e3, 2)} for p, b in windows.items()}
    speedup = loop_ratio = obj_diff = None
    if "fp32_mixed" in solves and "fp64" in solves:
        speedup = round(solves["fp64"]["time_to_tol_s"] / solves["fp32_mixed"]["time_to_tol_s"], 3)
        loop_ratio = round(solves["fp64"]["loop_s"] / solves["fp32_mixed"]["loop_s"], 3)
        o64 = solves["fp64"]["objective"]
        obj_diff = abs(solves["fp32_mixed"]["objective"] - o64) / (1.0 + abs(o64))
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "PDHG iters/s", "n_gpus": 1, "steps": iters,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if head == "fp32_mixed" else "f64", "data": "synthetic",
        "config": {"workload": "skewed LP m=500000 n=1000000 nnz=%d, Zipf(1.1) rows, 1+Poisson(4) per column, "
                               "half = / half <=, seed 0" % lp.nnz, "precision": head, "parallelism": "1 GPU"},
        "windows": other, "time_to_tolerance": solves, "fp64_over_fp32_time_ratio": speedup,
        "fp64_over_fp32_loop_ratio": loop_ratio,
        "fp32_vs_fp64_objective_rel_diff": obj_diff, "spmv_autotune": tuning or None,
        "roofline": roofline, "stream": stream, "cpu_baseline": cpu, "other_workloads": extras,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
