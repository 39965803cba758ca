"""The SpMV tile planner's host logic (cuopt_amd/csrc/tile_plan.cpp), compiled with
g++ beside a check program (tests/plan_check/plan_check.cpp): a plan built from the
column indices at the chunk midpoints alone equals the plan built from the whole
index array, tiles cover every nonzero once, and an incomplete sample is refused."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_plan_from_column_sample_equals_full_plan(tmp_path):
    exe = str(tmp_path / "plan_check")
    src = [os.path.join(ROOT, "tests", "plan_check", "plan_check.cpp"),
           os.path.join(ROOT, "cuopt_amd", "csrc", "tile_plan.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "cuopt_amd", "csrc"), *src,
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr
