"""A C99 program compiled against include/cuopt_c.h and linked to libcuopt_amd.so
(tests/c_driver/c_api_driver.c) runs the reference C API test cases
(cpp/tests/linear_programming/c_api_tests/c_api_test.c, expectations from
c_api_tests.cpp).  The CPU cases need no GPU; the solve cases are marked gpu and use
the driver that build() compiled."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from cuopt_amd import _build  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(driver, *args, timeout=120):
    r = subprocess.run([driver, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    return r


@pytest.fixture(scope="module")
def fresh_driver(tmp_path_factory):
    """Compiled here from the source with the same cc line as build(): the header and the
    library's exported symbols are what a reference C program needs."""
    if not os.path.exists(os.path.join(ROOT, "cuopt_amd", "libcuopt_amd.so")):
        pytest.skip("libcuopt_amd.so is not built")
    out = str(tmp_path_factory.mktemp("c_driver") / "c_api_driver")
    r = subprocess.run(_build.c_driver_command(_build.C_DRIVER_SRC, out), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


@pytest.mark.parametrize("case", ["sizes", "bad_parameter", "missing_file", "reference_parameters", "infeasible"])
def test_c_driver_cpu_cases(fresh_driver, case):
    assert run(fresh_driver, case).returncode == 0


@pytest.mark.parametrize("fname", ["afiro_original.mps", "good-mps-fixed-ranges.mps", "lp_model_with_var_bounds.mps",
                                   "max_offset.mps"])
def test_c_driver_write_read(fresh_driver, tmp_path, fname):
    assert run(fresh_driver, "write_read", os.path.join(GOLDEN, fname), str(tmp_path / "out.mps")).returncode == 0


def test_c_driver_header_compiles_as_cxx(tmp_path):
    # The reference header serves C++ callers too (CUOPT_INFINITY via <limits>).
    src = tmp_path / "h.cpp"
    src.write_text('#include "cuopt_c.h"\nint main() { return CUOPT_INFINITY > 0 ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), str(src), "-o",
                        str(tmp_path / "h")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["ranged", "maximize_duals"])
def test_c_driver_gpu_solves(case):
    assert os.path.exists(_build.C_DRIVER), "run __graft_entry__.build() first"
    print(f"c_api_driver {case}", flush=True)
    assert run(_build.C_DRIVER, case).returncode == 0


@pytest.mark.gpu
def test_c_driver_afiro():
    assert os.path.exists(_build.C_DRIVER), "run __graft_entry__.build() first"
    print("c_api_driver afiro", flush=True)
    assert run(_build.C_DRIVER, "afiro", os.path.join(GOLDEN, "afiro_original.mps")).returncode == 0
