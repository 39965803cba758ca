"""Compiles libcuopt_amd.so (HIP for gfx950 + C++ host code) in-tree with hipcc, and the
plain C caller of its C ABI that the tests run (tests/c_driver)."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcuopt_amd.so")
OBJ = os.path.join(HERE, "_build")

HIP_SOURCES = ["pdlp_solver.hip", "batch_solver.hip", "unrolled.hip"]
CXX_SOURCES = ["tile_plan.cpp", "settings.cpp", "mps_io.cpp", "batch_size.cpp", "presolve.cpp", "cuopt_c.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]
# gfx950 only; packed FP32 instructions disabled (they are never emitted by hand here).
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def _rocm_include():
    root = os.environ.get("ROCM_PATH") or os.path.dirname(os.path.dirname(os.path.realpath(_hipcc()))) \
        if os.path.isabs(_hipcc()) else "/opt/rocm"
    return os.path.join(root, "include")


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        jobs.append([hipcc, *COMMON, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in CXX_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        # host-only C++ (hipcc would otherwise compile .cpp files as HIP for a default GPU)
        jobs.append([hipcc, *COMMON, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + _rocm_include(), "-c",
                     os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    procs = [(j, subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for j in jobs]
    for j, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if verbose and text.strip():
            print(text)
        if p.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(j)}\n{text}")
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB, *objs]
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(link)}\n{r.stdout.decode(errors='replace')}")
    return LIB


REPO = os.path.dirname(HERE)
C_DRIVER_SRC = os.path.join(REPO, "tests", "c_driver", "c_api_driver.c")
C_DRIVER = os.path.join(REPO, "tests", "c_driver", "build", "c_api_driver")


def c_driver_command(src: str, out: str) -> list:
    """cc line of a C99 program against include/cuopt_c.h, linked to libcuopt_amd.so."""
    return ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(REPO, "include"), src,
            "-o", out, "-L" + HERE, "-lcuopt_amd", "-Wl,-rpath," + HERE, "-lm"]


def build_c_driver() -> str:
    os.makedirs(os.path.dirname(C_DRIVER), exist_ok=True)
    r = subprocess.run(c_driver_command(C_DRIVER_SRC, C_DRIVER), stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("C driver build failed:\n" + r.stdout.decode(errors="replace"))
    return C_DRIVER
