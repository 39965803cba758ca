"""MI355X-native PDLP LP engine behind the cuOpt C API.

The compute path is libcuopt_amd.so (hand-written HIP kernels for gfx950 plus a
C++ host driver).  This package loads it (``c_api``), mirrors the reference's
Python LP interface over it (``linear_programming``), wraps the differentiable
unrolled PDHG for PyTorch (``autograd``) and provides LP generators and an MPS
reader; there is no Python or CPU fallback solver.
"""

__all__ = ["c_api", "linear_programming", "autograd", "generators", "mps"]
