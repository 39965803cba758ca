"""Differentiable unrolled PDHG as a torch.autograd.Function.

``UnrolledPDHG(lp, ridge, steps, checkpoint_every, precision)`` binds one
constraint matrix and variable bounds; calling it with a batch of objectives
``c`` (B x n) and constraint bounds ``lc``, ``uc`` (m, or B x m) returns
``x_K`` (B x n) and ``y_K`` (B x m) after ``steps`` PDHG steps

    x+ = clamp((x - tau (c - A^T y)) / (1 + tau ridge), l, u)
    y+ = sigma (t - clamp(t, -uc, -lc)),   t = y / sigma - A (2 x+ - x)

on the GPU (cuOptAmdUnrolledForward), and its backward is the hand-written
adjoint (cuOptAmdUnrolledBackward): transposed SpMV plus projection masks,
checkpointed every ``checkpoint_every`` steps, gradients accumulated in FP64
(in FP32 mode each checkpoint segment is summed in FP32 and added into FP64).
An equality-constrained problem passes the same tensor ``b`` as ``lc`` and
``uc``; autograd then adds both gradients into ``b``.

The library must be built (``__graft_entry__.build()``); there is no CPU path.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import c_api

PRECISION = {"fp64": 0, "fp32_mixed": 1, "fp32": 1}


def _lib():
    L = c_api.lib()
    if not getattr(L, "_unrolled_bound", False):
        vp, i32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        L.cuOptAmdUnrolledCreate.argtypes = [i32, i32, vp, vp, vp, vp, vp, f64, i32, ctypes.POINTER(vp)]
        L.cuOptAmdUnrolledDestroy.argtypes = [ctypes.POINTER(vp)]
        L.cuOptAmdUnrolledDestroy.restype = None
        L.cuOptAmdUnrolledStepSize.argtypes = [vp, ctypes.POINTER(f64)]
        L.cuOptAmdUnrolledForward.argtypes = [vp, i32, i32, i32, vp, vp, vp, i32, vp, vp, vp]
        L.cuOptAmdUnrolledBackward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.cuOptAmdUnrolledOptimalBatchSize.argtypes = [vp, i32, ctypes.POINTER(i32)]
        L._unrolled_bound = True
    return L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class UnrolledPDHG:
    """A K-step PDHG map on one constraint matrix; see the module docstring."""

    def __init__(self, lp, ridge: float = 1e-3, steps: int = 2000, checkpoint_every: int = 45,
                 precision: str = "fp32_mixed", device: str = "cuda"):
        L = _lib()
        self.m, self.n = lp.m, lp.n
        self.steps, self.every = int(steps), int(checkpoint_every)
        self.device = torch.device(device)
        self.handle = ctypes.c_void_p()
        ptr = np.ascontiguousarray(lp.row_offsets, np.int32)
        idx = np.ascontiguousarray(lp.col_indices, np.int32)
        val = np.ascontiguousarray(lp.values, np.float64)
        lb = np.ascontiguousarray(lp.var_lb, np.float64)
        ub = np.ascontiguousarray(lp.var_ub, np.float64)
        rc = L.cuOptAmdUnrolledCreate(self.m, self.n, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data,
                                      lb.ctypes.data, ub.ctypes.data, float(ridge), PRECISION[precision],
                                      ctypes.byref(self.handle))
        c_api._check(rc, "cuOptAmdUnrolledCreate")
        self._generation = 0

    @property
    def step_size(self) -> float:
        eta = ctypes.c_double()
        c_api._check(_lib().cuOptAmdUnrolledStepSize(self.handle, ctypes.byref(eta)))
        return eta.value

    def optimal_batch_size(self, max_batch: int) -> int:
        """Batch size in [1, max_batch] with the least SpMM time per batch member on this
        matrix (cuOptAmdUnrolledOptimalBatchSize); larger batches are best run in chunks
        of that size."""
        out = ctypes.c_int32()
        c_api._check(_lib().cuOptAmdUnrolledOptimalBatchSize(self.handle, int(max_batch), ctypes.byref(out)),
                     "cuOptAmdUnrolledOptimalBatchSize")
        return out.value

    def __call__(self, c: torch.Tensor, lc: torch.Tensor, uc: torch.Tensor | None = None):
        return _UnrolledFn.apply(c, lc, lc if uc is None else uc, self)

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            _lib().cuOptAmdUnrolledDestroy(ctypes.byref(self.handle))


class _UnrolledFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, lc, uc, op: UnrolledPDHG):
        if c.dim() != 2 or c.shape[1] != op.n:
            raise ValueError("c must be (batch, n)")
        B = c.shape[0]
        per_batch = lc.dim() == 2
        if uc.dim() != lc.dim():
            raise ValueError("lc and uc must have the same shape")
        dev = op.device
        c_nb = c.detach().to(dev, torch.float64).t().contiguous()
        lc_d = lc.detach().to(dev, torch.float64)
        uc_d = uc.detach().to(dev, torch.float64)
        lc_d = lc_d.t().contiguous() if per_batch else lc_d.contiguous()
        uc_d = uc_d.t().contiguous() if per_batch else uc_d.contiguous()
        x = torch.empty(op.n, B, dtype=torch.float64, device=dev)
        y = torch.empty(op.m, B, dtype=torch.float64, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        rc = _lib().cuOptAmdUnrolledForward(op.handle, B, op.steps, op.every, _ptr(c_nb), _ptr(lc_d), _ptr(uc_d),
                                             int(per_batch), _ptr(x), _ptr(y), stream)
        c_api._check(rc, "cuOptAmdUnrolledForward")
        op._generation += 1
        ctx.op, ctx.generation, ctx.per_batch, ctx.B = op, op._generation, per_batch, B
        return x.t(), y.t()

    @staticmethod
    def backward(ctx, gx, gy):
        op = ctx.op
        if op._generation != ctx.generation:
            raise RuntimeError("UnrolledPDHG: backward must follow its own forward (the checkpoints were replaced)")
        dev = op.device
        B = ctx.B
        gx_nb = (gx if gx is not None else torch.zeros(B, op.n, dtype=torch.float64, device=dev))
        gx_nb = gx_nb.to(dev, torch.float64).t().contiguous()
        gy_nb = gy.to(dev, torch.float64).t().contiguous() if gy is not None else None
        gc = torch.empty(op.n, B, dtype=torch.float64, device=dev)
        glc = torch.empty(op.m, B, dtype=torch.float64, device=dev)
        guc = torch.empty(op.m, B, dtype=torch.float64, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        rc = _lib().cuOptAmdUnrolledBackward(op.handle, _ptr(gx_nb), _ptr(gy_nb), _ptr(gc), _ptr(glc), _ptr(guc),
                                              stream)
        c_api._check(rc, "cuOptAmdUnrolledBackward")
        if ctx.per_batch:
            return gc.t(), glc.t(), guc.t(), None
        return gc.t(), glc.sum(dim=1), guc.sum(dim=1), None
