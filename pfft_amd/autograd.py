"""Thin torch.autograd.Function over a pfft plan and its hand-written
adjoint (pfft_execute_adjoint), used to check gradients.

Forward runs the plan through the C ABI on the tensor's device memory;
backward runs the plan's adjoint pipeline (conjugate-transposed passes:
reversed order, opposite sign, pad <-> truncate swapped, r2c <-> c2r with
the Hermitian weights).  Gradients follow torch's convention for complex
tensors: grad = dL/dRe + i dL/dIm, i.e. the adjoint with respect to the
real inner product on interleaved (re, im) storage.

The library runs on its own HIP stream and every execute call returns
after that stream has finished, so the wrapper synchronises torch's
stream before handing it memory.
"""
import torch

import pfft_amd as P


def _complex_dtype(prec):
    return torch.complex128 if prec == "f64" else torch.complex64


def _real_dtype(prec):
    return torch.float64 if prec == "f64" else torch.float32


class PlanFunction(torch.autograd.Function):
    """y = A x for a fixed plan A (c2c, r2c or c2r)."""

    @staticmethod
    def forward(ctx, x, plan, out_shape, out_complex):
        x = x.contiguous()
        dt = _complex_dtype(plan.prec) if out_complex else _real_dtype(plan.prec)
        y = torch.empty(out_shape, dtype=dt, device=x.device)
        torch.cuda.synchronize(x.device)
        plan.execute_on(x.data_ptr(), y.data_ptr())
        ctx.plan = plan
        ctx.in_shape = x.shape
        ctx.in_dtype = x.dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        gy = gy.contiguous()
        gx = torch.zeros(ctx.in_shape, dtype=ctx.in_dtype, device=gy.device)
        torch.cuda.synchronize(gy.device)
        ctx.plan.execute_adjoint(gy.data_ptr(), gx.data_ptr())
        return gx, None, None, None


def apply_plan(x, plan, out_shape, out_complex):
    return PlanFunction.apply(x, plan, out_shape, out_complex)
