"""Test-only ``torch.autograd.Function`` over the C-ABI FIR and its adjoint
kernels.  It exists so the adjoints can be checked against torch's own
autograd; it is not part of the drop-in path (the reference has no autograd).

``b`` is in lfilter order (``y[n] = sum_k b[k] x[n-k]``).
"""
from __future__ import annotations

import torch

from . import device


class FIRFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, b, precision="f32"):
        x = x.contiguous()
        b = b.contiguous()
        ctx.save_for_backward(x, b)
        ctx.precision = precision
        return device.fir(x, b, precision)

    @staticmethod
    def backward(ctx, gy):
        x, b = ctx.saved_tensors
        gy = gy.contiguous()
        dx = device.vjp_dx(gy, b, ctx.precision) if ctx.needs_input_grad[0] else None
        db = device.vjp_db(gy, x, b.numel()) if ctx.needs_input_grad[1] else None
        return dx, db, None


def fir(x, b, precision="f32"):
    """Differentiable batched FIR on the GPU (test helper)."""
    return FIRFunction.apply(x, b, precision)
