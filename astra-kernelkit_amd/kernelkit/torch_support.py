"""PyTorch integration (drop-in for ``kernelkit/torch_support.py``).

``AutogradOperator`` (:10-34) applies an operator to ``(B, C, *domain)``
tensors and uses the operator's ``.T`` in the backward pass -- for
``XrayTransform`` that is the voxel-driven backprojector, which is *not* the
exact transpose of the ray-driven forward projector.
``XrayForwardProjection`` / ``XrayBackprojection`` (:37-56) wrap it as
modules.
"""
from __future__ import annotations

from itertools import product

import torch
from torch.autograd import Function

from kernelkit.operator import BaseOperator, XrayTransform


class AutogradOperator(Function):
    """Autograd function for ``Operator`` objects (unmatched adjoint)."""

    @staticmethod
    def forward(ctx, input, op: BaseOperator):
        B, C = input.shape[:2]
        ctx.operator = op
        out = input.new_empty(B, C, *op.range_shape, dtype=torch.float32)
        x = input.detach()
        for b, c in product(range(B), range(C)):
            op(x[b, c], out=out[b, c])
        return out

    @staticmethod
    def backward(ctx, grad_output):
        op = ctx.operator
        B, C = grad_output.shape[:2]
        grad_input = grad_output.new_zeros(B, C, *op.domain_shape,
                                           dtype=torch.float32)
        g = grad_output.detach()
        for b, c in product(range(B), range(C)):
            op.T(g[b, c].contiguous(), out=grad_input[b, c])
        return grad_input, None


class XrayForwardProjection(torch.nn.Module):
    """Module applying ``XrayTransform`` (``torch_support.py:37-45``)."""

    def __init__(self, *args, **kwargs):
        super().__init__()
        self.op = XrayTransform(*args, **kwargs)

    def forward(self, input):
        return AutogradOperator.apply(input, self.op)


class XrayBackprojection(torch.nn.Module):
    """Module applying ``XrayTransform.T`` (``torch_support.py:48-56``)."""

    def __init__(self, *args, **kwargs):
        super().__init__()
        self.op = XrayTransform(*args, **kwargs)

    def forward(self, input):
        return AutogradOperator.apply(input, self.op.T)
