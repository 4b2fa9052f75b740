"""PyTorch integration (drop-in for ``kernelkit/torch_support.py``).

``AutogradOperator`` (:10-34) applies an operator to ``(B, C, *domain)``
tensors and uses the operator's ``.T`` in the backward pass -- for
``XrayTransform`` that is the voxel-driven backprojector, which is *not* the
exact transpose of the ray-driven forward projector.
``XrayForwardProjection`` / ``XrayBackprojection`` (:37-56) wrap it as
modules.

Added beside the reference's API (no counterpart there):

* ``MatchedAutogradOperator`` -- the same contract, but the backward pass
  applies the exact transpose ``XrayTransform.matched_T`` (and the
  transpose's backward applies the forward projection), so gradients are
  exact; ``XrayForwardProjection(..., matched=True)`` selects it;
* ``unrolled_reconstruction`` -- differentiable unrolled gradient-descent /
  SIRT-weighted iterations with checkpointed recomputation of the
  intermediate volumes (only the iterates' inputs are kept between the
  forward and the backward pass).
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


class MatchedAutogradOperator(Function):
    """``AutogradOperator`` with the exact adjoint in the backward pass.

    ``op`` is an ``XrayTransform`` (backward: ``op.matched_T``) or its
    ``matched_T`` (backward: the forward projection ``op.matched_T.T``).
    """

    @staticmethod
    def forward(ctx, input, op: BaseOperator):
        B, C = input.shape[:2]
        ctx.operator = op
        out = input.new_empty(B, C, *op.range_shape, dtype=torch.float32)
        x = input.detach()
        for b, c in product(range(B), range(C)):
            op(x[b, c].contiguous(), out=out[b, c])
        return out

    @staticmethod
    def backward(ctx, grad_output):
        op = ctx.operator
        adj = op.matched_T if hasattr(op, "matched_T") else op.T
        B, C = grad_output.shape[:2]
        grad_input = grad_output.new_empty(B, C, *op.domain_shape,
                                           dtype=torch.float32)
        g = grad_output.detach()
        for b, c in product(range(B), range(C)):
            adj(g[b, c].contiguous(), out=grad_input[b, c])
        return grad_input, None


class XrayForwardProjection(torch.nn.Module):
    """Module applying ``XrayTransform`` (``torch_support.py:37-45``).

    ``matched=True`` (not in the reference) back-propagates through the
    exact transpose instead of the voxel-driven backprojector.
    """

    def __init__(self, *args, matched: bool = False, **kwargs):
        super().__init__()
        self.op = XrayTransform(*args, **kwargs)
        self.matched = matched

    def forward(self, input):
        if self.matched:
            return MatchedAutogradOperator.apply(input, self.op)
        return AutogradOperator.apply(input, self.op)


class XrayBackprojection(torch.nn.Module):
    """Module applying ``XrayTransform.T`` (``torch_support.py:48-56``)."""

    def __init__(self, *args, **kwargs):
        super().__init__()
        self.op = XrayTransform(*args, **kwargs)

    def forward(self, input):
        return AutogradOperator.apply(input, self.op.T)


def _apply(fn, x, op):
    """``op`` on one volume/sinogram through an autograd function."""
    return fn.apply(x[None, None], op)[0, 0]


def unrolled_reconstruction(y, op: XrayTransform, iters: int = 5, x0=None,
                            step=None, weighting: str = "sirt",
                            min_constraint: float | None = None,
                            checkpoint: bool = True, callback=None):
    """Differentiable unrolled reconstruction ``x_{k+1} = P(x_k - s_k W
    A^T(R (A x_k - y)))`` with the exact transpose ``A^T``.

    ``weighting='sirt'`` uses the SIRT weights ``W = 1/A^T 1`` and
    ``R = 1/A 1`` (zero where infinite, as ``algo.sirt``) with unit steps;
    ``'none'`` is plain gradient descent on ``||A x - y||^2 / 2`` (then
    ``step`` is required).  ``step`` may be a float, a tensor of ``iters``
    (learnable) step sizes, or None.  ``P`` clamps at ``min_constraint``
    when given.  With ``checkpoint`` every iteration runs under
    ``torch.utils.checkpoint``: its intermediate projections and volumes
    are recomputed in the backward pass instead of being stored, so the
    memory held for the backward pass is one volume per iteration.
    Gradients flow to ``y``, ``x0`` and ``step``.  ``callback(k, x_k)``,
    when given, is called after iteration ``k`` (as ``algo.sirt``'s
    callback, ``algo.py:363-364``).
    """
    from torch.utils.checkpoint import checkpoint as _ckpt
    dev = y.device
    if weighting not in ("sirt", "none"):
        raise ValueError("weighting must be 'sirt' or 'none'")
    if weighting == "none" and step is None:
        raise ValueError("plain gradient descent needs a step size")
    AT = op.matched_T
    if weighting == "sirt":
        with torch.no_grad():
            R = op(torch.ones(op.domain_shape, device=dev))
            W = AT(torch.ones(op.range_shape, device=dev))
            R = torch.where(R > 0, 1.0 / R, torch.zeros_like(R))
            W = torch.where(W > 0, 1.0 / W, torch.zeros_like(W))
    else:
        R = W = None
    if step is None:
        steps = [1.0] * iters
    elif isinstance(step, torch.Tensor) and step.ndim == 1:
        if step.numel() != iters:
            raise ValueError("need one step size per iteration")
        steps = [step[k] for k in range(iters)]
    else:
        steps = [step] * iters
    x = (torch.zeros(op.domain_shape, device=dev) if x0 is None else x0)

    def one(x, y, s):
        r = _apply(MatchedAutogradOperator, x, op) - y
        if R is not None:
            r = r * R
        g = _apply(MatchedAutogradOperator, r, AT)
        if W is not None:
            g = g * W
        x = x - s * g
        if min_constraint is not None:
            x = torch.clamp(x, min=min_constraint)
        return x

    for k in range(iters):
        s = steps[k]
        if not isinstance(s, torch.Tensor):
            s = torch.tensor(float(s), device=dev)
        if checkpoint:
            x = _ckpt(one, x, y, s, use_reentrant=False)
        else:
            x = one(x, y, s)
        if callback is not None:
            callback(k, x)
    return x
