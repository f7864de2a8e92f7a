"""Differentiable projection: the forward projection as a torch.autograd.Function whose
backward is the projector's own matched adjoint kernel.

  project(op, x)    A x for a volume tensor x (requires_grad allowed); d/dx <g, A x> = A^T g
  backproject(op, y)  A^T y for a sinogram tensor y; its backward is the forward projection

op is a projector of elsa_amd.projectors (JosephsMethodCUDA and the others). Tensors live on
the GPU with the operator's dtype and shape (volume ([z,] y, x), sinogram (poses, [v,] u)). A
leading batch dimension of size n is applied as one ProjectorBatch of op and n - 1 clones of it
(made once per op and n, kept while op lives): for the operators whose batches fuse (FP32
Joseph, Siddon and voxel projectors with the default kernels) every pass over the whole batch is
one launch per kernel, each element's result bit for bit op's own. Other operators are applied
element by element. There is no CPU path.
"""
from __future__ import annotations

import weakref

import torch

from .batch import ProjectorBatch

# op -> {n: ProjectorBatch of op and n - 1 clones}
_batches: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def _batch_of(op, n):
    per_op = _batches.setdefault(op, {})
    B = per_op.get(n)
    if B is None:
        B = ProjectorBatch([op] + [op.clone() for _ in range(n - 1)])
        per_op[n] = B
    return B


def _apply(op, t, adjoint):
    shape_out = op.getDomainDescriptor().shape if adjoint else op.getRangeDescriptor().shape
    inner = len(op.getRangeDescriptor().shape if adjoint else op.getDomainDescriptor().shape)
    t = t.contiguous()
    if t.dim() == inner:
        return op.applyAdjoint(t) if adjoint else op.apply(t)
    if t.dim() != inner + 1:
        raise ValueError(f"expected a tensor of {inner} or {inner + 1} dimensions, got {t.dim()}")
    n = t.shape[0]
    if n > 1 and hasattr(op, "clone"):
        B = _batch_of(op, n)
        if B.fused:
            flat = B.applyAdjoint(t.reshape(-1)) if adjoint else B.apply(t.reshape(-1))
            return flat.reshape((n,) + tuple(shape_out))
    out = torch.empty((n,) + tuple(shape_out), dtype=t.dtype, device=t.device)
    for i in range(n):
        if adjoint:
            op.applyAdjoint(t[i], out[i])
        else:
            op.apply(t[i], out[i])
    return out


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op):
        ctx.op = op
        return _apply(op, x.detach(), adjoint=False)

    @staticmethod
    def backward(ctx, g):
        return _Project_adjoint.apply(g, ctx.op), None


class _Project_adjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, op):
        ctx.op = op
        return _apply(op, y.detach(), adjoint=True)

    @staticmethod
    def backward(ctx, g):
        return _Project.apply(g, ctx.op), None


def project(op, x):
    """A x, differentiable in x (the gradient is the matched back projection)."""
    return _Project.apply(x, op)


def backproject(op, y):
    """A^T y, differentiable in y (the gradient is the forward projection)."""
    return _Project_adjoint.apply(y, op)
