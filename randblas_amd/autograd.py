"""Differentiable left sketch, for checking the adjoint kernels.

left_sketch(A, S) returns B = S * A for a torch tensor A of shape (m, n) or
(m,) on the GPU and a full DenseSkOp S of shape d x m.  Its backward pass is
dA = S^T * dB, computed by rbx_[sd]lskge3_adjoint, which regenerates S from
its Philox counters: nothing but the operator's descriptor is kept between
the forward and backward passes.  Both passes run the library's kernels on
row-major views (layout 'R'); torch only supplies memory and streams.
"""
from __future__ import annotations

import torch

from . import DenseSkOp, Error, lskge3, lskge3_adjoint


class LeftSketch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A: torch.Tensor, S: DenseSkOp) -> torch.Tensor:
        if not A.is_cuda:
            raise Error("left_sketch: A must be a GPU tensor")
        vec = A.dim() == 1
        A2 = (A.unsqueeze(1) if vec else A).contiguous()
        m, n = A2.shape
        if S.n_cols != m:
            raise Error(f"left_sketch: operator is {S.n_rows} x {S.n_cols}, A has {m} rows")
        d = S.n_rows
        B = torch.empty((d, n), dtype=A.dtype, device=A.device)
        lskge3("R", "N", "N", d, n, m, 1.0, S, 0, 0, A2, n, 0.0, B, n)
        ctx.S, ctx.vec, ctx.shape = S, vec, (d, n, m)
        return B.squeeze(1) if vec else B

    @staticmethod
    def backward(ctx, G: torch.Tensor):
        d, n, m = ctx.shape
        G2 = (G.unsqueeze(1) if ctx.vec else G).contiguous()
        dA = torch.empty((m, n), dtype=G.dtype, device=G.device)
        lskge3_adjoint("R", "N", d, n, m, 1.0, ctx.S, 0, 0, G2, n, 0.0, dA, n)
        return (dA.squeeze(1) if ctx.vec else dA), None


def left_sketch(A: torch.Tensor, S: DenseSkOp) -> torch.Tensor:
    return LeftSketch.apply(A, S)
