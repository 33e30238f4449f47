"""torch.autograd.Function around the overlay area and its hand-written
adjoint (libprt, through ctypes).  Used for gradient checks.

    area = overlay_area(xyA, offA, xyB, offB, pairs)   # (npairs,)
    area.sum().backward()                               # xyA.grad, xyB.grad

Forward: FP64 -> the scanbeam area of the batched kernel (bit-identical to
the CPU oracle); FP32 -> the Green's-theorem area the adjoint kernel
computes from the same boundary parts.  Backward: the adjoint kernel with
the incoming gradient as per-pair weights, crossings recomputed.
"""
from __future__ import annotations

import torch

from . import rtree


def _max_verts(off: torch.Tensor) -> int:
    return int((off[1:] - off[:-1]).max().item()) if off.numel() > 1 else 0


class OverlayArea(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyA, xyB, pairs, offA, offB):
        A = rtree.PolyDev(xyA.detach().contiguous(), offA, _max_verts(offA))
        B = rtree.PolyDev(xyB.detach().contiguous(), offB, _max_verts(offB))
        if xyA.dtype == torch.float64:
            area = rtree.overlay_batched(pairs, A, B)
        else:
            _, _, area = rtree.area_grad(pairs, A, B, green=True)
        ctx.save_for_backward(A.xy, B.xy, pairs, offA, offB)
        ctx.mv = (A.max_verts, B.max_verts)
        return area

    @staticmethod
    def backward(ctx, grad_area):
        xyA, xyB, pairs, offA, offB = ctx.saved_tensors
        A = rtree.PolyDev(xyA, offA, ctx.mv[0])
        B = rtree.PolyDev(xyB, offB, ctx.mv[1])
        gA, gB = rtree.area_grad(pairs, A, B, weights=grad_area)
        return gA, gB, None, None, None


def overlay_area(xyA: torch.Tensor, offA: torch.Tensor, xyB: torch.Tensor, offB: torch.Tensor,
                 pairs: torch.Tensor) -> torch.Tensor:
    return OverlayArea.apply(xyA, xyB, pairs, offA, offB)
