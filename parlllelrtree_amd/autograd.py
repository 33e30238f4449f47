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


def finite_difference_grad(pairs, S, h: float = 1e-6, device="cuda"):
    """Central differences (A(v + h) - A(v - h)) / 2h of the FP64 area of
    every pair (a, b) of `pairs` (into the CSR polygon set S) with respect to
    every vertex coordinate of both polygons, all from ONE batched overlay
    call (each perturbed polygon is paired with the other, unperturbed one).
    Returns a list with, per pair, (dA/dP of shape (nP, 2), dA/dQ (nQ, 2))."""
    import numpy as np

    from .gen import PolyCSR

    polys, prs, meta = [], [], []
    nrows = 0

    def add(p):
        polys.append(p)
        return len(polys) - 1

    for t, (a, b) in enumerate(pairs):
        P, Q = S.poly(int(a)), S.poly(int(b))
        ip, iq = add(P), add(Q)
        for side, base in ((0, P), (1, Q)):
            n = len(base)
            cp = np.repeat(base[None], 4 * n, axis=0)          # (4n, n, 2)
            k = np.arange(4 * n)
            v, c, sgn = k // 4, (k // 2) % 2, np.where(k % 2 == 0, 1.0, -1.0)
            cp[k, v, c] += sgn * h
            first = len(polys)
            polys.extend(cp)
            ids = np.arange(first, first + 4 * n)
            other = iq if side == 0 else ip
            prs.append(np.stack([ids, np.full(4 * n, other)] if side == 0 else
                                [np.full(4 * n, other), ids], 1))
            meta.append((t, side, n, nrows))
            nrows += 4 * n
    off = np.zeros(len(polys) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in polys])
    C = PolyCSR(np.ascontiguousarray(np.concatenate(polys)), off)
    allp = np.ascontiguousarray(np.concatenate(prs).astype(np.uint32))
    Cd = rtree.to_device_polys(C, device)
    area = rtree.overlay_batched(torch.from_numpy(allp.view(np.int32)).to(device), Cd, Cd)
    area = area.cpu().numpy()
    out = [[None, None] for _ in range(len(pairs))]
    for t, side, n, first in meta:
        a4 = area[first:first + 4 * n].reshape(n, 2, 2)     # (vertex, coord, +h / -h)
        out[t][side] = (a4[:, :, 0] - a4[:, :, 1]) / (2 * h)
    return [tuple(o) for o in out]
