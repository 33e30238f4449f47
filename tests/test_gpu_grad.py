"""GPU: adjoint of the intersection area (workload: independent star pairs,
16..128 vertices, unit radius, centres offset U[-0.3, 0.3]).  Checked
against the FP64 CPU oracle gradient, against central finite differences
(h = 1e-6) of the GPU forward area, FP32 against FP64, and through
torch.autograd."""
import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import autograd, gen, rtree, workloads

pytestmark = pytest.mark.gpu


def _pairs(n):
    pairs, S = workloads.w3(n)
    return pairs, S


def test_fp64_gradient_matches_oracle(dev):
    pairs, S = _pairs(2048)
    Sd = rtree.to_device_polys(S)
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    g1, g2, green = rtree.area_grad(pd, Sd, Sd, green=True)
    g = (g1 + g2).cpu().numpy()         # A and B are the same CSR set here
    ref = np.zeros_like(S.xy)
    for t, (a, b) in enumerate(pairs):
        gP, gQ = O.area_grad(S.poly(a), S.poly(b))
        ref[S.off[a]:S.off[a + 1]] += gP
        ref[S.off[b]:S.off[b + 1]] += gQ
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max() / scale
    assert err <= 1e-12, err
    exact = np.mean(g == ref)
    assert exact > 0.99, f"only {exact:.4f} of the gradient entries are bit-identical"
    # Green's area of the adjoint kernel agrees with the scanbeam area
    area = rtree.overlay_batched(pd, Sd, Sd).cpu().numpy()
    ga = green.cpu().numpy()
    rel = np.abs(ga - area) / np.maximum(np.abs(area), 1e-12)
    assert np.median(rel) < 1e-13 and rel.max() < 1e-9


def test_fp64_gradient_matches_finite_differences(dev):
    """Central differences with h = 1e-6 of the GPU forward area, on 160
    random pairs of the workload (every perturbed polygon is one more pair of
    a single batched overlay call), and on 12 pairs with the perturbed
    polygons built here one by one (no shared helper)."""
    pairs, S = _pairs(4096)
    sel = np.sort(np.random.default_rng(11).choice(len(pairs), 160, replace=False))
    Sd = rtree.to_device_polys(S)
    g1, g2 = rtree.area_grad(torch.from_numpy(pairs.view(np.int32)).cuda(), Sd, Sd)
    g = (g1 + g2).cpu().numpy()
    fd = autograd.finite_difference_grad(pairs[sel], S, 1e-6)
    errs = []
    for (a, b), (fP, fQ) in zip(pairs[sel], fd):
        for F, idx in ((fP, a), (fQ, b)):
            G = g[S.off[idx]:S.off[idx + 1]]
            errs.append(np.abs(F - G).max() / np.abs(G).max())
    assert max(errs) < 1e-5, max(errs)
    assert np.median(errs) < 1e-7, np.median(errs)

    pairs, S = _pairs(12)
    h = 1e-6
    meta = []
    polys = [S.poly(i) for i in range(S.n)]
    extra = []
    for t, (a, b) in enumerate(pairs):
        for side, idx in ((0, a), (1, b)):
            base = polys[idx]
            for v in range(len(base)):
                for c in range(2):
                    for sgn in (1.0, -1.0):
                        q = base.copy()
                        q[v, c] += sgn * h
                        extra.append(q)
                        meta.append((t, side, v, c, sgn))
    allp = polys + extra
    off = np.zeros(len(allp) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in allp])
    C = gen.PolyCSR(np.ascontiguousarray(np.concatenate(allp)), off)
    prs = []
    for k, (t, side, v, c, sgn) in enumerate(meta):
        a, b = pairs[t]
        pid = len(polys) + k
        prs.append((pid, b) if side == 0 else (a, pid))
    prs = np.array(prs, np.uint32)
    Cd = rtree.to_device_polys(C)
    areas = rtree.overlay_batched(torch.from_numpy(prs.view(np.int32)).cuda(), Cd, Cd).cpu().numpy()
    fdd = {}
    for k, (t, side, v, c, sgn) in enumerate(meta):
        fdd[(t, side, v, c)] = fdd.get((t, side, v, c), 0.0) + sgn * areas[k] / (2 * h)
    Sd = rtree.to_device_polys(S)
    g1, g2 = rtree.area_grad(torch.from_numpy(pairs.view(np.int32)).cuda(), Sd, Sd)
    g = (g1 + g2).cpu().numpy()
    errs = []
    for t, (a, b) in enumerate(pairs):
        for side, idx in ((0, a), (1, b)):
            G = g[S.off[idx]:S.off[idx + 1]]
            F = np.array([[fdd[(t, side, v, c)] for c in range(2)] for v in range(len(G))])
            errs.append(np.abs(F - G).max() / np.abs(G).max())
    assert max(errs) < 1e-5, max(errs)


def test_fp32_against_fp64(dev):
    pairs, S = _pairs(4096)
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    S64 = rtree.to_device_polys(S)
    S32 = rtree.to_device_polys(S, dtype=torch.float32)
    a64, b64, ar64 = rtree.area_grad(pd, S64, S64, green=True)
    a32, b32, ar32 = rtree.area_grad(pd, S32, S32, green=True)
    g64 = (a64 + b64).cpu().numpy()
    g32 = (a32 + b32).double().cpu().numpy()
    assert np.abs(g32 - g64).max() / np.abs(g64).max() < 1e-4
    rel = np.abs(ar32.double().cpu().numpy() - ar64.cpu().numpy()) / np.abs(ar64.cpu().numpy())
    assert np.median(rel) < 1e-5


def test_autograd_function(dev):
    pairs, S = _pairs(64)
    xy = torch.from_numpy(S.xy).cuda().requires_grad_(True)
    off = torch.from_numpy(S.off).cuda()
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    w = torch.linspace(0.5, 1.5, len(pairs), dtype=torch.float64, device="cuda")
    area = autograd.overlay_area(xy, off, xy, off, pd)
    (area * w).sum().backward()
    Sd = rtree.to_device_polys(S)
    g1, g2 = rtree.area_grad(pd, Sd, Sd, weights=w)
    torch.testing.assert_close(xy.grad, g1 + g2, rtol=0, atol=0)
    ref = O.overlay(pairs, S, S)
    assert np.array_equal(area.detach().cpu().numpy(), ref)


def test_autograd_gradcheck_small(dev):
    pairs, S = _pairs(3)
    xy = torch.from_numpy(S.xy).cuda().requires_grad_(True)
    off = torch.from_numpy(S.off).cuda()
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    f = lambda x: autograd.overlay_area(x, off, x, off, pd)  # noqa: E731
    assert torch.autograd.gradcheck(f, (xy,), eps=1e-6, atol=1e-6, rtol=1e-4)


def _comb_pairs(n, seed=5):
    """Each pair: a comb of 12 teeth and a quadrilateral whose lower edge
    crosses the teeth (up to 24 crossings on one edge), at random heights and
    slopes; the lower edge sometimes runs along the comb's tooth tops or
    through its vertices."""
    rng = np.random.default_rng(seed)
    top = []
    for k in range(12):
        top += [(k, 1.0), (k + 0.5, 1.0), (k + 0.5, 0.2), (k + 1, 0.2)]
    comb = [(0.0, 0.0), (12.0, 0.0)] + top[::-1]
    A, B = [], []
    for i in range(n):
        y = [1.0, 0.2, 0.5][i % 3] if i % 4 == 0 else rng.uniform(0.1, 1.1)
        x0, x1 = rng.uniform(-1, 3), rng.uniform(9, 13)
        y1 = y if i % 4 == 0 else y + rng.uniform(-0.3, 0.3)
        A.append([(x0, y), (x1, y1), (x1, 2.0), (x0, 2.0)])
        B.append(comb)
    return A, B


def test_fp64_gradient_bit_exact_on_many_crossings_and_degenerate(dev):
    """Edges that cross the other polygon more than four times (the kernel
    keeps up to four crossings per edge in registers and takes a pass per
    crossing beyond), and lattice polygons with shared vertices, collinear
    and coincident edges: every gradient entry equals the oracle's."""
    from test_gpu_degenerate import csr, lattice_polys
    A, B = _comb_pairs(300)
    A += lattice_polys(11, 400)
    B += lattice_polys(12, 400)
    L = lattice_polys(13, 100)
    A += L
    B += L                                    # coincident boundaries
    CA, CB = csr(A), csr(B)
    pairs = np.stack([np.arange(len(A)), np.arange(len(A))], 1).astype(np.uint32)
    dA, dB = rtree.to_device_polys(CA), rtree.to_device_polys(CB)
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    gA, gB, green = rtree.area_grad(pd, dA, dB, green=True)
    gA, gB, green = gA.cpu().numpy(), gB.cpu().numpy(), green.cpu().numpy()
    for k in range(len(A)):
        P, Q = CA.poly(k), CB.poly(k)
        rP, rQ = O.area_grad(P, Q)
        assert np.array_equal(gA[CA.off[k]:CA.off[k + 1]], rP), k
        assert np.array_equal(gB[CB.off[k]:CB.off[k + 1]], rQ), k
        ga = O.green_area(P, Q)
        assert abs(green[k] - ga) <= 1e-12 * max(abs(ga), 1.0), (k, green[k], ga)
