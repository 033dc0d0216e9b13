"""GPU parity at the full benchmark workload sizes (needs an MI355X).

Every workload runs its full-size operator on the GPU; the FP64 oracle is
evaluated on sampled views (forward) or on a sinogram that is zero outside
the sampled views (adjoint), so the full-size launches are checked without
a full-size CPU run:

* fan-arc 512x512, 720 views, 736 bins, batch 16, planned batch path (the
  16-instance tile with the quad-major copies);
* 64 uneven fan-arc instances (N in {256,384,512}, V in [96,720],
  D = ceil(1.44 N)) in one launch per op and phase;
* cone-flat 256^3, 360 views, 512 x 384 flat detector, batch 2;
* helical 512x512x64, 3 turns of 1152 views, 736 x 64 curved detector,
  batch 4, through the projectorsp autograd Functions.

Bars: relative L2 <= 1e-5 vs the FP64 oracle (max-abs printed), adjoint
dot-product tests <= 1e-6 with FP64 inner products.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pridediff_amd.phantom import ellipses2d, ellipsoids3d

pytestmark = pytest.mark.gpu

PIX, DET, SID, SDD = 0.7, 1.2858, 595.0, 1085.6


def report(tag, got, want):
    r, m = O.rel_l2(got, want), O.max_abs(got, want)
    print(f"{tag}: rel_l2={r:.3e} max_abs={m:.3e}", flush=True)
    return r


def dot_rel(ax, y, x, aty):
    """|<Ax, y> - <x, A^T y>| / |<Ax, y>|, inner products in FP64 on the GPU."""
    lhs = sum(float((a.double() * b.double()).sum()) for a, b in zip(ax, y))
    rhs = sum(float((a.double() * b.double()).sum()) for a, b in zip(x, aty))
    return abs(lhs - rhs) / abs(lhs), lhs, rhs


def sparse_views(y, sel):
    """Copy of sinogram ``y`` (views on dim 2) that is zero outside ``sel``."""
    out = torch.zeros_like(y)
    out[:, :, sel] = y[:, :, sel]
    return out


# ------------------------------------------------- fan-arc 512^2, batch 16
def test_fanarc_batch16_planned_full_size():
    from pridediff_amd.batched import BatchedProjector2d
    from pridediff_amd.deeprecon.torch.projector import projector2d
    n, d, v, b = 512, 736, 720, 16
    P = projector2d(n, d, PIX, DET, SID, SDD, scan_type="arc").cuda()
    ang_np = (np.arange(v, dtype=np.float32) * np.float32(2 * np.pi / v)).astype(np.float32)
    ang = torch.from_numpy(ang_np).cuda()
    BP = BatchedProjector2d(P, ang, b)
    print(f"fan-arc batch 16: tile {BP.tile} variant {BP.variant}", flush=True)
    x_np = ellipses2d(n, b, seed=2024, scale=0.04)
    x = torch.from_numpy(x_np)[:, None].cuda()
    g = O.Geom2D(n, n, d, PIX, PIX, DET / SDD, SID, SDD)
    with torch.no_grad():
        s = BP.projection(x)
    pick = [0, 15]
    want = O.op2d("arc", "proj", x_np[pick], ang_np, g, "f64")
    assert report("fan-arc 512^2 batch-16 forward (slices 0, 15) vs FP64", s[pick, 0].cpu().numpy(), want) <= 1e-5
    gen = torch.Generator(device="cuda").manual_seed(5)
    y = torch.rand(b, 1, v, d, device="cuda", generator=gen)
    with torch.no_grad():
        aty = BP.projection_t(y)
    want_t = O.op2d("arc", "proj_t", y[pick, 0].cpu().numpy(), ang_np, g, "f64")
    assert report("fan-arc 512^2 batch-16 adjoint (slices 0, 15) vs FP64", aty[pick, 0].cpu().numpy(),
                  want_t) <= 1e-5
    xr = torch.rand(b, 1, n, n, device="cuda", generator=gen)
    with torch.no_grad():
        rel, lhs, rhs = dot_rel(BP.projection(xr), y, xr, aty)
    print(f"fan-arc 512^2 dot-product <Ax,y>={lhs:.10e} <x,A^T y>={rhs:.10e} rel={rel:.2e}", flush=True)
    assert rel <= 1e-6


# ------------------------------------------- 64 uneven fan-arc instances
def _c2_instances():
    """The 64-instance set of the benchmark (same seeds)."""
    from pridediff_amd.deeprecon.torch.projector import projector2d
    rng = np.random.default_rng(2025)
    projs, angs_np, imgs = [], [], []
    for i in range(64):
        n = int(rng.choice([256, 384, 512]))
        v = int(rng.integers(96, 721))
        d = int(math.ceil(1.44 * n))
        projs.append(projector2d(n, d, PIX, DET, SID, SDD, scan_type="arc").cuda())
        angs_np.append((np.arange(v, dtype=np.float32) * np.float32(2 * np.pi / v)).astype(np.float32))
        imgs.append(ellipses2d(n, 1, seed=3000 + i, scale=0.04)[0])
    return projs, angs_np, imgs


@pytest.fixture(scope="module")
def c2():
    from pridediff_amd.batched import UnevenBatch2d
    projs, angs_np, imgs = _c2_instances()
    angs = [torch.from_numpy(a).cuda() for a in angs_np]
    ub = UnevenBatch2d(projs, angs)
    return ub, projs, angs, angs_np, imgs


def _c2_sel(v):
    return np.unique(np.array([0, v // 3, (2 * v) // 3 + 1, v - 1]))


def test_uneven_64_forward_every_instance_vs_fp64(c2):
    ub, projs, angs, angs_np, imgs = c2
    x = ub.pack_images([torch.from_numpy(im).cuda() for im in imgs])
    with torch.no_grad():
        sinos = ub.unpack_sinos(ub.projection(x))
    worst = 0.0
    for i, ((h, w, v, d), a, im, s) in enumerate(zip(ub.shapes, angs_np, imgs, sinos)):
        sel = _c2_sel(v)
        g = O.Geom2D(w, h, d, PIX, PIX, DET / SDD, SID, SDD)
        want = O.op2d("arc", "proj", im[None], a[sel], g, "f64")
        r = O.rel_l2(s[sel].cpu().numpy()[None], want)
        worst = max(worst, r)
        assert r <= 1e-5, (i, (h, w, v, d), r)
    print(f"uneven batch forward, 64 instances x 4 views vs FP64: worst rel_l2 {worst:.3e}", flush=True)


def test_uneven_64_adjoint_every_instance_vs_fp64(c2):
    ub, projs, angs, angs_np, imgs = c2
    gen = torch.Generator(device="cuda").manual_seed(9)
    ys = []
    for (h, w, v, d) in ub.shapes:
        y = torch.zeros(v, d, device="cuda")
        sel = torch.from_numpy(_c2_sel(v)).cuda()
        y[sel] = torch.rand(len(sel), d, device="cuda", generator=gen)
        ys.append(y)
    with torch.no_grad():
        backs = ub.unpack_images(ub.projection_t(ub.pack_sinos(ys)))
    worst = 0.0
    for i, ((h, w, v, d), a, y, b) in enumerate(zip(ub.shapes, angs_np, ys, backs)):
        sel = _c2_sel(v)
        g = O.Geom2D(w, h, d, PIX, PIX, DET / SDD, SID, SDD)
        want = O.op2d("arc", "proj_t", y[sel].cpu().numpy()[None], a[sel], g, "f64")
        r = O.rel_l2(b.cpu().numpy()[None], want)
        worst = max(worst, r)
        assert r <= 1e-5, (i, (h, w, v, d), r)
    print(f"uneven batch adjoint, 64 instances x 4 views vs FP64: worst rel_l2 {worst:.3e}", flush=True)


def test_uneven_64_dot_product_per_instance(c2):
    ub = c2[0]
    gen = torch.Generator(device="cuda").manual_seed(10)
    n_img = sum(h * w for h, w, _, _ in ub.shapes)
    n_sino = sum(v * d for _, _, v, d in ub.shapes)
    x = torch.rand(n_img, device="cuda", generator=gen)
    y = torch.rand(n_sino, device="cuda", generator=gen)
    with torch.no_grad():
        ax = ub.unpack_sinos(ub.projection(x))
        aty = ub.unpack_images(ub.projection_t(y))
    worst = 0.0
    for i, (a, yy, xx, b) in enumerate(zip(ax, ub.unpack_sinos(y), ub.unpack_images(x), aty)):
        rel, lhs, rhs = dot_rel([a], [yy], [xx], [b])
        worst = max(worst, rel)
        assert rel <= 1e-6, (i, ub.shapes[i], lhs, rhs)
    print(f"uneven batch dot-product, 64 instances: worst rel {worst:.2e}", flush=True)


def test_uneven_64_running_sums_vs_fp64_and_dot(c2):
    """The running-sum gathers (DD_FWD_PREFIX, DD_ADJ_BLOCK_PREFIX) on all 64
    instances: forward on 4 views each and adjoint of 4 sparse views each vs
    the FP64 oracle <= 1e-5 (worst printed), and the dot-product test of the
    pair <= 1e-6 per instance."""
    from pridediff_amd import load_radontorch
    rt = load_radontorch()
    ub, projs, angs, angs_np, imgs = c2
    x = ub.pack_images([torch.from_numpy(im).cuda() for im in imgs])
    gen = torch.Generator(device="cuda").manual_seed(13)
    y = torch.rand(sum(v * d for _, _, v, d in ub.shapes), device="cuda", generator=gen)
    ys = []
    for (h, w, v, d) in ub.shapes:
        yy = torch.zeros(v, d, device="cuda")
        sel = torch.from_numpy(_c2_sel(v)).cuda()
        yy[sel] = torch.rand(len(sel), d, device="cuda", generator=gen)
        ys.append(yy)
    try:
        ub.plan.set_variant(rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX)
        with torch.no_grad():
            sinos = ub.unpack_sinos(ub.projection(x))
            aty = ub.unpack_images(ub.projection_t(y))
            backs = ub.unpack_images(ub.projection_t(ub.pack_sinos(ys)))
    finally:
        ub.plan.set_variant(rt.ADJ_QUAD)
    worst_f = worst_a = 0.0
    for i, ((h, w, v, d), a, im, s, yy, b) in enumerate(zip(ub.shapes, angs_np, imgs, sinos,
                                                          ys, backs)):
        sel = _c2_sel(v)
        g = O.Geom2D(w, h, d, PIX, PIX, DET / SDD, SID, SDD)
        want = O.op2d("arc", "proj", im[None], a[sel], g, "f64")
        r = O.rel_l2(s[sel].cpu().numpy()[None], want)
        worst_f = max(worst_f, r)
        assert r <= 1e-5, ("forward", i, (h, w, v, d), r)
        want = O.op2d("arc", "proj_t", yy[sel].cpu().numpy()[None], a[sel], g, "f64")
        r = O.rel_l2(b.cpu().numpy()[None], want)
        worst_a = max(worst_a, r)
        assert r <= 1e-5, ("adjoint", i, (h, w, v, d), r)
    worst_dot = 0.0
    for i, (a, yy, xx, b) in enumerate(zip(sinos, ub.unpack_sinos(y), ub.unpack_images(x), aty)):
        rel, lhs, rhs = dot_rel([a], [yy], [xx], [b])
        worst_dot = max(worst_dot, rel)
        assert rel <= 1e-6, (i, ub.shapes[i], lhs, rhs)
    print(f"uneven batch running sums, 64 instances x 4 views vs FP64: forward worst rel_l2 {worst_f:.3e}, "
          f"adjoint {worst_a:.3e}; dot-product worst rel {worst_dot:.2e}", flush=True)


def test_uneven_64_small_grid_strides_items(c2):
    """A grid far smaller than the item count makes every workgroup stride
    over many items (and reuse its LDS across them): results equal the
    default one-workgroup-per-item launch bit for bit, for every op and
    adjoint variant."""
    from pridediff_amd import load_radontorch
    rt = load_radontorch()
    ub, projs, angs, angs_np, imgs = c2
    x = ub.pack_images([torch.from_numpy(im).cuda() for im in imgs])
    gen = torch.Generator(device="cuda").manual_seed(12)
    y = torch.rand(sum(v * d for _, _, v, d in ub.shapes), device="cuda", generator=gen)
    ops = ("projection", "backprojection_t")
    adj = ("projection_t", "backprojection", "backprojection_w")
    with torch.no_grad():
        ref = {o: getattr(ub, o)(x).clone() for o in ops}
        variants = (rt.ADJ_PER_PIXEL, rt.ADJ_TILE16, rt.ADJ_QUAD, rt.ADJ_BLOCK_WINDOW)
        ref_adj = {}
        for var in variants:
            ub.plan.set_variant(var)
            ref_adj[var] = {o: getattr(ub, o)(y).clone() for o in adj}
        ub.plan.set_grid(3)
        try:
            for o in ops:
                assert torch.equal(getattr(ub, o)(x), ref[o]), o
            for var in variants:
                ub.plan.set_variant(var)
                for o in adj:
                    assert torch.equal(getattr(ub, o)(y), ref_adj[var][o]), (var, o)
        finally:
            ub.plan.set_grid(0)
            ub.plan.set_variant(rt.ADJ_QUAD)


# ----------------------------------------------- cone-flat 256^3, batch 2
@pytest.fixture(scope="module")
def c3():
    from pridediff_amd.deeprecon.torch.projector import projector3d
    W = H = Dz = 256
    C, R, V, B = 512, 384, 360, 2
    P = projector3d((W, H, Dz), (C, R), 1.0, 1.0, 1000.0, 1500.0, scan_type="flat").cuda()
    ang_np = (np.arange(V) * (2 * np.pi / V)).astype(np.float32)
    g = O.Geom3D(W, H, Dz, C, R, 1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 1500.0)
    x_np = ellipsoids3d(W, H, Dz, B, count=20, seed=77) * np.float32(0.02)
    return P, ang_np, g, x_np


C3_SEL = np.array([0, 45, 118, 225, 301])   # x- and y-axis views and both 45-degree diagonals


def test_cone_flat_full_size_forward_vs_fp64(c3):
    P, ang_np, g, x_np = c3
    ang = torch.from_numpy(ang_np).cuda()
    x = torch.from_numpy(x_np)[:, None].cuda()
    with torch.no_grad():
        s = P.projection(x, ang)
    got = s[:, 0, C3_SEL].cpu().numpy()
    want = O.op3d("cone_flat", "proj", x_np, ang_np[C3_SEL], g)
    assert report("cone-flat 256^3 forward, 5 views x 2 volumes vs FP64", got, want) <= 1e-5


def test_cone_flat_full_size_adjoint_vs_fp64(c3):
    P, ang_np, g, _ = c3
    ang = torch.from_numpy(ang_np).cuda()
    gen = torch.Generator(device="cuda").manual_seed(13)
    y = sparse_views(torch.rand(2, 1, len(ang_np), g.R, g.C, device="cuda", generator=gen),
                     torch.from_numpy(C3_SEL).cuda())
    x = torch.zeros(2, 1, g.Dz, g.H, g.W, device="cuda", requires_grad=True)
    P.projection(x, ang).backward(y)          # the autograd backward is ProjTrans
    want = O.op3d("cone_flat", "proj_t", y[:, 0, C3_SEL].cpu().numpy(), ang_np[C3_SEL], g)
    assert report("cone-flat 256^3 adjoint (autograd), 5 views vs FP64",
                  x.grad[:, 0].cpu().numpy(), want) <= 1e-5


def test_cone_flat_full_size_dot_product(c3):
    P, ang_np, g, _ = c3
    ang = torch.from_numpy(ang_np).cuda()
    gen = torch.Generator(device="cuda").manual_seed(14)
    x = torch.rand(2, 1, g.Dz, g.H, g.W, device="cuda", generator=gen)
    y = torch.rand(2, 1, len(ang_np), g.R, g.C, device="cuda", generator=gen)
    with torch.no_grad():
        rel, lhs, rhs = dot_rel(P.projection(x, ang), y, x, P.projection_t(y, ang))
    print(f"cone-flat 256^3 dot-product <Ax,y>={lhs:.10e} <x,A^T y>={rhs:.10e} rel={rel:.2e}", flush=True)
    assert rel <= 1e-6


# -------------------------------------------- helical 512x512x64, batch 4
@pytest.fixture(scope="module")
def c4():
    from pridediff_amd.deeprecon.torch.projector import projectorsp
    W, H, Dz = 512, 512, 64
    C, R, B = 736, 64, 4
    V = 3 * 1152
    P = projectorsp((W, H, Dz), (C, R), (0.7, 0.7, 1.0), (1.2858, 1.82454), SID, SDD,
                    scan_type="arc").cuda()
    ang_np = (np.arange(V) * (2 * np.pi / 1152)).astype(np.float32)
    spos = np.zeros((4, V), dtype=np.float32)
    spos[0] = np.linspace(-96.0, 96.0, V, dtype=np.float32)
    g = O.Geom3D(W, H, Dz, C, R, 0.7, 0.7, 1.0, 1.2858 / SDD, 1.82454, SID, SDD)
    x_np = ellipsoids3d(W, H, Dz, B, count=20, seed=77) * np.float32(0.02)
    return P, ang_np, spos, g, x_np


C4_SEL = np.array([0, 700, 1728, 2301, 3455])   # both ends of the helix, the middle, mixed axes


def test_helical_full_size_forward_vs_fp64(c4):
    P, ang_np, spos, g, x_np = c4
    ang, sp = torch.from_numpy(ang_np).cuda(), torch.from_numpy(spos).cuda()
    x = torch.from_numpy(x_np)[:, None].cuda()
    with torch.no_grad():
        s = P.projection(x, ang, sp)
    pick = [0, 3]
    got = s[pick, 0][:, C4_SEL].cpu().numpy()
    del s
    want = O.op3d("spiral", "proj", x_np[pick], ang_np[C4_SEL], g,
                  np.ascontiguousarray(spos[:, C4_SEL]))
    assert report("helical 512x512x64 forward, 5 views x 2 volumes vs FP64", got, want) <= 1e-5


def test_helical_full_size_adjoint_vs_fp64(c4):
    P, ang_np, spos, g, _ = c4
    ang, sp = torch.from_numpy(ang_np).cuda(), torch.from_numpy(spos).cuda()
    gen = torch.Generator(device="cuda").manual_seed(15)
    y = torch.zeros(4, 1, len(ang_np), g.R, g.C, device="cuda")
    sel = torch.from_numpy(C4_SEL).cuda()
    y[:, :, sel] = torch.rand(4, 1, len(C4_SEL), g.R, g.C, device="cuda", generator=gen)
    x = torch.zeros(4, 1, g.Dz, g.H, g.W, device="cuda", requires_grad=True)
    P.projection(x, ang, sp).backward(y)      # ProjTransSpiralDis3d with source_pos
    pick = [1, 2]
    want = O.op3d("spiral", "proj_t", y[pick, 0][:, sel].cpu().numpy(), ang_np[C4_SEL], g,
                  np.ascontiguousarray(spos[:, C4_SEL]))
    assert report("helical 512x512x64 adjoint (autograd), 5 views vs FP64",
                  x.grad[pick, 0].cpu().numpy(), want) <= 1e-5


def test_helical_full_size_dot_product(c4):
    P, ang_np, spos, g, _ = c4
    ang, sp = torch.from_numpy(ang_np).cuda(), torch.from_numpy(spos).cuda()
    gen = torch.Generator(device="cuda").manual_seed(16)
    x = torch.rand(4, 1, g.Dz, g.H, g.W, device="cuda", generator=gen)
    y = torch.rand(4, 1, len(ang_np), g.R, g.C, device="cuda", generator=gen)
    with torch.no_grad():
        ax = P.projection(x, ang, sp)
        aty = P.projection_t(y, ang, sp)
        rel, lhs, rhs = dot_rel(ax, y, x, aty)
    print(f"helical 512x512x64 dot-product <Ax,y>={lhs:.10e} <x,A^T y>={rhs:.10e} rel={rel:.2e}", flush=True)
    assert rel <= 1e-6
