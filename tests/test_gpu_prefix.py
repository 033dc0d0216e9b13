"""Running-sum gathers of uneven batches (needs an MI355X).

DD_FWD_PREFIX: PROJ and BPROJ_T integrate each bin over each line as a
difference of the line's running pixel sums.  DD_ADJ_BLOCK_PREFIX: PROJ_T
and BPROJ take each pixel's value from a view as a difference of the view's
running integral over the bins at the pixel's two edges.  Both keep their
sums as FP32 (hi, lo) pairs of an FP64 accumulation.  Same operators as the
walks with different rounding, so they are held to the FP64 oracle and to
the walks by tolerance:

* every op vs the FP64 oracle <= 1e-6 relative L2 (the walk's distance is
  printed beside it) on para / fan-flat / fan-arc instances with shifted
  images and detectors, rectangular images, bins much narrower and much
  wider than a pixel (a bin edge far outside the line), and detectors
  reaching past the image;
* inputs on a large constant offset (running sums ~1e3 times the values
  they are differences of): an uncompensated FP32 running sum would be
  ~1e-4 off there, the hi/lo pairs keep the error at the walk's level;
* dot-product tests of each running-sum op against the other direction
  <= 1e-6;
* the weighted ops keep the walks: equal bit for bit;
* a graph-captured replay equals the eager step bit for bit, and a striding
  grid equals one workgroup per item bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pridediff_amd.phantom import ellipses2d

pytestmark = pytest.mark.gpu

# (W=H, D, det scale, pixel size, pixel shift, bin shift)
CASES = ((64, 93, 2.0, 2.8, (0.0, 0.0), 0.0),
         (80, 300, 0.25, 1.4, (3.1, -2.2), 0.7),
         (52, 25, 6.0, 1.4, (0.0, 1.5), -2.0),
         (96, 141, 1.0, 1.1, (-4.0, 2.5), 0.3),
         (40, 400, 0.5, 3.0, (0.0, 0.0), 0.0))


def _instances(scan, seed=5):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    rng = np.random.default_rng(seed)
    projs, angs, geoms, imgs = [], [], [], []
    for i, (n, d, ds, px, sh, bs) in enumerate(CASES):
        h = n if i % 2 == 0 else n + 9  # some rectangular images
        if scan == "para":
            P = projector2d((n, h), d, px, 1.2858 * ds, 1.0, 1.0, pixshift=sh, binshift=bs,
                            scan_type="para")
        else:
            P = projector2d((n, h), d, px, 1.2858 * ds, 595.0, 1085.6, pixshift=sh,
                            binshift=bs, scan_type=scan)
        projs.append(P.cuda())
        v = int(rng.integers(17, 60))
        a = (np.arange(v) * (2 * np.pi / v) + 0.05 * i).astype(np.float32)
        angs.append(torch.from_numpy(a).cuda())
        opt = P.host_options()
        geoms.append(O.Geom2D(int(opt[0]), int(opt[1]), int(opt[2]), *[float(t) for t in opt[3:]]))
        imgs.append(torch.from_numpy(ellipses2d(max(n, h), 1, seed=400 + i, scale=0.04)[0, :h, :n]
                                     .copy()).cuda())
    return projs, angs, geoms, imgs


def _batch(scan, **kw):
    from pridediff_amd.batched import UnevenBatch2d
    projs, angs, geoms, imgs = _instances(scan, **kw)
    return UnevenBatch2d(projs, angs), projs, angs, geoms, imgs


def _rt():
    from pridediff_amd import load_radontorch
    return load_radontorch()


OPS = (("projection", "proj", "image"), ("backprojection_t", "bproj_t", "image"),
       ("projection_t", "proj_t", "sino"), ("backprojection", "bproj", "sino"))


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
@pytest.mark.parametrize("offset", [0.0, 1000.0])
def test_running_sum_ops_match_fp64(scan, offset):
    rt = _rt()
    ub, projs, angs, geoms, imgs = _batch(scan)
    imgs = [im + offset for im in imgs]
    gen = torch.Generator(device="cuda").manual_seed(2)
    sinos = [torch.rand(v, d, device="cuda", generator=gen) + offset
             for (_, _, v, d) in ub.shapes]
    walk_var, pref_var = rt.ADJ_BLOCK_WINDOW, rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX
    worst = {}
    with torch.no_grad():
        for op, name, kind in OPS:
            inp = ub.pack_images(imgs) if kind == "image" else ub.pack_sinos(sinos)
            unpack = ub.unpack_sinos if kind == "image" else ub.unpack_images
            ub.plan.set_variant(walk_var)
            walk = unpack(getattr(ub, op)(inp))
            ub.plan.set_variant(pref_var)
            pref = unpack(getattr(ub, op)(inp))
            srcs = imgs if kind == "image" else sinos
            for i, (g, a, src, got, w) in enumerate(zip(geoms, angs, srcs, pref, walk)):
                want = O.op2d(scan, name, src.cpu().numpy()[None].astype(np.float64),
                              a.cpu().numpy(), g, "f64")[0]
                r, rw = O.rel_l2(got.cpu().numpy(), want), O.rel_l2(w.cpu().numpy(), want)
                worst[(op, i)] = (r, rw)
                assert r <= 1e-6, (scan, offset, op, i, r, rw)
    ub.plan.set_variant(rt.ADJ_QUAD)
    for op, _, _ in OPS:
        top = max(((k, v) for k, v in worst.items() if k[0] == op), key=lambda kv: kv[1][0])
        print(f"{scan} offset {offset} {op}: worst running-sum rel_l2 {top[1][0]:.2e} "
              f"(walk {top[1][1]:.2e}, instance {top[0][1]})", flush=True)


@pytest.mark.parametrize("scan", ["flat", "arc"])
def test_running_sums_on_signed_inputs(scan):
    """Residual-like inputs (both signs, zero mean): still <= 1e-6 of FP64."""
    rt = _rt()
    ub, projs, angs, geoms, imgs = _batch(scan, seed=6)
    gen = torch.Generator(device="cuda").manual_seed(3)
    imgs = [torch.randn(im.shape, device="cuda", generator=gen) for im in imgs]
    sinos = [torch.randn(v, d, device="cuda", generator=gen) for (_, _, v, d) in ub.shapes]
    with torch.no_grad():
        ub.plan.set_variant(rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX)
        fwd = ub.unpack_sinos(ub.projection(ub.pack_images(imgs)))
        adj = ub.unpack_images(ub.projection_t(ub.pack_sinos(sinos)))
    ub.plan.set_variant(rt.ADJ_QUAD)
    for i, (g, a, im, sn, f, b) in enumerate(zip(geoms, angs, imgs, sinos, fwd, adj)):
        for name, src, got in (("proj", im, f), ("proj_t", sn, b)):
            want = O.op2d(scan, name, src.cpu().numpy()[None].astype(np.float64),
                          a.cpu().numpy(), g, "f64")[0]
            r = O.rel_l2(got.cpu().numpy(), want)
            assert r <= 1e-6, (scan, name, i, r)


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_running_sum_dot_products_and_weighted_walks(scan):
    rt = _rt()
    ub, projs, angs, geoms, imgs = _batch(scan, seed=7)
    gen = torch.Generator(device="cuda").manual_seed(4)
    x = torch.rand(ub.plan.image_size, device="cuda", generator=gen)
    y = torch.rand(ub.plan.sino_size, device="cuda", generator=gen)

    def check(fwd_var, adj_var, tag):
        ub.plan.set_variant(fwd_var)
        ax = ub.unpack_sinos(ub.projection(x))
        ub.plan.set_variant(adj_var)
        aty = ub.unpack_images(ub.projection_t(y))
        for i, (a, yy, xx, b) in enumerate(zip(ax, ub.unpack_sinos(y), ub.unpack_images(x), aty)):
            lhs = float((a.double() * yy.double()).sum())
            rhs = float((xx.double() * b.double()).sum())
            assert abs(lhs - rhs) / abs(rhs) <= 1e-6, (scan, tag, i, lhs, rhs)

    with torch.no_grad():
        check(rt.ADJ_QUAD | rt.FWD_PREFIX, rt.ADJ_QUAD, "prefix forward / walk adjoint")
        check(rt.ADJ_QUAD, rt.ADJ_BLOCK_PREFIX, "walk forward / prefix adjoint")
        check(rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX, rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX, "both")
        if scan != "para":
            # the weighted ops keep the walks under both flags
            ub.plan.set_variant(rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX)
            tw, w = ub.plan.run(rt.BPROJ_TW, x).clone(), ub.plan.run(rt.BPROJ_W, y).clone()
            ub.plan.set_variant(rt.ADJ_BLOCK_WINDOW)
            assert torch.equal(tw, ub.plan.run(rt.BPROJ_TW, x))
            assert torch.equal(w, ub.plan.run(rt.BPROJ_W, y))
    ub.plan.set_variant(rt.ADJ_QUAD)


def test_running_sum_graph_replay_and_small_grid():
    """A captured step with both running-sum gathers replays bit for bit
    (eager work between replays), and a grid that strides over the items
    gives the one-workgroup-per-item results bit for bit."""
    from pridediff_amd.batched import GraphStep, data_consistency_grad
    rt = _rt()
    ub, projs, angs, geoms, imgs = _batch("arc", seed=8)
    ub.plan.set_variant(rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX)
    x = ub.pack_images(imgs)
    with torch.no_grad():
        y = ub.projection(x) * 1.01
        ref, ref_t = ub.projection(x).clone(), ub.projection_t(y).clone()
        ub.plan.set_grid(3)
        strided, strided_t = ub.projection(x).clone(), ub.projection_t(y).clone()
        ub.plan.set_grid(0)
    assert torch.equal(strided, ref)
    assert torch.equal(strided_t, ref_t)
    xe = x.clone().requires_grad_(True)
    eager = data_consistency_grad(ub.projection, xe, y).clone()
    step = GraphStep(lambda a, b: data_consistency_grad(ub.projection, a, b),
                     x.clone().requires_grad_(True), y)
    for i in range(3):
        got = step.replay(x, y).clone()
        torch.cuda.synchronize()
        _ = torch.randn(100, device="cuda").sum()
        assert torch.equal(got, eager), f"replay {i}"
    ub.plan.set_variant(rt.ADJ_QUAD)
