"""Running-sum adjoint of single calls (needs an MI355X).

At instance tile 1, DD_ADJ_BLOCK_PREFIX runs ProjTrans and BackProj of
dd2d_run / planned runs through per-view running integrals of the sinogram
(FP64 knots), with each 64 x 64 pixel tile's views split over workgroups and
the partial images added in order; it is chosen explicitly (or by the
autotuner), the automatic choice keeping the walks.  Same operator as the
walks, different rounding, so it is held to the FP64 oracle and to the walks
by tolerance:

* ProjTrans and BackProj vs the FP64 oracle <= 1e-6 relative L2 on para /
  fan-flat / fan-arc with shifted and rectangular images, bins narrower and
  wider than a pixel, and inputs on a 1e3 offset (running sums ~1e3 times
  the values they are differences of); single calls and a batch of 3 at a
  forced tile of 1;
* with DD_FWD_PREFIX OR-ed in, Proj and BackProjTrans likewise through the
  lines' running pixel sums (FP32 hi/lo pairs of an FP64 scan), <= 1e-6 of
  FP64 on the same cases;
* dot products <= 1e-6 for every pairing of walk and running-sum forward
  and adjoint;
* the weighted ops keep their walks: equal bit for bit under the flags;
* a 512^2 x 720-view slice within 1e-6 of FP64 and of the walk, and the
  default choice there equal bit for bit to the walk;
* planned runs (both directions) equal the single calls bit for bit, their
  epilogues (A x - y in the forward's reduction, x - A^T y in the
  adjoint's write) equal the compositions bit for bit, and a captured graph
  of x - A^T (A x - y) replays the eager result bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pridediff_amd.phantom import ellipses2d

pytestmark = pytest.mark.gpu

OPNAMES = {"proj": "Proj", "proj_t": "ProjTrans", "bproj": "BackProj",
           "bproj_t": "BackProjTrans", "bproj_w": "BackProjWeighted",
           "bproj_tw": "BackProjTransWeighted"}
GEOMNAMES = {"para": "ParaDis2d", "flat": "FanFlatDis2d", "arc": "FanArcDis2d"}

# (W, H, D, pixel, det scale, pixel shift, bin shift, views)
CASES = ((96, 96, 141, 1.1, 1.0, (0.0, 0.0), 0.0, 120),
         (120, 87, 300, 1.4, 0.25, (3.1, -2.2), 0.7, 97),
         (70, 70, 25, 1.4, 6.0, (0.0, 1.5), -2.0, 64),
         (130, 140, 185, 0.7, 1.2858 / 0.7 * 0.7, (0.0, 0.0), 0.0, 150))


@pytest.fixture(scope="module")
def rt():
    from pridediff_amd import load_radontorch
    return load_radontorch()


def _geom(scan, case):
    w, h, d, px, ds, sh, bs, v = case
    if scan == "para":
        g = O.Geom2D(w, h, d, px, px, 1.2858 * ds, 1.0, 1.0, sh[0], sh[1], bs)
    elif scan == "flat":
        g = O.Geom2D(w, h, d, px, px, 1.2858 * ds, 595.0, 1085.6, sh[0], sh[1], bs)
    else:
        g = O.Geom2D(w, h, d, px, px, 1.2858 * ds / 1085.6, 595.0, 1085.6, sh[0], sh[1],
                     bs / 1085.6)
    ang = (np.arange(v) * (2 * np.pi / v) + 0.05).astype(np.float32)
    return g, ang


def _call(rt, scan, op, x, ang, g):
    fn = getattr(rt, OPNAMES[op] + GEOMNAMES[scan])
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))[:, None].cuda()
    out = fn(xt, torch.tensor(g.options(), dtype=torch.float32),
             torch.from_numpy(ang).cuda())
    torch.cuda.synchronize()
    return out[:, 0].cpu().numpy()


def _with_variant(rt, var, fn):
    rt.set_adjoint_variant(var)
    try:
        return fn()
    finally:
        rt.set_adjoint_variant(-1)


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
@pytest.mark.parametrize("offset", [0.0, 1000.0])
def test_single_call_running_sums_match_fp64(rt, scan, offset):
    rng = np.random.default_rng(21)
    try:
        for batch, tile in ((1, 0), (3, 1)):
            rt.set_instance_tile(tile)
            for i, case in enumerate(CASES):
                g, ang = _geom(scan, case)
                y = rng.random((batch, len(ang), g.D)).astype(np.float32) + np.float32(offset)
                for op in ("proj_t", "bproj"):
                    got = _with_variant(rt, rt.ADJ_BLOCK_PREFIX,
                                        lambda: _call(rt, scan, op, y, ang, g))
                    walk = _with_variant(rt, rt.ADJ_PER_PIXEL,
                                         lambda: _call(rt, scan, op, y, ang, g))
                    want = O.op2d(scan, op, y.astype(np.float64), ang, g, "f64")
                    r, rw = O.rel_l2(got, want), O.rel_l2(walk, want)
                    print(f"{scan} offset {offset} batch {batch} case {i} {op}: running sums "
                          f"{r:.2e}, walk {rw:.2e}", flush=True)
                    assert r <= 1e-6, (scan, offset, batch, i, op, r, rw)
    finally:
        rt.set_instance_tile(0)


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
@pytest.mark.parametrize("offset", [0.0, 1000.0])
def test_single_call_running_sum_forward_matches_fp64(rt, scan, offset):
    """DD_FWD_PREFIX: Proj and BackProjTrans of single calls through the
    lines' running pixel sums (hi/lo pairs of an FP64 scan)."""
    rng = np.random.default_rng(25)
    try:
        for batch, tile in ((1, 0), (3, 1)):
            rt.set_instance_tile(tile)
            for i, case in enumerate(CASES):
                g, ang = _geom(scan, case)
                x = rng.random((batch, g.H, g.W)).astype(np.float32) + np.float32(offset)
                for op in ("proj", "bproj_t"):
                    got = _with_variant(rt, rt.ADJ_PER_PIXEL | rt.FWD_PREFIX,
                                        lambda: _call(rt, scan, op, x, ang, g))
                    walk = _call(rt, scan, op, x, ang, g)
                    want = O.op2d(scan, op, x.astype(np.float64), ang, g, "f64")
                    r, rw = O.rel_l2(got, want), O.rel_l2(walk, want)
                    print(f"{scan} offset {offset} batch {batch} case {i} {op}: running sums "
                          f"{r:.2e}, walk {rw:.2e}", flush=True)
                    assert r <= 1e-6, (scan, offset, batch, i, op, r, rw)
    finally:
        rt.set_instance_tile(0)


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_single_call_running_sums_dot_products(rt, scan):
    rng = np.random.default_rng(22)
    both = rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX
    for i, case in enumerate(CASES):
        g, ang = _geom(scan, case)
        x = rng.random((1, g.H, g.W)).astype(np.float32)
        y = rng.random((1, len(ang), g.D)).astype(np.float32)
        for fwd_var, adj_var in ((-1, rt.ADJ_BLOCK_PREFIX), (both, -1), (both, both)):
            ax = _with_variant(rt, fwd_var,
                               lambda: _call(rt, scan, "proj", x, ang, g)).astype(np.float64)
            aty = _with_variant(rt, adj_var,
                                lambda: _call(rt, scan, "proj_t", y, ang, g)).astype(np.float64)
            lhs, rhs = float(np.sum(ax * y)), float(np.sum(x * aty))
            assert abs(lhs - rhs) / abs(rhs) <= 1e-6, (scan, i, fwd_var, adj_var, lhs, rhs)


@pytest.mark.parametrize("scan", ["flat", "arc"])
def test_weighted_ops_keep_their_walks(rt, scan):
    rng = np.random.default_rng(23)
    g, ang = _geom(scan, CASES[1])
    y = rng.random((1, len(ang), g.D)).astype(np.float32)
    x = rng.random((1, g.H, g.W)).astype(np.float32)
    both = rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX
    for op, inp in (("bproj_w", y), ("bproj_tw", x)):
        a = _with_variant(rt, both, lambda: _call(rt, scan, op, inp, ang, g))
        b = _with_variant(rt, rt.ADJ_PER_PIXEL, lambda: _call(rt, scan, op, inp, ang, g))
        assert np.array_equal(a, b), op


def test_full_size_slice_and_default(rt):
    """A 512^2 fan-arc slice of 720 views: the running sums within 1e-6 of
    FP64 and of the walk; the default (automatic) choice keeps the walk, so
    a slice run alone gets the bits it gets in a batch."""
    rng = np.random.default_rng(24)
    g = O.Geom2D(512, 512, 736, 0.7, 0.7, 1.2858 / 1085.6, 595.0, 1085.6)
    ang = (np.arange(720) * (2 * np.pi / 720)).astype(np.float32)
    y = rng.random((1, 720, 736)).astype(np.float32)
    pref = _with_variant(rt, rt.ADJ_BLOCK_PREFIX, lambda: _call(rt, "arc", "proj_t", y, ang, g))
    walk = _with_variant(rt, rt.ADJ_PER_PIXEL, lambda: _call(rt, "arc", "proj_t", y, ang, g))
    auto = _call(rt, "arc", "proj_t", y, ang, g)
    want = O.op2d("arc", "proj_t", y.astype(np.float64), ang, g, "f64")
    r, rw = O.rel_l2(pref, want), O.rel_l2(walk, want)
    print(f"512^2 x 720 proj_t: running sums {r:.2e}, walk {rw:.2e}", flush=True)
    assert r <= 1e-6
    assert O.rel_l2(pref, walk) <= 1e-6
    assert np.array_equal(auto, walk)


def test_planned_epilogue_and_graph_replay(rt):
    from pridediff_amd.batched import GraphStep
    from pridediff_amd.deeprecon.torch.projector import projector2d
    P = projector2d((120, 87), 300, 1.4, 1.2858 * 0.25, 595.0, 1085.6, pixshift=(3.1, -2.2),
                    binshift=0.7, scan_type="arc").cuda()
    v = 97
    ang = torch.from_numpy((np.arange(v) * (2 * np.pi / v) + 0.05).astype(np.float32)).cuda()
    both = rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX
    plan = rt.Plan2d(rt.SCAN[P.scan_type], P.host_options(), ang, 1, 1, both)
    gen = torch.Generator(device="cuda").manual_seed(5)
    y = torch.rand(1, 1, v, 300, device="cuda", generator=gen)
    x = torch.from_numpy(ellipses2d(120, 1, seed=9)[:, :87, :]).cuda()[:, None].contiguous()
    with torch.no_grad():
        planned = plan.run(rt.PROJ_T, y)
        single = _with_variant(rt, both, lambda: P.projection_t(y, ang))
        assert torch.equal(planned, single)
        planned_fp = plan.run(rt.PROJ, x)
        assert torch.equal(planned_fp, _with_variant(rt, both, lambda: P.projection(x, ang)))
        fused = plan.run_ep(rt.PROJ_T, y, c=x, alpha=-1.0, beta=1.0)
        assert torch.equal(fused, x - planned)
        # the residual A x - y in the forward's chunk reduction
        resid = plan.run_ep(rt.PROJ, x, c=y, alpha=1.0, beta=-1.0)
        assert torch.equal(resid, planned_fp - y)

        def dc_step(a, b):
            r = plan.run_ep(rt.PROJ, a, c=b, alpha=1.0, beta=-1.0)
            return plan.run_ep(rt.PROJ_T, r, c=a, alpha=-1.0, beta=1.0)
        eager = dc_step(x, y)
        step = GraphStep(dc_step, x.clone(), y.clone())
        for i in range(3):
            got = step.replay(x, y).clone()
            torch.cuda.synchronize()
            _ = torch.randn(100, device="cuda").sum()
            assert torch.equal(got, eager), f"replay {i}"
