"""Edge cases of the HIP path (needs an MI355X): empty and ragged inputs,
degenerate sizes, axis-switch angles, partial instance tiles, odd 3D sizes.
Every result is compared with the FP64 oracle at rel-L2 <= 1e-5 (max-abs
printed), or checked for the exact shape / error the case calls for."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPNAMES = {"proj": "Proj", "proj_t": "ProjTrans", "bproj": "BackProj", "bproj_t": "BackProjTrans",
           "bproj_w": "BackProjWeighted", "bproj_tw": "BackProjTransWeighted"}
GEOMNAMES = {"para": "ParaDis2d", "flat": "FanFlatDis2d", "arc": "FanArcDis2d"}


@pytest.fixture(scope="module")
def rt():
    from pridediff_amd import load_radontorch
    return load_radontorch()


def _run(rt, geom, op, x, ang, g):
    fn = getattr(rt, OPNAMES[op] + GEOMNAMES[geom])
    out = fn(torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))[:, None].cuda(),
             torch.tensor(g.options(), dtype=torch.float32),
             torch.from_numpy(np.ascontiguousarray(ang, dtype=np.float32)).cuda())
    torch.cuda.synchronize()
    return out[:, 0].cpu().numpy()


def _check(rt, geom, g, ang, b=2, seed=0, ops=None):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (b, g.H, g.W)).astype(np.float32)
    y = rng.uniform(0, 1, (b, len(ang), g.D)).astype(np.float32)
    ops = ops or (["proj", "proj_t", "bproj", "bproj_t"] +
                  ([] if geom == "para" else ["bproj_w", "bproj_tw"]))
    for op in ops:
        inp = x if op in O.TO_SINO else y
        got = _run(rt, geom, op, inp, ang, g)
        want = O.op2d(geom, op, inp, ang, g, "f64")
        r, m = O.rel_l2(got, want), O.max_abs(got, want)
        print(f"{geom}/{op} W={g.W} H={g.H} D={g.D} V={len(ang)}: rel_l2={r:.3e} max_abs={m:.3e}",
              flush=True)
        assert r <= 1e-5, (geom, op, r)


def _geom(geom, w, h, d, pix=1.0, det=1.3, shx=0.0, shy=0.0, bin_=0.0):
    sdd = 260.0
    return O.Geom2D(w, h, d, pix, pix * 1.1, det / sdd if geom == "arc" else det, 150.0, sdd,
                    shx, shy, bin_ / sdd if geom == "arc" else bin_)


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
def test_empty_batch_and_no_views(rt, geom):
    g = _geom(geom, 16, 12, 21)
    opt = torch.tensor(g.options(), dtype=torch.float32)
    fn = getattr(rt, "Proj" + GEOMNAMES[geom])
    out = fn(torch.zeros(0, 1, 12, 16).cuda(), opt, torch.zeros(5).cuda())
    assert tuple(out.shape) == (0, 1, 5, 21)
    fnt = getattr(rt, "ProjTrans" + GEOMNAMES[geom])
    out = fnt(torch.zeros(3, 1, 0, 21).cuda(), opt, torch.zeros(0).cuda())
    assert tuple(out.shape) == (3, 1, 12, 16) and float(out.abs().sum()) == 0.0


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
@pytest.mark.parametrize("w,h,d,v", [(37, 29, 53, 7), (1, 1, 1, 1), (5, 64, 3, 4), (64, 5, 9, 13),
                                     (48, 48, 11, 6), (44, 40, 63, 5), (44, 40, 64, 5),
                                     (30, 34, 126, 3), (30, 34, 127, 3)])
def test_ragged_and_degenerate_sizes(rt, geom, w, h, d, v):
    """Odd and non-square images, one pixel, one bin, tall and wide images,
    a detector far narrower than the image, and bin counts either side of
    the forward's 63-bin wavefronts."""
    g = _geom(geom, w, h, d, pix=2.0 if max(w, h) <= 5 else 1.0)
    ang = (np.arange(v) * (2 * np.pi / v) + 0.3).astype(np.float32)
    _check(rt, geom, g, ang, seed=w * 100 + h)


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
def test_axis_switch_and_out_of_range_angles(rt, geom):
    """Views exactly at multiples of 45 degrees (where the common axis
    switches), negative angles and angles beyond 2 pi; shifted image and
    detector."""
    g = _geom(geom, 40, 36, 61, shx=1.5, shy=-2.25, bin_=0.7)
    ang = np.array([k * np.pi / 4 for k in range(8)] + [-0.3, -np.pi / 4, 2 * np.pi + 0.2,
                                                        7.5 * np.pi / 4, 13.0],
                   dtype=np.float32)
    _check(rt, geom, g, ang, seed=3)


def test_partial_instance_tiles(rt):
    """19 instances (not a multiple of the 16-instance tile) in one call
    equal 19 single calls bit for bit."""
    g = _geom("arc", 48, 40, 67)
    ang = (np.arange(30) * (2 * np.pi / 30)).astype(np.float32)
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (19, 40, 48)).astype(np.float32)
    y = rng.uniform(0, 1, (19, 30, 67)).astype(np.float32)
    for op, inp in (("proj", x), ("proj_t", y), ("bproj_w", y)):
        whole = _run(rt, "arc", op, inp, ang, g)
        single = np.concatenate([_run(rt, "arc", op, inp[i:i + 1], ang, g) for i in range(19)])
        assert np.array_equal(whole, single), op


def test_channels_fold_into_the_batch(rt):
    """[B, C, ...] inputs with C > 1 are B*C instances."""
    g = _geom("flat", 24, 24, 35)
    ang = (np.arange(12) * (2 * np.pi / 12)).astype(np.float32)
    x = torch.rand(2, 3, 24, 24).cuda()
    opt = torch.tensor(g.options(), dtype=torch.float32)
    a = torch.from_numpy(ang).cuda()
    out = rt.ProjFanFlatDis2d(x, opt, a)
    assert tuple(out.shape) == (2, 3, 12, 35)
    flat = rt.ProjFanFlatDis2d(x.reshape(6, 1, 24, 24).contiguous(), opt, a)
    assert torch.equal(out.reshape(6, 1, 12, 35), flat)


@pytest.mark.parametrize("scan", ["cone_flat", "cone_arc", "spiral"])
def test_3d_odd_sizes(rt, scan):
    sdd = 230.0
    col = 1.2 if scan == "cone_flat" else 1.2 / sdd
    g = O.Geom3D(13, 9, 5, 17, 3, 1.1, 0.9, 1.3, col, 1.4, 130.0, sdd, 0.4, -0.2,
                 0.3 if scan != "spiral" else 0.0, 0.0, 0.1)
    ang = np.array([0.0, np.pi / 4, 1.1, 3 * np.pi / 4, 2.0, 4.0, -0.5], dtype=np.float32)
    spos = None
    if scan == "spiral":
        spos = np.stack([np.linspace(-2, 2, 7), 0.3 * np.ones(7), 0.01 * np.ones(7),
                         0.1 * np.ones(7)]).astype(np.float32)
    names = {"cone_flat": "ConeFlatDis3d", "cone_arc": "ConeArcDis3d", "spiral": "SpiralDis3d"}
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (1, g.Dz, g.H, g.W)).astype(np.float32)
    y = rng.uniform(0, 1, (1, len(ang), g.R, g.C)).astype(np.float32)
    for op in OPNAMES:
        inp = x if op in O.TO_SINO else y
        fn = getattr(rt, OPNAMES[op] + names[scan])
        args = [torch.from_numpy(inp)[:, None].cuda(),
                torch.tensor(g.options(spiral=scan == "spiral"), dtype=torch.float32),
                torch.from_numpy(ang).cuda()]
        if spos is not None:
            args.append(torch.from_numpy(spos).cuda())
        got = fn(*args)[:, 0].cpu().numpy()
        want = O.op3d(scan, op, inp, ang, g, spos)
        r = O.rel_l2(got, want)
        print(f"{scan}/{op} odd sizes: rel_l2={r:.3e} max_abs={O.max_abs(got, want):.3e}",
              flush=True)
        assert r <= 1e-5, (scan, op, r)


@pytest.mark.parametrize("scan", ["cone_flat", "cone_arc", "spiral"])
def test_3d_partial_instance_tiles(rt, scan):
    """6 instances (tile 4, padded batch 8: the instance stride exceeds the
    tile) in one call equal 6 single calls, for all six ops."""
    sdd = 230.0
    col = 1.2 if scan == "cone_flat" else 1.2 / sdd
    g = O.Geom3D(14, 12, 6, 19, 4, 1.1, 0.9, 1.3, col, 1.4, 130.0, sdd, 0.2, -0.1,
                 0.2 if scan != "spiral" else 0.0, 0.0, 0.1)
    ang = (np.arange(9) * (2 * np.pi / 9) + 0.1).astype(np.float32)
    spos = None
    if scan == "spiral":
        spos = np.stack([np.linspace(-2, 2, 9), 0.3 * np.ones(9), 0.01 * np.ones(9),
                         0.1 * np.ones(9)]).astype(np.float32)
    names = {"cone_flat": "ConeFlatDis3d", "cone_arc": "ConeArcDis3d", "spiral": "SpiralDis3d"}
    rng = np.random.default_rng(8)
    x = rng.uniform(0, 1, (6, g.Dz, g.H, g.W)).astype(np.float32)
    y = rng.uniform(0, 1, (6, len(ang), g.R, g.C)).astype(np.float32)

    def run(op, inp):
        fn = getattr(rt, OPNAMES[op] + names[scan])
        args = [torch.from_numpy(np.ascontiguousarray(inp))[:, None].cuda(),
                torch.tensor(g.options(spiral=scan == "spiral"), dtype=torch.float32),
                torch.from_numpy(ang).cuda()]
        if spos is not None:
            args.append(torch.from_numpy(spos).cuda())
        return fn(*args)[:, 0].cpu().numpy()

    for op in OPNAMES:
        inp = x if op in O.TO_SINO else y
        whole = run(op, inp)
        single = np.concatenate([run(op, inp[i:i + 1]) for i in range(6)])
        r = O.rel_l2(whole, single)
        print(f"{scan}/{op} batch 6 vs singles: rel_l2={r:.3e} "
              f"bitwise={np.array_equal(whole, single)}", flush=True)
        assert r <= 1e-6, (scan, op, r)


def test_views2d_rejects_too_many_views(rt):
    g = _geom("arc", 8, 8, 9)
    opt = torch.tensor(g.options(), dtype=torch.float32)
    with pytest.raises(RuntimeError):
        rt.views2d(rt.SCAN["arc"], rt.BPROJ, torch.zeros(1, 1, 65536, 9).cuda(), opt,
                   torch.zeros(65536).cuda())


def test_device_resident_options_are_cached_and_refreshed(rt):
    """A GPU options tensor is read once per version: reuse gives the same
    result, an in-place change is picked up, a graph capture needs no read."""
    g = _geom("arc", 32, 32, 45)
    ang = torch.from_numpy((np.arange(20) * (2 * np.pi / 20)).astype(np.float32)).cuda()
    x = torch.rand(2, 1, 32, 32).cuda()
    opt = torch.tensor(g.options(), dtype=torch.float32).cuda()
    host = torch.tensor(g.options(), dtype=torch.float32)
    a = rt.ProjFanArcDis2d(x, opt, ang)
    assert torch.equal(a, rt.ProjFanArcDis2d(x, opt, ang))
    assert torch.equal(a, rt.ProjFanArcDis2d(x, host, ang))
    opt[3] = opt[3] * 1.5  # pixel size: in-place change bumps the version
    host2 = host.clone()
    host2[3] = host2[3] * 1.5
    b = rt.ProjFanArcDis2d(x, opt, ang)
    assert torch.equal(b, rt.ProjFanArcDis2d(x, host2, ang))
    assert not torch.equal(a, b)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        rt.ProjFanArcDis2d(x, opt, ang)
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(graph):
        out = rt.ProjFanArcDis2d(x, opt, ang)
    graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(out, b)
