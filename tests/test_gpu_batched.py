"""Batched path: planned ops and graph replay (needs an MI355X).

* a Plan2d op equals the per-call radontorch op bit for bit;
* a captured data-consistency step, replayed several times with eager work
  and synchronisations between replays, equals the eager step every time.
"""
import numpy as np
import pytest
import torch

from pridediff_amd.phantom import ellipses2d

pytestmark = pytest.mark.gpu


def _setup(n=128, d=185, v=120, b=5, scan="arc"):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    P = projector2d(n, d, 0.7 * 4, 1.2858 * 4, 595.0, 1085.6, scan_type=scan).cuda()
    ang = torch.from_numpy((np.arange(v, dtype=np.float32) * np.float32(2 * np.pi / v))).cuda()
    x = torch.from_numpy(ellipses2d(n, b, seed=31, scale=0.04))[:, None].cuda()
    y = torch.rand(b, 1, v, d, device="cuda")
    return P, ang, x, y


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_plan_ops_equal_plain_ops(scan):
    from pridediff_amd.batched import BatchedProjector2d
    P, ang, x, y = _setup(scan=scan)
    BP = BatchedProjector2d(P, ang, x.shape[0])
    with torch.no_grad():
        assert torch.equal(BP.projection(x), P.projection(x, ang))
        assert torch.equal(BP.projection_t(y), P.projection_t(y, ang))
        assert torch.equal(BP.backprojection(y), P.backprojection(y, ang))
        assert torch.equal(BP.backprojection_t(x), P.backprojection_t(x, ang))
        assert torch.equal(BP.backprojection_w(y), P.backprojection_w(y, ang))


@pytest.mark.parametrize("n", [100, 99])
@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_adjoint_kernel_variants_agree_bitwise(scan, n):
    """Per-pixel, LDS-tiled (16x16, 8x8) and 2x2-quad sinogram -> image
    kernels sum the same overlaps in the same order, so they agree bit for
    bit; sizes that are not tile multiples (odd, too) exercise the partial
    tiles."""
    from pridediff_amd import load_radontorch
    from pridediff_amd.batched import BatchedProjector2d
    rt = load_radontorch()
    P, ang, x, y = _setup(n=n, d=151, v=90, b=3, scan=scan)
    outs = {}
    for var in (rt.ADJ_PER_PIXEL, rt.ADJ_TILE16, rt.ADJ_TILE8, rt.ADJ_QUAD):
        BP = BatchedProjector2d(P, ang, x.shape[0], variant=var)
        with torch.no_grad():
            outs[var] = (BP.projection_t(y), BP.backprojection(y), BP.backprojection_w(y))
    for var in (rt.ADJ_TILE16, rt.ADJ_TILE8, rt.ADJ_QUAD):
        for op, a, b in zip(("proj_t", "bproj", "bproj_w"), outs[rt.ADJ_PER_PIXEL], outs[var]):
            assert torch.equal(a, b), (scan, var, op, (a - b).abs().max().item())


def _run_angles(kind):
    if kind == "full":       # full turn from 0.3 rad
        a = 0.3 + np.arange(90) * (2 * np.pi / 90)
    elif kind == "short":    # half turn plus fan, from a negative angle
        a = -0.2 + np.arange(70) * ((np.pi + 0.6) / 70)
    elif kind == "shuffled":  # views out of order
        a = np.random.default_rng(5).permutation(np.arange(60) * (2 * np.pi / 60))
    else:                    # one view next to the 45 degree axis switch
        a = np.array([0.7854])
    return torch.from_numpy(a.astype(np.float32)).cuda()


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
@pytest.mark.parametrize("size,batch,tile,angles", [
    ((130, 70), 16, 16, "full"),    # W != H, one 16-instance tile
    ((130, 70), 16, 8, "full"),     # two instance tiles in one launch
    ((64, 127), 9, 4, "short"),     # padded tiles
    ((63, 126), 3, 0, "shuffled"),
    ((100, 100), 1, 0, "single"),
])
def test_adjoint_variants_agree_on_irregular_scans(scan, size, batch, tile, angles):
    """The per-pixel kernel (each view's search started one view ahead from a
    guess taken at the pixel centre) and the 2x2-quad and 8x8-tile kernels
    (guess from the line's edge map, found view by view) sum the same
    overlaps in the same order: equal bit for bit on rectangular images,
    instance tiles, padded batches, views out of order and a single view."""
    from pridediff_amd import load_radontorch
    from pridediff_amd.batched import BatchedProjector2d
    from pridediff_amd.deeprecon.torch.projector import projector2d
    rt = load_radontorch()
    W, H = size
    P = projector2d((W, H), 151, 0.7 * 4, 1.2858 * 4, 595.0, 1085.6, scan_type=scan).cuda()
    ang = _run_angles(angles)
    gen = torch.Generator(device="cuda").manual_seed(11)
    y = torch.rand(batch, 1, ang.numel(), 151, device="cuda", generator=gen)
    outs = {}
    for var in (rt.ADJ_PER_PIXEL, rt.ADJ_QUAD, rt.ADJ_TILE8):
        BP = BatchedProjector2d(P, ang, batch, tile=tile, variant=var)
        with torch.no_grad():
            outs[var] = (BP.projection_t(y), BP.backprojection(y), BP.backprojection_w(y))
    for var in (rt.ADJ_QUAD, rt.ADJ_TILE8):
        for op, a, b in zip(("proj_t", "bproj", "bproj_w"), outs[rt.ADJ_PER_PIXEL], outs[var]):
            assert torch.equal(a, b), (scan, size, angles, var, op, (a - b).abs().max().item())


def test_graph_replay_matches_eager():
    from pridediff_amd.batched import BatchedProjector2d, GraphStep, data_consistency_grad
    P, ang, x, y = _setup()
    BP = BatchedProjector2d(P, ang, x.shape[0])
    step = GraphStep(lambda a, b: data_consistency_grad(BP.projection, a, b),
                     x.clone().requires_grad_(True), y)
    for i in range(4):
        xi = (x + 0.01 * i).contiguous()
        yi = (y * (1.0 + 0.1 * i)).contiguous()
        xe = xi.clone().requires_grad_(True)
        eager = data_consistency_grad(BP.projection, xe, yi).clone()
        torch.cuda.synchronize()
        _ = torch.randn(1000, device="cuda").sum()  # unrelated eager work between replays
        got = step.replay(xi, yi).clone()
        torch.cuda.synchronize()
        assert torch.equal(got, eager), f"replay {i} differs"


def _uneven(n_inst=8, seed=50):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    rng = np.random.default_rng(seed)
    projs, angs, imgs = [], [], []
    for i in range(n_inst):
        n = int(rng.choice([64, 96, 128, 75]))
        v = int(rng.integers(24, 91))
        d = int(np.ceil(1.44 * n))
        projs.append(projector2d(n, d, 0.7 * 4, 1.2858 * 4, 595.0, 1085.6, scan_type="arc").cuda())
        angs.append(torch.from_numpy((np.arange(v) * (2 * np.pi / v)).astype(np.float32)).cuda())
        imgs.append(torch.from_numpy(ellipses2d(n, 1, seed=100 + i, scale=0.04))[0].cuda())
    return projs, angs, imgs


def test_uneven_batch_equals_per_instance_ops():
    from pridediff_amd.batched import UnevenBatch2d
    projs, angs, imgs = _uneven()
    ub = UnevenBatch2d(projs, angs)
    with torch.no_grad():
        sinos = ub.unpack_sinos(ub.projection(ub.pack_images(imgs)))
        for P, a, im, s in zip(projs, angs, imgs, sinos):
            assert torch.equal(s, P.projection(im[None, None], a)[0, 0])
        ys = [torch.rand_like(s) for s in sinos]
        from pridediff_amd import load_radontorch
        rt = load_radontorch()
        for var in (rt.ADJ_PER_PIXEL, rt.ADJ_TILE16, rt.ADJ_QUAD, rt.ADJ_BLOCK_WINDOW):
            ub.plan.set_variant(var)
            backs = ub.unpack_images(ub.projection_t(ub.pack_sinos(ys)))
            for P, a, y, b in zip(projs, angs, ys, backs):
                assert torch.equal(b, P.projection_t(y[None, None], a)[0, 0]), var
        bw = ub.unpack_images(ub.backprojection_w(ub.pack_sinos(ys)))
        bt = ub.unpack_sinos(ub.backprojection_t(ub.pack_images(imgs)))
        for P, a, y, im, b1, b2 in zip(projs, angs, ys, imgs, bw, bt):
            assert torch.equal(b1, P.backprojection_w(y[None, None], a)[0, 0])
            assert torch.equal(b2, P.backprojection_t(im[None, None], a)[0, 0])


def test_uneven_bin_counts_at_wave_boundaries():
    """Instances whose bin count sits either side of the forward's 63-bin
    wavefronts (62, 63, 64, 126, 127) equal the per-instance ops bit for bit."""
    from pridediff_amd.batched import UnevenBatch2d
    from pridediff_amd.deeprecon.torch.projector import projector2d
    projs, angs, imgs = [], [], []
    for i, d in enumerate((62, 63, 64, 126, 127)):
        n = 40 + 4 * i
        projs.append(projector2d(n, d, 0.7 * 4, 1.2858 * 2, 595.0, 1085.6, scan_type="arc").cuda())
        v = 20 + 3 * i
        angs.append(torch.from_numpy((np.arange(v) * (2 * np.pi / v) + 0.1).astype(np.float32)).cuda())
        imgs.append(torch.from_numpy(ellipses2d(n, 1, seed=300 + i, scale=0.04))[0].cuda())
    ub = UnevenBatch2d(projs, angs)
    with torch.no_grad():
        sinos = ub.unpack_sinos(ub.projection(ub.pack_images(imgs)))
        for P, a, im, s in zip(projs, angs, imgs, sinos):
            assert torch.equal(s, P.projection(im[None, None], a)[0, 0])


def test_uneven_graph_replay_and_autotune():
    from pridediff_amd.autotune import tune_uneven
    from pridediff_amd.batched import GraphStep, UnevenBatch2d, data_consistency_grad
    projs, angs, imgs = _uneven(seed=51)
    ub = UnevenBatch2d(projs, angs)
    x = ub.pack_images(imgs)
    with torch.no_grad():
        y = ub.projection(x) * 1.01
    info = {}
    tune_uneven(ub, use_cache=False, report=info)
    print("autotuned grid", info, flush=True)
    # the tuner may pick the running-sum forward, which rounds differently
    # from the pixel walk: the eager reference runs the tuned configuration
    xe = x.clone().requires_grad_(True)
    eager = data_consistency_grad(ub.projection, xe, y).clone()
    step = GraphStep(lambda a, b: data_consistency_grad(ub.projection, a, b),
                     x.clone().requires_grad_(True), y)
    for i in range(3):
        got = step.replay(x, y).clone()
        torch.cuda.synchronize()
        _ = torch.randn(100, device="cuda").sum()
        assert torch.equal(got, eager), f"replay {i}"


def test_uniform_autotune_keeps_results():
    """Over the walk variants the tuned plan gives the untuned bits; with the
    single-call running sums in the sweep it stays within 1e-6 of them."""
    from pridediff_amd.autotune import tune_uniform
    from pridediff_amd.batched import BatchedProjector2d
    P, ang, x, y = _setup(b=6)
    ref = BatchedProjector2d(P, ang, 6, tile=1)
    info = {}
    tuned = tune_uniform(P, ang, 6, use_cache=False, report=info, running_sums=False)
    print("autotuned over the walks", info, flush=True)
    with torch.no_grad():
        assert torch.equal(tuned.projection(x), ref.projection(x))
        assert torch.equal(tuned.projection_t(y), ref.projection_t(y))
    info = {}
    tuned = tune_uniform(P, ang, 6, use_cache=False, report=info)
    print("autotuned with running sums", info, flush=True)
    with torch.no_grad():
        for a, b in ((tuned.projection(x), ref.projection(x)),
                     (tuned.projection_t(y), ref.projection_t(y))):
            assert float((a - b).double().norm() / b.double().norm()) <= 1e-6


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_uneven_block_window_paths_bitwise(scan):
    """The windowed 4 x 4 blocks (DD_ADJ_BLOCK_WINDOW) against the
    per-instance ops bit for bit, on instances whose bins are much narrower
    than a pixel (a 4-pixel run meets more bins than the window holds, so
    the serial tail runs), much wider (few bins per run), and tiny detectors
    (fewer bins than the window: the serial path)."""
    from pridediff_amd import load_radontorch
    from pridediff_amd.batched import UnevenBatch2d
    from pridediff_amd.deeprecon.torch.projector import projector2d
    rt = load_radontorch()
    projs, angs, ys = [], [], []
    gen = torch.Generator(device="cuda").manual_seed(77)
    for i, (n, d, det) in enumerate(((40, 300, 1.2858 / 4), (52, 25, 1.2858 * 6),
                                     (36, 5, 1.2858 * 10), (64, 93, 1.2858 * 2))):
        projs.append(projector2d(n, d, 0.7 * 2, det, 595.0, 1085.6, scan_type=scan).cuda())
        v = 17 + 5 * i
        angs.append(torch.from_numpy((np.arange(v) * (2 * np.pi / v) + 0.03).astype(np.float32)).cuda())
        ys.append(torch.rand(v, d, device="cuda", generator=gen))
    ub = UnevenBatch2d(projs, angs)
    ops = ("projection_t", "backprojection") + (() if scan == "para" else ("backprojection_w",))
    with torch.no_grad():
        for var in (rt.ADJ_BLOCK_WINDOW, rt.ADJ_QUAD):
            ub.plan.set_variant(var)
            for op in ops:
                outs = ub.unpack_images(getattr(ub, op)(ub.pack_sinos(ys)))
                for P, a, y, o in zip(projs, angs, ys, outs):
                    want = getattr(P, op)(y[None, None], a)[0, 0]
                    assert torch.equal(o, want), (scan, var, op, (o - want).abs().max().item())


def test_workgroup_shapes_are_bitwise_equal():
    """Every forward (waves x views per workgroup) and per-pixel adjoint
    (pixel-tile width) workgroup shape the autotuner sweeps gives the default
    shape's results bit for bit, at instance tiles 1, 4 and 16, with the
    epilogue path too."""
    from pridediff_amd.autotune import FWD_SHAPES, PIX_WIDTHS
    from pridediff_amd.batched import BatchedProjector2d
    from pridediff_amd import load_radontorch
    rt = load_radontorch()
    P, ang, x, y = _setup(b=16)
    for tile in (1, 4, 16):
        ref = BatchedProjector2d(P, ang, 16, tile=tile, variant=rt.ADJ_PER_PIXEL)
        with torch.no_grad():
            a0, t0, w0 = ref.projection(x), ref.projection_t(y), ref.backprojection_w(y)
            for wv in FWD_SHAPES:
                for pw in PIX_WIDTHS:
                    bp = BatchedProjector2d(P, ang, 16, tile=tile, variant=rt.ADJ_PER_PIXEL,
                                            shape=(wv[0], wv[1], pw))
                    assert torch.equal(bp.projection(x), a0), (tile, wv, pw)
                    assert torch.equal(bp.projection_t(y), t0), (tile, wv, pw)
                    assert torch.equal(bp.backprojection_w(y), w0), (tile, wv, pw)
