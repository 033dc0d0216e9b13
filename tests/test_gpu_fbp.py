"""FBP and PrideDiff's data-consistency step on the GPU (needs an MI355X).

FP64 reference pipeline, independent of the product path: torch's CPU
conv2d / conv3d in float64 (the reference's own filter call,
projector.py:207, :385), the FP64 oracle's BackProjWeighted
(oracle/dd_oracle*.cpp), and the angular scale |theta_last - theta_0|/(V-1),
halved when redundant (projector.py:210-212).  Bar: rel-L2 <= 1e-5.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pridediff_amd.phantom import ellipses2d

pytestmark = pytest.mark.gpu


def _fbp64_2d(P, g, scan, sino, ang_np, redundant=True):
    taps = P.filter.detach().cpu().double()
    w = P.weight.detach().cpu().double()
    d = sino.shape[-1]
    filt = torch.nn.functional.conv2d(torch.from_numpy(sino).double()[:, None] * w, taps,
                                      padding=(0, d - 1))[:, 0].numpy()
    op = "bproj" if scan == "para" else "bproj_w"
    rec = O.op2d(scan, op, filt, ang_np, g, "f64")
    a = ang_np.astype(np.float64)
    s = abs(a[-1] - a[0]) / (len(a) - 1)
    return rec * (s / 2 if redundant else s)


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_fbp_matches_fp64_pipeline(scan):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    n, d, v = 64, 97, 90
    det = 1.2 if scan != "arc" else 1.2858 * 2
    P = projector2d(n, d, 1.4, det, 300.0, 520.0, scan_type=scan).cuda()
    g = O.Geom2D(n, n, d, 1.4, 1.4, det / 520.0 if scan == "arc" else det, 300.0, 520.0)
    ang_np = (np.arange(v) * (2 * np.pi / v)).astype(np.float32)
    x = ellipses2d(n, 2, seed=8, scale=0.04)
    sino = O.op2d(scan, "proj", x, ang_np, g, "f64").astype(np.float32)
    got = P.filtered_backprojection(torch.from_numpy(sino)[:, None].cuda(),
                                    torch.from_numpy(ang_np).cuda())[:, 0].cpu().numpy()
    want = _fbp64_2d(P, g, scan, sino, ang_np)
    r = O.rel_l2(got, want)
    print(f"{scan} FBP vs FP64 pipeline: rel_l2={r:.3e}", flush=True)
    assert r <= 1e-5


def test_planned_fbp_and_fidelity_step_graph():
    from pridediff_amd.batched import BatchedProjector2d, GraphStep
    from pridediff_amd.deeprecon.torch.projector import projector2d
    from pridediff_amd.fbp import PlannedFBP2d, fidelity_step
    n, d, v, b = 96, 139, 120, 4
    P = projector2d(n, d, 0.7 * 5, 1.2858 * 5, 595.0, 1085.6, scan_type="arc").cuda()
    ang = torch.from_numpy((np.arange(v) * (2 * np.pi / v)).astype(np.float32)).cuda()
    x = torch.from_numpy(ellipses2d(n, b, seed=9, scale=0.04))[:, None].cuda()
    y = (P.projection(x, ang) * 1.01).contiguous()
    BP = BatchedProjector2d(P, ang, b)
    fbp = PlannedFBP2d(P, ang, b, plan=BP)
    with torch.no_grad():
        planned = fbp(y)
        mirror = P.filtered_backprojection(y, ang)
    r = O.rel_l2(planned.cpu().numpy(), mirror.cpu().numpy())
    print(f"planned FBP vs mirror FBP: rel_l2={r:.3e}", flush=True)
    assert r <= 1e-6
    w = torch.tensor(0.3, device="cuda")
    with torch.no_grad():
        eager = fidelity_step(BP.projection, fbp, x, y, w).clone()
        step = GraphStep(lambda a, c: fidelity_step(BP.projection, fbp, a, c, w), x, y)
        for i in range(3):
            got = step.replay(x, y).clone()
            torch.cuda.synchronize()
            _ = torch.randn(1000, device="cuda").sum()  # unrelated eager work between replays
            assert torch.equal(got, eager), f"replay {i}"
    # the update moves x towards consistency: ||A x' - y|| < ||A x - y||
    with torch.no_grad():
        r0 = (P.projection(x, ang) - y).norm().item()
        r1 = (P.projection(eager, ang) - y).norm().item()
    assert r1 < r0


def test_fbp3d_matches_fp64_pipeline():
    from pridediff_amd.deeprecon.torch.projector import projector3d
    g = O.Geom3D(24, 22, 10, 40, 12, 1.0, 1.0, 1.0, 1.3, 1.3, 150.0, 240.0)
    P = projector3d((24, 22, 10), (40, 12), 1.0, 1.3, 150.0, 240.0, scan_type="flat").cuda()
    ang_np = (np.arange(48) * (2 * np.pi / 48)).astype(np.float32)
    rng = np.random.default_rng(10)
    sino = rng.uniform(0, 1, (1, 48, 12, 40)).astype(np.float32)
    got = P.filtered_backprojection(torch.from_numpy(sino)[:, None].cuda(),
                                    torch.from_numpy(ang_np).cuda())[:, 0].cpu().numpy()
    taps = P.filter.detach().cpu().double()
    w = P.weight.detach().cpu().double()
    filt = torch.nn.functional.conv3d(torch.from_numpy(sino).double()[:, None] * w, taps,
                                      padding=(0, 0, 39))[:, 0].numpy()
    rec = O.op3d("cone_flat", "bproj_w", filt, ang_np, g)
    a = ang_np.astype(np.float64)
    want = rec * abs(a[-1] - a[0]) / (len(a) - 1) / 2
    r = O.rel_l2(got, want)
    print(f"cone-flat FBP vs FP64 pipeline: rel_l2={r:.3e}", flush=True)
    assert r <= 1e-5


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_fused_fidelity_step_matches_fp64_composition(scan):
    """x - w FBP(A x - y) with the residual fused into the forward's reduction
    and the update into the backprojection's write, against the same step
    composed in FP64 from the oracle's Proj, torch's CPU float64 conv2d and
    the oracle's BackProjWeighted (model/pridediff.py:56-68); and equal to the
    unfused planned composition.  Replayed from a graph after the weight w
    changes on the device."""
    from pridediff_amd.batched import BatchedProjector2d, GraphStep
    from pridediff_amd.deeprecon.torch.projector import projector2d
    from pridediff_amd.fbp import PlannedFBP2d, fidelity_step, fidelity_step_fused
    n, d, v, b = 80, 117, 96, 3
    det = 1.2 * 4 if scan != "arc" else 1.2858 * 4
    P = projector2d(n, d, 0.7 * 4, det, 595.0, 1085.6, scan_type=scan).cuda()
    g = O.Geom2D(n, n, d, 0.7 * 4, 0.7 * 4, det / 1085.6 if scan == "arc" else det, 595.0, 1085.6)
    ang_np = (np.arange(v) * (2 * np.pi / v)).astype(np.float32)
    ang = torch.from_numpy(ang_np).cuda()
    x_np = ellipses2d(n, b, seed=21, scale=0.04)
    x = torch.from_numpy(x_np)[:, None].cuda()
    y_np = (O.op2d(scan, "proj", x_np, ang_np, g, "f64") * 1.02
            + 0.01 * np.random.default_rng(22).standard_normal((b, v, d))).astype(np.float32)
    y = torch.from_numpy(y_np)[:, None].cuda()
    BP = BatchedProjector2d(P, ang, b)
    fbp = PlannedFBP2d(P, ang, b, plan=BP)
    w = torch.tensor([0.35], device="cuda")
    with torch.no_grad():
        fused = fidelity_step_fused(fbp, x, y, w).clone()
        unfused = fidelity_step(BP.projection, fbp, x, y, w[0]).clone()
    d_fu = O.rel_l2(fused.cpu().numpy(), unfused.cpu().numpy())
    r64 = O.op2d(scan, "proj", x_np, ang_np, g, "f64") - y_np
    want = x_np - 0.35 * _fbp64_2d(P, g, scan, r64, ang_np)
    r = O.rel_l2(fused[:, 0].cpu().numpy(), want)
    # the step itself (x - x'), where the filtered backprojection lives
    rs = O.rel_l2((x - fused)[:, 0].cpu().numpy(), x_np - want)
    print(f"{scan} fused DC step: vs FP64 {r:.3e} (update part {rs:.3e}), vs unfused {d_fu:.3e}",
          flush=True)
    assert r <= 1e-6 and rs <= 1e-5 and d_fu <= 1e-7
    with torch.no_grad():
        step = GraphStep(lambda a, c: fidelity_step_fused(fbp, a, c, w), x, y)
        for i, wv in enumerate((0.35, 0.1, 0.6)):
            w.fill_(wv)
            eager = fidelity_step_fused(fbp, x, y, w).clone()
            got = step.replay(x, y).clone()
            torch.cuda.synchronize()
            _ = torch.randn(1000, device="cuda").sum()
            assert torch.equal(got, eager), f"replay {i} (w = {wv})"
