"""View-resolved backprojection (needs an MI355X).

Reference behaviour: ``bprj_sv_fun`` (deeprecon/deeprecon/torch/model/
vvbp_tensor_net.py:11-32) loops ``projector.backprojection(input[:, :, v:v+1],
angles[v:v+1])`` over views; its backward applies ``backprojection_t`` to the
view-summed gradient.  ``radontorch.views2d`` must give each view's image bit
for bit equal to that single-view call, and the views must sum to the
all-view op.
"""
import time

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(scan, n=80, d=117, v=60, b=3):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    P = projector2d(n, d, 0.7 * 6, 1.2858 * 6, 595.0, 1085.6, scan_type=scan).cuda()
    ang = torch.from_numpy((np.arange(v) * (2 * np.pi / v) + 0.01).astype(np.float32)).cuda()
    y = torch.rand(b, 1, v, d, device="cuda")
    return P, ang, y


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
@pytest.mark.parametrize("op", ["bprojection", "projection_t", "bprojection_w"])
def test_views_equal_single_view_calls(scan, op):
    from pridediff_amd.views import backprojection_per_view
    P, ang, y = _setup(scan)
    with torch.no_grad():
        per = backprojection_per_view(P, y, ang, op)
        fn = {"bprojection": P.backprojection, "projection_t": P.projection_t,
              "bprojection_w": P.backprojection_w}[op]
        for v in (0, 7, 31, 59):
            single = fn(y[:, :, v:v + 1].contiguous(), ang[v:v + 1].contiguous())
            assert torch.equal(per[:, v:v + 1], single), (scan, op, v)
        full = fn(y, ang)
    r = O.rel_l2(per.sum(1, keepdim=True).double().cpu().numpy(), full.double().cpu().numpy())
    assert r <= 1e-6, r


def test_single_views_function_forward_backward_and_speed():
    from pridediff_amd.views import BackProjSingleViews
    P, ang, y = _setup("arc", n=128, d=185, v=180, b=2)
    x = y.clone().requires_grad_(True)
    out = BackProjSingleViews.apply(x, P, ang)
    g = torch.rand_like(out)
    out.backward(g)
    with torch.no_grad():
        want = P.backprojection_t(g.sum(1, keepdim=True), ang)
    assert torch.equal(x.grad, want)
    # the reference's loop of V single-view calls vs one launch
    with torch.no_grad():
        loop_out = torch.zeros_like(out)
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for v in range(ang.numel()):
                loop_out[:, v:v + 1] = P.backprojection(y[:, :, v:v + 1].contiguous(),
                                                        ang[v:v + 1].contiguous())
            torch.cuda.synchronize()
            t_loop = time.perf_counter() - t0
            t0 = time.perf_counter()
            one = BackProjSingleViews.apply(y, P, ang)
            torch.cuda.synchronize()
            t_one = time.perf_counter() - t0
    assert torch.equal(one, loop_out)
    print(f"{ang.numel()} views: loop {t_loop * 1e3:.2f} ms, one launch {t_one * 1e3:.2f} ms",
          flush=True)


@pytest.mark.parametrize("scan", ["para", "flat", "arc"])
def test_views_match_fp64_single_view_ops(scan):
    """Each view's image of the one-launch view-resolved backprojection
    against the FP64 oracle's op on that one view (bprj_sv_fun,
    vvbp_tensor_net.py:11-24), for the three sinogram -> image ops."""
    from pridediff_amd.views import backprojection_per_view
    P, ang, y = _setup(scan, n=64, d=93, v=24, b=2)
    det = 1.2858 * 6
    g = O.Geom2D(64, 64, 93, 0.7 * 6, 0.7 * 6, det / 1085.6 if scan == "arc" else det,
                 595.0, 1085.6)
    ang_np = ang.cpu().numpy()
    y_np = y[:, 0].cpu().numpy()
    ops = {"bprojection": "bproj", "projection_t": "proj_t", "bprojection_w": "bproj_w"}
    for op, oop in ops.items():
        if scan == "para" and oop == "bproj_w":
            oop = "bproj"   # the reference maps para's weighted BP to BackProj (projector.py:49)
        with torch.no_grad():
            per = backprojection_per_view(P, y, ang, op).cpu().numpy()
        worst = 0.0
        for v in range(0, 24, 5):
            want = O.op2d(scan, oop, np.ascontiguousarray(y_np[:, v:v + 1]), ang_np[v:v + 1], g, "f64")
            r = O.rel_l2(per[:, v], want)
            worst = max(worst, r)
            assert r <= 1e-5, (scan, op, v, r)
        print(f"{scan} views2d {op} vs FP64 single-view op: worst rel_l2 {worst:.3e}", flush=True)
