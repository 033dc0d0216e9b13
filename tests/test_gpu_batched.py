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
