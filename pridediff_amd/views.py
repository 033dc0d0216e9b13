"""View-resolved backprojection (one image per view) in one launch.

deeprecon's VVBP-tensor network builds a [B, V, H, W] stack of single-view
backprojections with a Python loop of V ``projector.backprojection`` calls
(deeprecon/deeprecon/torch/model/vvbp_tensor_net.py:11-24, ``bprj_sv_fun``).
``backprojection_per_view`` computes the same stack with one
``radontorch.views2d`` launch; ``BackProjSingleViews`` mirrors
``bprj_sv_fun`` including its backward (:26-32), which backprojects the
view-summed gradient with BackProjTrans over all angles.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from pridediff_amd import load_radontorch

rt = load_radontorch()

_OPS = {"bprojection": "BPROJ", "projection_t": "PROJ_T", "bprojection_w": "BPROJ_W"}


def backprojection_per_view(projector, sino: torch.Tensor, angles: torch.Tensor,
                            op: str = "bprojection") -> torch.Tensor:
    """[B, C, V, D] -> [B, C*V, H, W]: entry (b, c*V + v) is ``op`` applied to
    view v alone; ``op`` is 'bprojection', 'projection_t' or 'bprojection_w'
    (parallel beam maps 'bprojection_w' to 'bprojection', as the reference's
    dispatch table does)."""
    name = _OPS[op]
    if projector.scan_type == "para" and name == "BPROJ_W":
        name = "BPROJ"
    return rt.views2d(rt.SCAN[projector.scan_type], getattr(rt, name), sino.contiguous(),
                      projector.host_options(), angles.contiguous())


class BackProjSingleViews(Function):
    """``bprj_sv_fun``: forward [B, 1, V, D] -> [B, V, H, W] single-view
    backprojections; backward sums the gradient over views and applies
    BackProjTrans with all angles."""

    @staticmethod
    @torch.amp.custom_fwd(device_type="cuda", cast_inputs=torch.float32)
    def forward(ctx, input, projector, angles):
        ctx.projector = projector
        ctx.save_for_backward(angles)
        return backprojection_per_view(projector, input, angles)

    @staticmethod
    @torch.amp.custom_bwd(device_type="cuda")
    def backward(ctx, grad_output):
        (angles,) = ctx.saved_tensors
        temp = grad_output.contiguous().sum(1, keepdim=True)
        return ctx.projector.backprojection_t(temp, angles), None, None
