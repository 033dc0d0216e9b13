"""Autograd Functions over the radontorch ops.

Same public names and semantics as deeprecon/deeprecon/torch/projector/
function.py:7-553: every Function's forward is one op, its backward is the
transposed op (Proj <-> ProjTrans, BackProj <-> BackProjTrans,
BackProjWeighted -> BackProjTransWeighted), inputs are cast to float32 under
autocast, and options / angles / source_pos get no gradient.  Because the
transposed ops here are the exact (bit-consistent) adjoints, the gradient of
``||A x - y||^2`` is the true gradient of the discrete operator.
"""
import torch
from torch.autograd import Function

from pridediff_amd import load_radontorch

radontorch = load_radontorch()

_fwd = torch.amp.custom_fwd(device_type="cuda", cast_inputs=torch.float32)
_bwd = torch.amp.custom_bwd(device_type="cuda")


def _function(name, op, adjoint, n_geometry):
    """Builds a Function calling `op` forward and `adjoint` backward.

    `n_geometry` is the number of non-differentiable geometry tensors that
    follow the input: 2 (options, angles) or 3 (+ source_pos, spiral).
    """
    fwd_op = getattr(radontorch, op)
    bwd_op = getattr(radontorch, adjoint)

    def forward(ctx, input, *geometry):
        if not input.is_contiguous():
            input = input.contiguous()
        ctx.save_for_backward(*geometry)
        return fwd_op(input, *geometry)

    def backward(ctx, grad_output):
        if not grad_output.is_contiguous():
            grad_output = grad_output.contiguous()
        grad = bwd_op(grad_output, *ctx.saved_tensors)
        return (grad,) + (None,) * n_geometry

    return type(name, (Function,), {
        "forward": staticmethod(_fwd(forward)),
        "backward": staticmethod(_bwd(backward)),
        "__doc__": f"radontorch.{op} with backward radontorch.{adjoint}.",
    })


def _family(suffix, n_geometry, weighted=True):
    pairs = [("Proj", "ProjTrans"), ("ProjTrans", "Proj"),
             ("BackProj", "BackProjTrans"), ("BackProjTrans", "BackProj")]
    if weighted:
        pairs.append(("BackProjWeighted", "BackProjTransWeighted"))
    return {f"{a}{suffix}": _function(f"{a}{suffix}", f"{a}{suffix}", f"{b}{suffix}", n_geometry)
            for a, b in pairs}


_ALL = {}
_ALL.update(_family("FanFlatDis2d", 2))
_ALL.update(_family("FanArcDis2d", 2))
_ALL.update(_family("ParaDis2d", 2, weighted=False))
_ALL.update(_family("ConeFlatDis3d", 2))
_ALL.update(_family("ConeArcDis3d", 2))
_ALL.update(_family("SpiralDis3d", 3))
globals().update(_ALL)


def _sysmatrix(op):
    fn = getattr(radontorch, op)

    def get(option, angles):
        # the reference builds the matrix from CPU copies (function.py:535-553)
        return fn(option.to("cpu").to(torch.float32), angles.to("cpu").to(torch.float32))

    get.__name__ = op
    return get


GetSysMatrixFanFlat2d = _sysmatrix("GetSysMatrixFanFlat2d")
GetSysMatrixFanArc2d = _sysmatrix("GetSysMatrixFanArc2d")
GetSysMatrixPara2d = _sysmatrix("GetSysMatrixPara2d")

__all__ = sorted(_ALL) + ["GetSysMatrixFanFlat2d", "GetSysMatrixFanArc2d", "GetSysMatrixPara2d"]
