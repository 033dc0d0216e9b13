"""Batched projector path: one planned launch per op for a whole batch.

``BatchedProjector2d`` wraps ``radontorch.Plan2d``: the per-view geometry is
built once on the device, workspaces are allocated up front, and each op is
two stream-ordered launches (layout + gather) over every instance of the
batch at once, sharing one geometry evaluation across up to 16 instances.
Everything it launches is capturable, so a whole data-consistency step
(``A x``, residual, ``A^T``) can be recorded once with ``GraphStep`` and
replayed as a HIP graph.

No counterpart in the reference, which re-derives geometry and re-creates a
texture on every call (deeprecon/cuda/dis2d/FanArcDisCUDA2d.cu:758-814).
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from pridediff_amd import load_radontorch

rt = load_radontorch()


class _PlannedOp(Function):
    @staticmethod
    def forward(ctx, x, plan, op, adjoint):
        ctx.plan, ctx.adjoint = plan, adjoint
        return plan.run(op, x.contiguous())

    @staticmethod
    def backward(ctx, grad):
        return ctx.plan.run(ctx.adjoint, grad.contiguous()), None, None, None


class BatchedProjector2d(torch.nn.Module):
    """All six 2D ops of ``projector`` for a fixed angle set and batch size.

    ``projector`` is a ``projector2d``; ``batch`` counts B*C instances.
    ``tile``, ``variant`` and ``shape`` = (forward waves per workgroup,
    forward views per workgroup, adjoint pixel-tile width; 0 = built-in)
    are launch choices (``autotune.tune_uniform`` sweeps them); none changes
    a result bit.
    """

    _PAIRS = {"projection": ("PROJ", "PROJ_T"), "projection_t": ("PROJ_T", "PROJ"),
              "backprojection": ("BPROJ", "BPROJ_T"), "backprojection_t": ("BPROJ_T", "BPROJ"),
              "backprojection_w": ("BPROJ_W", "BPROJ_TW")}

    def __init__(self, projector, angles: torch.Tensor, batch: int, tile: int = 0,
                 variant: int = -1, shape=(0, 0, 0)):
        super().__init__()
        self.scan_type = projector.scan_type
        fwd_waves, fwd_views, pix_width = shape
        self.plan = rt.Plan2d(rt.SCAN[projector.scan_type], projector.host_options(),
                              angles.contiguous().to(torch.float32), int(batch), int(tile),
                              int(variant), int(fwd_waves), int(fwd_views), int(pix_width))
        self.batch = int(batch)
        self.tile = self.plan.tile
        self.variant = self.plan.variant
        self.shape = tuple(self.plan.shape)

    def _op(self, name, x):
        a, b = self._PAIRS[name]
        if self.scan_type == "para" and name == "backprojection_w":
            a, b = "BPROJ", "BPROJ_T"
        return _PlannedOp.apply(x, self.plan, getattr(rt, a), getattr(rt, b))

    def projection(self, image):
        return self._op("projection", image)

    def projection_t(self, sino):
        return self._op("projection_t", sino)

    def backprojection(self, sino):
        return self._op("backprojection", sino)

    def backprojection_t(self, image):
        return self._op("backprojection_t", image)

    def backprojection_w(self, sino):
        return self._op("backprojection_w", sino)

    def forward(self, image):
        return self.projection(image)


class UnevenBatch2d(torch.nn.Module):
    """Instances of different image size, detector and view count, all run
    by one launch per op and phase (``radontorch.PlanBatch2d``).

    Images and sinograms travel packed into flat buffers (``pack_images`` /
    ``unpack_sinos`` and friends); the ops are differentiable.
    """

    def __init__(self, projectors, angles, grid: int = 0):
        super().__init__()
        scans = {p.scan_type for p in projectors}
        assert len(scans) == 1, "all instances of a batch must share the scan type"
        self.scan_type = scans.pop()
        self.plan = rt.PlanBatch2d(rt.SCAN[self.scan_type], [p.host_options() for p in projectors],
                                   [a.contiguous().to(torch.float32) for a in angles], int(grid))
        self.shapes = [tuple(s) for s in self.plan.shapes]   # (H, W, V, D)
        self.image_offsets = list(self.plan.image_offsets)
        self.sino_offsets = list(self.plan.sino_offsets)

    def __len__(self):
        return len(self.shapes)

    def pack_images(self, images):
        return torch.cat([im.reshape(-1) for im in images]).contiguous()

    def pack_sinos(self, sinos):
        return torch.cat([s.reshape(-1) for s in sinos]).contiguous()

    def unpack_images(self, flat):
        return [flat[o:o + h * w].view(h, w) for o, (h, w, _, _) in zip(self.image_offsets, self.shapes)]

    def unpack_sinos(self, flat):
        return [flat[o:o + v * d].view(v, d) for o, (_, _, v, d) in zip(self.sino_offsets, self.shapes)]

    def projection(self, flat_images):
        return _PlannedOp.apply(flat_images, self.plan, rt.PROJ, rt.PROJ_T)

    def projection_t(self, flat_sinos):
        return _PlannedOp.apply(flat_sinos, self.plan, rt.PROJ_T, rt.PROJ)

    def backprojection(self, flat_sinos):
        return _PlannedOp.apply(flat_sinos, self.plan, rt.BPROJ, rt.BPROJ_T)

    def backprojection_t(self, flat_images):
        return _PlannedOp.apply(flat_images, self.plan, rt.BPROJ_T, rt.BPROJ)

    def backprojection_w(self, flat_sinos):
        if self.scan_type == "para":  # the reference maps it to BackProj (projector.py:49)
            return self.backprojection(flat_sinos)
        return _PlannedOp.apply(flat_sinos, self.plan, rt.BPROJ_W, rt.BPROJ_TW)

    def forward(self, flat_images):
        return self.projection(flat_images)


def data_consistency_grad(proj, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """Gradient of ||A x - y||^2 through autograd, without a reduction.

    Drives autograd with d/dr ||r||^2 = 2 r on the residual, so the step
    contains only the projector ops and elementwise kernels (nothing that
    records a memset when captured).  Returns ``x.grad``.
    """
    r = proj(x) - y
    torch.autograd.backward(r, grad_tensors=2.0 * r)
    return x.grad


class GraphStep:
    """Records ``fn(*static_inputs)`` into a HIP graph and replays it.

    ``fn`` must only launch capturable work (the projector ops and
    elementwise torch ops).  Inputs are copied into the static buffers
    before each replay; ``replay()`` returns the static outputs.
    """

    def __init__(self, fn, *example_inputs, warmup: int = 2):
        self.static_inputs = [t.detach().clone().requires_grad_(t.requires_grad)
                              for t in example_inputs]
        self.fn = fn
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self._zero_grads()
                fn(*self.static_inputs)
        torch.cuda.current_stream().wait_stream(side)
        self._zero_grads()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.static_outputs = fn(*self.static_inputs)

    def _zero_grads(self):
        for t in self.static_inputs:
            t.grad = None

    def replay(self, *inputs):
        with torch.no_grad():
            for s, t in zip(self.static_inputs, inputs):
                s.copy_(t)
        self.graph.replay()
        return self.static_outputs
