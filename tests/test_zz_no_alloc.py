"""dd2d_run_noalloc (the allocation-free entry point) against dd2d_run, both
called through the C ABI (needs an MI355X).

The no-allocation call takes its scratch from a fixed per-device buffer in
the library; its results must equal dd2d_run's (caller workspace) bit for
bit, on two calls in a row where the second is smaller (the buffer keeps the
first call's contents), for every op of every 2D geometry, and with the
single-call running sums selected.  Problems larger than the buffer
are rejected with DD_TOO_LARGE before any device work.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pridediff_amd", "libdd.so")
DD_OK, DD_TOO_LARGE = 0, 6
TO_SINO = {0, 3, 5}


class Geom(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("num_det", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("pix_x", "pix_y", "det", "iso_source", "source_detector",
                                              "shift_x", "shift_y", "bin_shift")]


@pytest.fixture(scope="module")
def lib():
    import pridediff_amd
    pridediff_amd.load_radontorch()          # libdd binds torch's HIP runtime
    lib = ctypes.CDLL(LIB)
    vp, i = ctypes.c_void_p, ctypes.c_int
    lib.dd2d_workspace_size.restype = ctypes.c_size_t
    lib.dd2d_workspace_size.argtypes = [i, i, vp, i, i]
    lib.dd2d_run.restype = i
    lib.dd2d_run.argtypes = [i, i, vp, vp, vp, vp, i, i, vp, ctypes.c_size_t, vp]
    lib.dd2d_run_noalloc.restype = i
    lib.dd2d_run_noalloc.argtypes = [i, i, vp, vp, vp, vp, i, i, vp]
    return lib


def _geom(scan, n, d, scale):
    if scan == 0:
        return Geom(n, n, d, 1.0 * scale, 1.0 * scale, 1.0 * scale, 1.0, 1.0, 0.3, -0.2, 0.1)
    if scan == 1:
        return Geom(n, n, d, 0.7 * scale, 0.7 * scale, 1.2858 * scale, 595.0, 1085.6, 0.3, -0.2, 0.4)
    return Geom(n, n, d, 0.7 * scale, 0.7 * scale, 1.2858 * scale / 1085.6, 595.0, 1085.6, 0.3, -0.2,
                0.4 / 1085.6)


def _run_both(lib, scan, op, g, batch, views, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    img_shape = (batch, g.height, g.width)
    sino_shape = (batch, views, g.num_det)
    to_sino = op in TO_SINO
    x = torch.rand(img_shape if to_sino else sino_shape, device="cuda", generator=gen)
    ang = (torch.arange(views, device="cuda", dtype=torch.float32) * (2 * np.pi / views) + 0.05)
    out_shape = sino_shape if to_sino else img_shape
    # outputs start as NaN so an unwritten value cannot pass for a result
    out_ws = torch.full(out_shape, float("nan"), device="cuda")
    out_na = torch.full(out_shape, float("nan"), device="cuda")
    ws_bytes = lib.dd2d_workspace_size(scan, op, ctypes.byref(g), batch, views)
    assert ws_bytes > 0
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.dd2d_run(scan, op, ctypes.byref(g), x.data_ptr(), out_ws.data_ptr(), ang.data_ptr(),
                      batch, views, ws.data_ptr(), ws_bytes, stream)
    assert rc == DD_OK, rc
    rc = lib.dd2d_run_noalloc(scan, op, ctypes.byref(g), x.data_ptr(), out_na.data_ptr(),
                              ang.data_ptr(), batch, views, stream)
    assert rc == DD_OK, rc
    torch.cuda.synchronize()
    return out_ws, out_na, ws_bytes


@pytest.mark.parametrize("scan", [0, 1, 2])
def test_noalloc_equals_workspace_call_twice(lib, scan):
    ops = [0, 1, 2, 3] + ([] if scan == 0 else [4, 5])
    for op in ops:
        # first call: a 512x512 slice with 720 views (the fan-arc slice
        # geometry); second call smaller, so stale scratch from the first
        # stays behind it
        for k, (n, d, views, batch, scale) in enumerate(((512, 736, 720, 1, 1.0),
                                                         (96, 141, 90, 3, 4.0))):
            g = _geom(scan, n, d, scale)
            print(f"scan {scan} op {op} call {k}: {n}x{n}, {views} views, {d} bins, batch {batch}",
                  flush=True)
            a, b, ws_bytes = _run_both(lib, scan, op, g, batch, views, seed=100 * scan + 10 * op + k)
            assert ws_bytes <= 64 << 20
            assert not torch.isnan(a).any()
            assert torch.equal(a, b), (scan, op, k, (a - b).abs().max().item())


@pytest.mark.parametrize("scan", [0, 2])
def test_noalloc_with_running_sum_adjoint(lib, scan):
    """With the single-call running sums selected (ProjTrans and BackProj at
    instance tile 1: knots, knot ranges and per-chunk partial images in the
    scratch too; Proj and BackProjTrans: the lines' running-sum tables), the
    same two calls agree bit for bit."""
    lib.dd2d_set_adjoint_variant.argtypes = [ctypes.c_int]
    lib.dd2d_set_adjoint_variant(5 | 0x100)  # DD_ADJ_BLOCK_PREFIX | DD_FWD_PREFIX
    try:
        for op in (0, 1, 2, 3):
            for k, (n, d, views, batch, scale) in enumerate(((512, 736, 720, 1, 1.0),
                                                             (96, 141, 90, 1, 4.0))):
                g = _geom(scan, n, d, scale)
                print(f"running sums: scan {scan} op {op} call {k}: {n}x{n}, {views} views",
                      flush=True)
                a, b, ws_bytes = _run_both(lib, scan, op, g, batch, views,
                                           seed=1000 + 100 * scan + 10 * op + k)
                assert ws_bytes <= 64 << 20
                assert not torch.isnan(a).any()
                assert torch.equal(a, b), (scan, op, k, (a - b).abs().max().item())
    finally:
        lib.dd2d_set_adjoint_variant(-1)


def test_noalloc_rejects_problems_beyond_its_buffer(lib):
    g = _geom(2, 512, 736, 1.0)
    need = lib.dd2d_workspace_size(2, 0, ctypes.byref(g), 16, 720)
    assert need > 64 << 20
    dummy = torch.zeros(1, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.dd2d_run_noalloc(2, 0, ctypes.byref(g), dummy.data_ptr(), dummy.data_ptr(),
                              dummy.data_ptr(), 16, 720, stream)
    assert rc == DD_TOO_LARGE
