"""C ABI and host-side logic, on the CPU (no GPU calls).

* libdd.so loads and exports every function declared in include/*.h;
* the radontorch module imports and registers the reference's 37 op names
  (deeprecon/interface/RadonTorch.cpp:1376-1418);
* the Python API packs Options as the reference does
  (deeprecon/deeprecon/torch/projector/projector.py:96-125, :263-298, :437-472)
  and builds the FBP filter and cosine weights (filter.py:12-30,
  projector.py:175-190, :350-371).
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pridediff_amd", "libdd.so")

REFERENCE_OPS = [f"{a}{g}" for g in ("FanFlatDis2d", "FanArcDis2d", "ConeFlatDis3d", "ConeArcDis3d", "SpiralDis3d")
                 for a in ("Proj", "ProjTrans", "BackProj", "BackProjTrans", "BackProjWeighted",
                           "BackProjTransWeighted")] + \
                [f"{a}ParaDis2d" for a in ("Proj", "ProjTrans", "BackProj", "BackProjTrans")] + \
                ["GetSysMatrixFanFlat2d", "GetSysMatrixFanArc2d", "GetSysMatrixPara2d"]


def _declared_functions():
    names = set()
    for h in ("dd.h", "dd3d.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(dd\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    declared = _declared_functions()
    assert len(declared) >= 15, declared
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    lib.dd_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.dd_version()


def test_workspace_queries_are_host_only():
    lib = ctypes.CDLL(LIB)

    class G(ctypes.Structure):
        _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("num_det", ctypes.c_int)] + \
                   [(n, ctypes.c_float) for n in ("pix_x", "pix_y", "det", "iso_source", "source_detector",
                                                  "shift_x", "shift_y", "bin_shift")]
    lib.dd2d_workspace_size.restype = ctypes.c_size_t
    g = G(512, 512, 736, 0.7, 0.7, 1.18e-3, 595.0, 1085.6, 0.0, 0.0, 0.0)
    fp = lib.dd2d_workspace_size(2, 0, ctypes.byref(g), 16, 720)
    adj = lib.dd2d_workspace_size(2, 1, ctypes.byref(g), 16, 720)
    # geometry tables + interleaved image copies / interleaved sinogram
    assert fp >= 2 * 512 * 512 * 16 * 4 and adj >= 720 * 736 * 16 * 4
    assert lib.dd2d_workspace_size(2, 0, ctypes.byref(g), 0, 720) == 0
    # invalid op/scan pairs are rejected before any device work
    lib.dd2d_run.restype = ctypes.c_int
    assert lib.dd2d_run(0, 4, ctypes.byref(g), None, None, None, 1, 1, None, 0, None) == 2  # para has no weighted BP
    assert lib.dd2d_run(9, 0, ctypes.byref(g), None, None, None, 1, 1, None, 0, None) == 1


def test_3d_workspace_query_and_size_limit():
    """3D workspace sizes are host arithmetic; a per-view detector block of
    2^30 floats or more is rejected (the adjoint addresses it with 32-bit
    offsets)."""
    lib = ctypes.CDLL(LIB)

    class G3(ctypes.Structure):
        _fields_ = [(n, ctypes.c_int) for n in ("width", "height", "depth", "num_col", "num_row")] + \
                   [(n, ctypes.c_float) for n in ("pix_x", "pix_y", "pix_z", "col_size", "row_size",
                                                  "iso_source", "source_detector", "shift_x", "shift_y",
                                                  "shift_z", "col_shift", "row_shift")]
    lib.dd3d_workspace_size.restype = ctypes.c_size_t
    g = G3(256, 256, 256, 512, 384, 1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 1500.0, 0, 0, 0, 0, 0)
    fp = lib.dd3d_workspace_size(3, 0, ctypes.byref(g), 2, 360)
    adj = lib.dd3d_workspace_size(3, 1, ctypes.byref(g), 2, 360)
    # two interleaved volume copies (+ one pixel of padding each) / row running sums
    assert fp >= 2 * (256 ** 3 * 2 + 2) * 4
    assert adj >= 360 * (384 + 2) * 512 * 2 * 4
    big = G3(64, 64, 64, 32768, 32768, 1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 1500.0, 0, 0, 0, 0, 0)
    assert lib.dd3d_workspace_size(3, 1, ctypes.byref(big), 1, 4) == 0
    big.num_row = (1 << 29) // 32768 - 3  # (rows + 2) * cols just below 2^29
    assert lib.dd3d_workspace_size(3, 1, ctypes.byref(big), 1, 4) > 0


def test_radontorch_registers_reference_ops():
    from pridediff_amd import load_radontorch
    rt = load_radontorch()
    assert len(REFERENCE_OPS) == 37
    missing = [n for n in REFERENCE_OPS if not hasattr(rt, n)]
    assert not missing, missing


def test_cpu_tensors_are_rejected():
    from pridediff_amd import load_radontorch
    rt = load_radontorch()
    opt = torch.tensor([8, 8, 10, 1, 1, 1, 1, 1, 0, 0, 0], dtype=torch.float32)
    with pytest.raises(RuntimeError, match="HIP device only"):
        rt.ProjParaDis2d(torch.zeros(1, 1, 8, 8), opt, torch.zeros(3))


def test_projector2d_options_and_weights():
    from pridediff_amd.deeprecon.torch.projector import projector2d
    P = projector2d(512, 736, 0.7, 1.2858, 595.0, 1085.6, pixshift=(0.5, -0.25), binshift=2.0, scan_type="arc")
    want = [512, 512, 736, 0.7, 0.7, 1.2858 / 1085.6, 595.0, 1085.6, 0.5, -0.25, 2.0 / 1085.6]
    np.testing.assert_allclose(P.options.detach().numpy(), np.array(want, dtype=np.float32))
    assert P.options.requires_grad is False
    u = (np.arange(736) - 368 + 0.5) * (1.2858 / 1085.6) + 2.0 / 1085.6
    np.testing.assert_allclose(P.weight.detach().numpy(), np.cos(u), rtol=1e-6)
    F = projector2d(64, 90, 1.0, 1.5, 100.0, 180.0, scan_type="flat")
    uf = ((np.arange(90) - 45 + 0.5) * 1.5) * 100.0 / 180.0
    np.testing.assert_allclose(F.weight.detach().numpy(), 100.0 / np.sqrt(uf ** 2 + 100.0 ** 2), rtol=1e-6)
    assert tuple(F.filter.shape) == (1, 1, 1, 179)


def test_ramp_and_shepp_logan_taps():
    from pridediff_amd.deeprecon.torch.projector.filter import filter_generate
    d, n = 0.8, 7
    r = filter_generate(n, d, "ramp", torch.float64).numpy()
    for i, off in enumerate(range(-(n - 1), n)):
        want = 1 / (4 * d * d) if off == 0 else (-1 / (math.pi ** 2 * d * d * off * off) if off % 2 else 0.0)
        assert r[i] == pytest.approx(want, rel=1e-12, abs=1e-15)
    s = filter_generate(n, d, "sheeplogan", torch.float64).numpy()
    offs = np.arange(-(n - 1), n)
    np.testing.assert_allclose(s, -2 / (math.pi ** 2 * d * d * (4 * offs ** 2 - 1)), rtol=1e-12)
    with pytest.raises(AssertionError):
        filter_generate(n, d, "hann")


def test_projector3d_and_spiral_options():
    from pridediff_amd.deeprecon.torch.projector import projector3d, projectorsp
    P = projector3d((256, 240, 200), (512, 384), (1.0, 1.1, 1.2), (1.0, 0.9), 1000.0, 1500.0,
                    pixshift=(0.1, 0.2, 0.3), binshift=(0.4, 0.5), scan_type="flat")
    np.testing.assert_allclose(P.options.detach().numpy(), np.array(
        [256, 240, 200, 512, 384, 1.0, 1.1, 1.2, 1.0, 0.9, 1000.0, 1500.0, 0.1, 0.2, 0.3, 0.4, 0.5],
        dtype=np.float32))
    S = projectorsp((512, 512, 64), (736, 64), (0.7, 0.7, 1.0), (1.2858, 1.82454), 595.0, 1085.6)
    o = S.options.detach().numpy()
    assert o.shape == (16,)
    np.testing.assert_allclose(o[8], np.float32(1.2858 / 1085.6))
    assert tuple(S.weight.shape) == (64, 736)
