"""Brute-force distance-driven matrices in numpy, FP64 (test helper).

Independent of the merge walk in oracle/dd_oracle.cpp: every (view, detector
bin, pixel) entry is the overlap of the pixel's and the bin's intervals on the
view's common axis, computed in detector order from the geometric definitions
(deeprecon/cpu/dis2d/FanArcDisCPU2d.cpp:72-93 for the fan-arc maps and weights,
FanFlatDisCPU2d.cpp:67-84 for the flat detector, ParaDisCPU2d.cpp:51-60 for
parallel beam; BackProj weights FanArcDisCPU2d.cpp:451,480,507,
FanFlatDisCPU2d.cpp:19-22,478, ParaDisCPU2d.cpp:202).  Only the common-axis
choice is taken the way the reference takes it (FP32 cosine squared in FP32).
Sizes must stay tiny: the matrices are dense.
"""
from __future__ import annotations

import numpy as np


def _axis(view32: np.float32):
    c = np.float32(np.cos(np.float64(view32)))
    return bool(np.float32(c * c) > 0.5)


def matrices(geom: str, g, angles: np.ndarray):
    """Returns dense ``(M_fp, M_bp, M_bpw)`` of shape ``[V*D, H*W]``."""
    V, D, W, H = len(angles), g.D, g.W, g.H
    # geometry at float32 precision, as the Options tensor carries it
    dx, dy, det, sid, sdd, shx, shy, bin_ = g.f32()
    M_fp = np.zeros((V * D, H * W))
    M_bp = np.zeros_like(M_fp)
    M_bpw = np.zeros_like(M_fp)
    for v, th32 in enumerate(np.asarray(angles, dtype=np.float32)):
        th = np.float64(th32)
        s, c = np.sin(th), np.cos(th)
        sx, sy = -s * sid, c * sid
        xaxis = _axis(th32)
        j = np.arange(D + 1, dtype=np.float64)
        # detector edges mapped onto the common axis, in detector order
        if geom == "arc":
            gam = (j - D / 2.0) * det + bin_ + th
            e = sx + sy * np.tan(gam) if xaxis else sy + sx / np.tan(gam)
        elif geom == "flat":
            vdet, vbin = sid / sdd * det, sid / sdd * bin_
            u = (j - D / 2.0) * vdet + vbin
            px, py = u * c, u * s
            e = (sx * py - sy * px) / (py - sy) if xaxis else (sy * px - sx * py) / (px - sx)
        else:
            u = (j - D / 2.0) * det + bin_
            e = u / c if xaxis else u / s
        lo_b, hi_b = np.minimum(e[:-1], e[1:]), np.maximum(e[:-1], e[1:])
        width = hi_b - lo_b
        mid = 0.5 * (e[:-1] + e[1:])
        if geom == "para":
            w_fp = np.full(D, (dy if xaxis else dx) / det)
        elif xaxis:
            w_fp = dy / (width * (abs(sy) / np.sqrt((mid - sx) ** 2 + sy ** 2)))
        else:
            w_fp = dx / (width * (abs(sx) / np.sqrt((mid - sy) ** 2 + sx ** 2)))
        for r in range(H):
            for col in range(W):
                xc = (col - W / 2.0 + 0.5) * dx + shx
                yc = (H / 2.0 - r - 0.5) * dy + shy
                if xaxis:
                    xe = np.array([xc - dx / 2, xc + dx / 2])
                    if geom == "para":
                        a = xe + yc * s / c
                    else:
                        a = (sx * yc - sy * xe) / (yc - sy)
                else:
                    ye = np.array([yc - dy / 2, yc + dy / 2])
                    if geom == "para":
                        a = ye + xc * c / s
                    else:
                        a = (sy * xc - sx * ye) / (xc - sx)
                lo_p, hi_p = min(a), max(a)
                ov = np.clip(np.minimum(hi_p, hi_b) - np.maximum(lo_p, lo_b), 0.0, None)
                pint = hi_p - lo_p
                if geom == "arc":
                    w_bp = sid * det / pint
                    fbp = 1.0 / ((sx - xc) ** 2 + (sy - yc) ** 2)
                elif geom == "flat":
                    w_bp = sid / sdd * det / pint
                    d = (sx * xc + sy * yc) / sid
                    fbp = sid * sid / (sid - d) ** 2
                else:
                    w_bp = det / pint
                    fbp = 1.0
                rows = v * D + np.arange(D)
                colidx = r * W + col
                M_fp[rows, colidx] = ov * w_fp
                M_bp[rows, colidx] = ov * w_bp
                M_bpw[rows, colidx] = ov * w_bp * fbp
    return M_fp, M_bp, M_bpw


def apply(geom: str, op: str, M, x: np.ndarray):
    """Applies the op through the dense matrices to a batch ``x``."""
    M_fp, M_bp, M_bpw = M
    B = x.shape[0]
    flat = x.reshape(B, -1).astype(np.float64)
    if op == "proj":
        return flat @ M_fp.T
    if op == "proj_t":
        return flat @ M_fp
    if op == "bproj":
        return flat @ M_bp
    if op == "bproj_t":
        return flat @ M_bp.T
    if op == "bproj_w":
        return flat @ M_bpw
    if op == "bproj_tw":
        return flat @ M_bpw.T
    raise ValueError(op)
