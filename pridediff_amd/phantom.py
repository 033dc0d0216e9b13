"""Seeded ellipse / ellipsoid phantoms used as synthetic projector inputs.

Coordinates span [-1, 1] over the grid; row 0 is the top (+y) as in the
reference's image convention (deeprecon/deeprecon/torch/utils/phantom.py:33-61
rasterises from y=+1 down to -1).  Each phantom is a sum of random ellipses
with centres uniform in the unit disk, semi-axes U[0.05, 0.4], rotation
U[0, pi) and intensity U[0.1, 1.0].
"""
from __future__ import annotations

import numpy as np


def ellipses2d(n: int, batch: int = 1, count: int = 10, seed: int = 0,
               scale: float | None = None, height: int | None = None) -> np.ndarray:
    """Returns ``[batch, height, n]`` float32 phantoms.

    ``scale`` rescales each phantom so its maximum equals ``scale``.
    """
    h = n if height is None else height
    rng = np.random.default_rng(seed)
    xs = (np.arange(n) + 0.5) / n * 2.0 - 1.0
    ys = 1.0 - (np.arange(h) + 0.5) / h * 2.0
    X, Y = np.meshgrid(xs, ys)
    out = np.zeros((batch, h, n), dtype=np.float64)
    for b in range(batch):
        for _ in range(count):
            r = np.sqrt(rng.uniform(0.0, 1.0))
            phi = rng.uniform(0.0, 2 * np.pi)
            cx, cy = r * np.cos(phi) * 0.6, r * np.sin(phi) * 0.6
            a, bb = rng.uniform(0.05, 0.4, size=2)
            t = rng.uniform(0.0, np.pi)
            val = rng.uniform(0.1, 1.0)
            ct, st = np.cos(t), np.sin(t)
            u = (X - cx) * ct + (Y - cy) * st
            v = -(X - cx) * st + (Y - cy) * ct
            out[b][(u / a) ** 2 + (v / bb) ** 2 <= 1.0] += val
        if scale is not None and out[b].max() > 0:
            out[b] *= scale / out[b].max()
    return out.astype(np.float32)


def ellipsoids3d(nx: int, ny: int, nz: int, batch: int = 1, count: int = 20,
                 seed: int = 0) -> np.ndarray:
    """Returns ``[batch, nz, ny, nx]`` float32 phantoms (slice 0 at the bottom)."""
    rng = np.random.default_rng(seed)
    xs = (np.arange(nx) + 0.5) / nx * 2.0 - 1.0
    ys = 1.0 - (np.arange(ny) + 0.5) / ny * 2.0
    zs = (np.arange(nz) + 0.5) / nz * 2.0 - 1.0

    def window(coords, lo, hi):
        # indices whose coordinate lies in [lo, hi] (coords monotone), one voxel of margin
        idx = np.nonzero((coords >= lo) & (coords <= hi))[0]
        if idx.size == 0:
            return slice(0, 0)
        return slice(max(int(idx[0]) - 1, 0), min(int(idx[-1]) + 2, coords.size))

    out = np.zeros((batch, nz, ny, nx), dtype=np.float32)
    for b in range(batch):
        acc = out[b]
        for _ in range(count):
            c = rng.uniform(-0.5, 0.5, size=3)
            ax = rng.uniform(0.05, 0.4, size=3)
            t = rng.uniform(0.0, np.pi)
            val = np.float32(rng.uniform(0.1, 1.0))
            ct, st = np.cos(t), np.sin(t)
            # the ellipsoid lies inside the box |x - cx|, |y - cy| <= max(a, b),
            # |z - cz| <= c; the inclusion test is evaluated only there
            rxy = max(ax[0], ax[1])
            sx = window(xs, c[0] - rxy, c[0] + rxy)
            sy = window(ys, c[1] - rxy, c[1] + rxy)
            sz = window(zs, c[2] - ax[2], c[2] + ax[2])
            Z, Y, X = np.meshgrid(zs[sz], ys[sy], xs[sx], indexing="ij")
            u = (X - c[0]) * ct + (Y - c[1]) * st
            v = -(X - c[0]) * st + (Y - c[1]) * ct
            w = Z - c[2]
            acc[sz, sy, sx][(u / ax[0]) ** 2 + (v / ax[1]) ** 2 + (w / ax[2]) ** 2 <= 1.0] += val
    return out
