"""Writes the 2D golden vectors from the reference's own CPU projector code.

Runs the reference's deeprecon/cpu/dis2d sources, compiled in place into
oracle/_ref/libdd_ref2d.so by ``make -C oracle ref``, on small seeded inputs
and stores inputs and outputs as ``tests/golden/dd2d_<geom>.npz`` (plain
arrays, no pickles).  The view counts stay at or below 128 so the reference
runs single-threaded and its transpose ops sum in a fixed order.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pridediff_amd.phantom import ellipses2d  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# Small geometries with non-square images, odd/even detector counts and
# non-zero pixel and bin shifts so every branch of the walk is exercised.
CASES = {
    "para": dict(g=O.Geom2D(24, 20, 37, 1.1, 0.9, 1.05, 1.0, 1.0, 0.6, -0.4, 0.3),
                 views=np.arange(64, dtype=np.float32) * np.float32(np.pi / 64)),
    "flat": dict(g=O.Geom2D(24, 20, 41, 1.1, 0.9, 1.7, 60.0, 110.0, 0.6, -0.4, 0.8),
                 views=np.arange(72, dtype=np.float32) * np.float32(2 * np.pi / 72)),
    "arc": dict(g=O.Geom2D(24, 20, 41, 1.1, 0.9, 1.7 / 110.0, 60.0, 110.0, 0.6, -0.4, 0.8 / 110.0),
                views=np.arange(72, dtype=np.float32) * np.float32(2 * np.pi / 72)),
}

OPS = {"para": ["proj", "proj_t", "bproj", "bproj_t"],
       "flat": list(O.OPS), "arc": list(O.OPS)}


def main():
    if not O.ref_available():
        raise SystemExit("build the reference library first: make -C oracle ref")
    for geom, case in CASES.items():
        g, views = case["g"], case["views"].astype(np.float32)
        x = ellipses2d(g.W, 2, count=6, seed=11, height=g.H)
        y = np.random.default_rng(12).uniform(0.0, 1.0, (2, len(views), g.D)).astype(np.float32)
        arrays = {"options": np.array(g.options(), dtype=np.float32), "angles": views,
                  "image": x, "sino": y}
        for op in OPS[geom]:
            inp = x if op in O.TO_SINO else y
            arrays[op] = O.op2d(geom, op, inp, views, g, mode="ref")
        rows, cols, vals = O.sysmatrix2d(geom, views[:8], g, mode="ref")
        arrays.update(sys_rows=rows, sys_cols=cols, sys_vals=vals)
        path = os.path.join(HERE, f"dd2d_{geom}.npz")
        np.savez_compressed(path, **arrays)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
