"""Pins the CPU restatement (oracle/dd_oracle.cpp) before it is trusted.

* golden vectors written by the reference's own CPU code (tests/golden):
  the FP32-mirror restatement must reproduce them bit for bit;
* the reference library itself (oracle/_ref) on fresh seeded inputs, when it
  has been built in this checkout;
* an independent brute-force FP64 matrix (tests/dd_brute.py) for every
  geometry and op, including detector flips, shifts and both common axes;
* closed forms: parallel-beam mass conservation and the chord length of a
  constant image; adjointness of each op pair.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pridediff_amd.phantom import ellipses2d
from tests import dd_brute

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GEOMS = ["para", "flat", "arc"]


def _golden(geom):
    z = np.load(os.path.join(GOLDEN, f"dd2d_{geom}.npz"))
    opt = z["options"]
    g = O.Geom2D(int(opt[0]), int(opt[1]), int(opt[2]), *[float(v) for v in opt[3:]])
    return z, g


def _ops(geom):
    return ["proj", "proj_t", "bproj", "bproj_t"] + ([] if geom == "para" else ["bproj_w", "bproj_tw"])


@pytest.mark.parametrize("geom", GEOMS)
def test_mirror_reproduces_golden_bitwise(geom):
    z, g = _golden(geom)
    for op in _ops(geom):
        inp = z["image"] if op in O.TO_SINO else z["sino"]
        out = O.op2d(geom, op, inp, z["angles"], g, mode="f32")
        assert np.array_equal(out, z[op]), f"{geom}/{op} differs from the reference"


@pytest.mark.parametrize("geom", GEOMS)
def test_mirror_sysmatrix_matches_golden(geom):
    z, g = _golden(geom)
    rows, cols, vals = O.sysmatrix2d(geom, z["angles"][:8], g)
    assert np.array_equal(rows, z["sys_rows"])
    assert np.array_equal(cols, z["sys_cols"])
    assert np.array_equal(vals, z["sys_vals"])


@pytest.mark.parametrize("geom", GEOMS)
def test_f64_oracle_close_to_golden(geom):
    z, g = _golden(geom)
    for op in _ops(geom):
        inp = z["image"] if op in O.TO_SINO else z["sino"]
        out = O.op2d(geom, op, inp, z["angles"], g, mode="f64")
        assert O.rel_l2(out, z[op]) < 2e-5, (geom, op, O.rel_l2(out, z[op]))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built in this checkout")
@pytest.mark.parametrize("geom", GEOMS)
def test_mirror_matches_reference_library(geom):
    g = {"para": O.Geom2D(40, 36, 61, 0.8, 0.8, 0.9, 1.0, 1.0, 0.3, 0.2, -0.2),
         "flat": O.Geom2D(40, 36, 61, 0.8, 0.8, 1.3, 95.0, 170.0, 0.3, 0.2, -0.4),
         "arc": O.Geom2D(40, 36, 61, 0.8, 0.8, 1.3 / 170.0, 95.0, 170.0, 0.3, 0.2, -0.4 / 170.0)}[geom]
    views = (np.arange(100, dtype=np.float32) * np.float32(2 * np.pi / 100)).astype(np.float32)
    x = ellipses2d(g.W, 2, seed=21, height=g.H)
    y = np.random.default_rng(22).uniform(0, 1, (2, 100, g.D)).astype(np.float32)
    for op in _ops(geom):
        inp = x if op in O.TO_SINO else y
        assert np.array_equal(O.op2d(geom, op, inp, views, g, "f32"),
                              O.op2d(geom, op, inp, views, g, "ref")), (geom, op)


BRUTE_CASES = {
    "para": O.Geom2D(7, 6, 11, 1.1, 0.9, 0.8, 1.0, 1.0, 0.35, -0.2, 0.15),
    "flat": O.Geom2D(7, 6, 13, 1.1, 0.9, 1.4, 20.0, 37.0, 0.35, -0.2, 0.3),
    "arc": O.Geom2D(7, 6, 13, 1.1, 0.9, 1.4 / 37.0, 20.0, 37.0, 0.35, -0.2, 0.3 / 37.0),
}


@pytest.mark.parametrize("geom", GEOMS)
def test_f64_oracle_equals_bruteforce_matrix(geom):
    g = BRUTE_CASES[geom]
    # 20 views over the full circle hit both axes and every flip; offsets keep
    # views away from the exact diagonals, where the two axes legitimately differ
    views = (np.arange(20, dtype=np.float64) * (2 * np.pi / 20) + 0.05).astype(np.float32)
    M = dd_brute.matrices(geom, g, views)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, (2, g.H, g.W))
    y = rng.uniform(0, 1, (2, len(views), g.D))
    for op in _ops(geom):
        inp = x if op in O.TO_SINO else y
        got = O.op2d(geom, op, inp, views, g, mode="f64").reshape(2, -1)
        want = dd_brute.apply(geom, op, M, inp)
        assert O.rel_l2(got, want) < 1e-9, (geom, op, O.rel_l2(got, want))


def test_bruteforce_catches_a_planted_mistake():
    """A dropped detector flip in the walk must fail the brute-force check."""
    g = BRUTE_CASES["arc"]
    views = (np.arange(20, dtype=np.float64) * (2 * np.pi / 20) + 0.05).astype(np.float32)
    M = dd_brute.matrices("arc", g, views)
    x = np.random.default_rng(6).uniform(0, 1, (1, g.H, g.W))
    good = O.op2d("arc", "proj", x, views, g, mode="f64")
    planted = good.copy()
    flipped = [v for v, t in enumerate(views) if np.cos(t) < -0.75]
    planted[:, flipped, :] = planted[:, flipped, ::-1]
    want = dd_brute.apply("arc", "proj", M, x).reshape(good.shape)
    assert O.rel_l2(good, want) < 1e-9
    assert O.rel_l2(planted, want) > 1e-2


@pytest.mark.parametrize("geom", GEOMS)
def test_adjoint_pairs(geom):
    g = BRUTE_CASES[geom]
    views = (np.arange(33, dtype=np.float32) * np.float32(2 * np.pi / 33)).astype(np.float32)
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, (1, g.H, g.W))
    y = rng.uniform(0, 1, (1, len(views), g.D))
    pairs = [("proj", "proj_t"), ("bproj_t", "bproj")]
    if geom != "para":
        pairs.append(("bproj_tw", "bproj_w"))
    for fwd, adj in pairs:
        ax = O.op2d(geom, fwd, x, views, g, mode="f64")
        aty = O.op2d(geom, adj, y, views, g, mode="f64")
        lhs, rhs = float(np.sum(ax * y)), float(np.sum(x * aty))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (geom, fwd, lhs, rhs)


def test_para_mass_conservation():
    """sum_k p[v,k] * det = dx * dy * sum(x) when the detector covers the image."""
    g = O.Geom2D(32, 32, 80, 1.0, 1.0, 1.0, 1.0, 1.0)
    views = (np.arange(45, dtype=np.float32) * np.float32(np.pi / 45)).astype(np.float32)
    x = ellipses2d(32, 1, seed=8)
    p = O.op2d("para", "proj", x, views, g, mode="f64")
    mass = p[0].sum(axis=1) * g.det
    np.testing.assert_allclose(mass, g.dx * g.dy * x.sum(dtype=np.float64), rtol=1e-12)


def test_para_constant_image_chord_length():
    """Central bins of a constant image equal the chord length H*dy/|cos|."""
    g = O.Geom2D(64, 64, 31, 1.0, 1.0, 1.0, 1.0, 1.0)
    th = np.array([0.0, 0.3, -0.5, np.pi - 0.2], dtype=np.float32)
    x = np.ones((1, 64, 64))
    p = O.op2d("para", "proj", x, th, g, mode="f64")[0]
    for v, t in enumerate(th.astype(np.float64)):
        np.testing.assert_allclose(p[v, 15], 64 / abs(np.cos(t)), rtol=1e-12)


def test_reference_fp32_floor_is_measured():
    """The FP32 mirror stays within the FP32 floor of the FP64 oracle."""
    g = O.Geom2D(64, 64, 92, 0.7, 0.7, 1.2858 / 1085.6, 595.0, 1085.6)
    views = (np.arange(90, dtype=np.float32) * np.float32(2 * np.pi / 90)).astype(np.float32)
    x = ellipses2d(64, 1, seed=9)
    a = O.op2d("arc", "proj", x, views, g, "f32")
    b = O.op2d("arc", "proj", x, views, g, "f64")
    assert O.rel_l2(a, b) < 1e-5
