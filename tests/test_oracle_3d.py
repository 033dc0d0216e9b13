"""Pins the FP64 3D restatement (oracle/dd_oracle3d.cpp).

The reference has no CPU 3D path, so the restatement is pinned by what the
mathematics fixes:
* one centred detector row and one slice that covers it: the 3D forward
  projection and its adjoint reduce to the pinned 2D fan-flat / fan-arc ops
  (the z overlap equals the row height and cancels the weight's row term;
  ConeFlatDisCUDA3d.cu:99-100 vs FanFlatDisCPU2d.cpp:87-93);
* a constant volume: every detector cell whose ray band crosses the whole
  volume in y and stays inside it in z equals the chord length H*dy/cos3D,
  with cos3D the reference's TriAngCos of the cell centre (:21-23, :100);
* exact adjointness of every op pair;
* spiral with zero source tables equals cone-arc with zero z shift
  (SpiralDisCUDA3d.cu:72-83 vs ConeArcDisCUDA3d.cu);
* one exact identity per free shift parameter (PixZShift, BinRowShift, the
  spiral's SourceShiftRadius / SourceShiftAngle / SourceShiftAxial), each
  also run against the oracle with a planted sign slip or dropped factor
  (oracle.planted3d), which it must fail.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _angles(v, off=0.05):
    return (np.arange(v, dtype=np.float64) * (2 * np.pi / v) + off).astype(np.float32)


@pytest.mark.parametrize("scan2d,scan3d", [("flat", "cone_flat"), ("arc", "cone_arc")])
def test_reduces_to_2d_with_one_row(scan2d, scan3d):
    sdd = 180.0
    col = 1.3 if scan2d == "flat" else 1.3 / sdd
    g3 = O.Geom3D(20, 18, 1, 33, 1, 1.1, 0.9, 400.0, col, 2.0, 95.0, sdd, 0.3, -0.2, 0.0,
                  0.4 if scan2d == "flat" else 0.4 / sdd, 0.0)
    g2 = O.Geom2D(20, 18, 33, 1.1, 0.9, col, 95.0, sdd, 0.3, -0.2, g3.csh)
    ang = _angles(24)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (2, 1, 18, 20))
    p3 = O.op3d(scan3d, "proj", x, ang, g3)[:, :, 0, :]
    p2 = O.op2d(scan2d, "proj", x[:, 0], ang, g2, "f64")
    assert O.rel_l2(p3, p2) < 1e-12
    y = rng.uniform(0, 1, (2, 24, 1, 33))
    a3 = O.op3d(scan3d, "proj_t", y, ang, g3)[:, 0]
    a2 = O.op2d(scan2d, "proj_t", y[:, :, 0, :], ang, g2, "f64")
    assert O.rel_l2(a3, a2) < 1e-12


def test_constant_volume_gives_chord_length():
    g = O.Geom3D(24, 24, 40, 41, 9, 1.0, 1.0, 1.0, 1.0, 1.0, 300.0, 450.0)
    ang = np.array([0.1, 0.3, np.pi - 0.2, np.pi + 0.4], dtype=np.float32)
    p = O.op3d("cone_flat", "proj", np.ones((1, 40, 24, 24)), ang, g)[0]
    sid, sdd = np.float64(np.float32(300.0)), np.float64(np.float32(450.0))
    for v, th in enumerate(ang.astype(np.float64)):
        s, c = np.sin(th), np.cos(th)
        sx, sy = -s * sid, c * sid
        flip = c < 0
        for r in range(3, 6):           # central rows stay inside the volume in z
            for kk in range(17, 24):    # central columns cross the whole volume in y
                k = 41 - 1 - kk if flip else kk
                off0, off1 = ((41 / 2 - k, 41 / 2 - k - 1) if flip else (k - 41 / 2, k + 1 - 41 / 2))
                offc = 0.5 * (off0 + off1)
                vd = sid / sdd
                def axis(off):
                    u = off * vd
                    px, py = u * c, u * s
                    return (sx * py - sy * px) / (py - sy)
                e0, e1 = axis(off0), axis(off1)
                m = sy / (sy - offc * vd * s)
                z0 = m * (r - 4.5) * vd
                z1 = m * (r + 1 - 4.5) * vd
                a, b = 0.5 * (e0 + e1) - sx, 0.5 * (z0 + z1)
                cos3 = abs(sy) / np.sqrt(a * a + b * b + sy * sy)
                assert p[v, r, kk] == pytest.approx(24 * 1.0 / cos3, rel=1e-9), (v, r, kk)


@pytest.mark.parametrize("scan", ["cone_flat", "cone_arc", "spiral"])
def test_adjoint_pairs_3d(scan):
    sdd = 210.0
    col = 1.4 if scan == "cone_flat" else 1.4 / sdd
    g = O.Geom3D(14, 12, 10, 25, 7, 1.0, 1.1, 1.2, col, 1.5, 120.0, sdd, 0.2, -0.3,
                 0.1 if scan != "spiral" else 0.0, 0.2 if scan == "cone_flat" else 0.2 / sdd, 0.3)
    ang = _angles(16)
    spos = None
    if scan == "spiral":
        spos = np.stack([np.linspace(-4, 4, 16), 0.5 * np.sin(np.arange(16)),
                         0.01 * np.cos(np.arange(16)), 0.2 * np.ones(16)]).astype(np.float32)
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, (1, 10, 12, 14))
    y = rng.uniform(0, 1, (1, 16, 7, 25))
    for fwd, adj in (("proj", "proj_t"), ("bproj_t", "bproj"), ("bproj_tw", "bproj_w")):
        ax = O.op3d(scan, fwd, x, ang, g, spos)
        aty = O.op3d(scan, adj, y, ang, g, spos)
        lhs, rhs = float(np.sum(ax * y)), float(np.sum(x * aty))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (scan, fwd, lhs, rhs)
        assert abs(lhs) > 0


def test_spiral_without_shifts_equals_cone_arc():
    sdd = 210.0
    g = O.Geom3D(14, 12, 10, 25, 7, 1.0, 1.1, 1.2, 1.4 / sdd, 1.5, 120.0, sdd, 0.2, -0.3, 0.0,
                 0.2 / sdd, 0.3)
    ang = _angles(12)
    zero = np.zeros((4, 12), dtype=np.float32)
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, (1, 10, 12, 14))
    y = rng.uniform(0, 1, (1, 12, 7, 25))
    for op, inp in (("proj", x), ("proj_t", y), ("bproj", y), ("bproj_w", y), ("bproj_tw", x)):
        a = O.op3d("spiral", op, inp, ang, g, zero)
        b = O.op3d("cone_arc", op, inp, ang, g)
        assert O.rel_l2(a, b) < 1e-13, op


def test_spiral_axial_shift_moves_the_projection():
    """Moving source and detector up by one slice equals moving the volume
    down: exact for the BackProj pair; for Proj only up to the FP weight's
    cosine, which the reference evaluates at the absolute row-centre z
    (SpiralDisCUDA3d.cu:111), not relative to the source."""
    sdd = 210.0
    g = O.Geom3D(12, 12, 12, 21, 5, 1.0, 1.0, 1.0, 1.4 / sdd, 1.2, 120.0, sdd)
    ang = _angles(8)
    x = np.zeros((1, 12, 12, 12))
    x[0, 3:8, 4:9, 3:9] = np.random.default_rng(3).uniform(0.5, 1, (5, 5, 6))
    shifted = np.roll(x, 1, axis=1)
    spos = np.zeros((4, 8), dtype=np.float32)
    spos[0, :] = 1.0
    zero = np.zeros((4, 8), dtype=np.float32)
    a = O.op3d("spiral", "bproj_t", shifted, ang, g, spos)
    b = O.op3d("spiral", "bproj_t", x, ang, g, zero)
    assert O.rel_l2(a, b) < 1e-12
    a = O.op3d("spiral", "proj", shifted, ang, g, spos)
    b = O.op3d("spiral", "proj", x, ang, g, zero)
    assert 1e-7 < O.rel_l2(a, b) < 1e-3


# ---------------------------------------------------------------------------
# Exact identities for the free shift parameters.  All geometry values are
# dyadic (exact in float32, and in every sum and product the identities
# rearrange), so the two sides differ only by FP64 rounding; each identity is
# also shown to fail when the oracle carries a planted sign slip.

def _dyadic_views(v, step=0.125, off=0.0625):
    return (np.arange(v, dtype=np.float64) * step + off).astype(np.float32)


def _g3(scan, **kw):
    sdd = 192.0
    base = dict(W=14, H=12, Dz=10, C=25, R=7, dx=1.0, dy=1.125, dz=1.25,
                col=1.5 if scan == "cone_flat" else 2.0 ** -6, row=1.5, sid=96.0, sdd=sdd,
                shx=0.25, shy=-0.375, shz=0.0,
                csh=0.25 if scan == "cone_flat" else 2.0 ** -7, rsh=0.0)
    base.update(kw)
    return O.Geom3D(**base)


def _rand(shape, seed):
    return np.random.default_rng(seed).uniform(0.0, 1.0, shape)


def _assert_identity(check, plant):
    """check() returns the relative L2 mismatch of the identity."""
    assert check() < 1e-12
    with O.planted3d(plant):
        assert check() > 1e-6, plant


@pytest.mark.parametrize("scan", ["cone_flat", "cone_arc"])
def test_pix_z_shift_moves_the_volume_one_slice(scan):
    """PixZShift = dz puts slice s where slice s+1 was
    (ConeFlatDisCUDA3d.cu:103-104, ConeArcDisCUDA3d.cu:105-106): projecting
    the volume moved down one slice with the shift equals projecting it
    unshifted, and the backprojections are the same volumes moved by one."""
    g0 = _g3(scan)
    g1 = _g3(scan, shz=g0.dz)
    ang = _dyadic_views(16)
    x = _rand((2, g0.Dz, g0.H, g0.W), 60)
    x[:, 0] = 0.0                                   # nothing leaves the grid
    x_down = np.roll(x, -1, axis=1)
    y = _rand((2, 16, g0.R, g0.C), 61)

    def check():
        errs = []
        for op in ("proj", "bproj_t", "bproj_tw"):
            errs.append(O.rel_l2(O.op3d(scan, op, x_down, ang, g1), O.op3d(scan, op, x, ang, g0)))
        for op in ("proj_t", "bproj", "bproj_w"):
            a = O.op3d(scan, op, y, ang, g1)[:, :-1]
            b = O.op3d(scan, op, y, ang, g0)[:, 1:]
            errs.append(O.rel_l2(a, b))
        return max(errs)

    _assert_identity(check, "pix_z_shift")


@pytest.mark.parametrize("scan", ["cone_flat", "cone_arc", "spiral"])
def test_bin_row_shift_moves_the_rows_by_one(scan):
    """BinRowShift = one row pitch makes row r the old row r+1
    (ConeFlatDisCUDA3d.cu:68, ConeArcDisCUDA3d.cu:73, SpiralDisCUDA3d.cu:82)."""
    g0 = _g3(scan)
    g1 = _g3(scan, rsh=g0.row)
    ang = _dyadic_views(16)
    spos = None
    if scan == "spiral":
        spos = np.stack([np.arange(16) * 0.25 - 2.0, np.zeros(16), np.zeros(16),
                         np.full(16, 0.5)]).astype(np.float32)
    x = _rand((2, g0.Dz, g0.H, g0.W), 62)
    y = _rand((2, 16, g0.R, g0.C), 63)
    y[:, :, 0] = 0.0
    y_down = np.roll(y, -1, axis=2)

    def check():
        errs = []
        for op in ("proj", "bproj_t", "bproj_tw"):
            a = O.op3d(scan, op, x, ang, g1, spos)[:, :, :-1]
            b = O.op3d(scan, op, x, ang, g0, spos)[:, :, 1:]
            errs.append(O.rel_l2(a, b))
        for op in ("proj_t", "bproj", "bproj_w"):
            errs.append(O.rel_l2(O.op3d(scan, op, y_down, ang, g1, spos),
                                 O.op3d(scan, op, y, ang, g0, spos)))
        return max(errs)

    _assert_identity(check, "bin_row_shift")


ALL_OPS = ("proj", "proj_t", "bproj", "bproj_t", "bproj_w", "bproj_tw")


def _all_ops_rel(run_a, run_b, x, y):
    return max(O.rel_l2(run_a(op, x if op in O.TO_SINO else y),
                        run_b(op, x if op in O.TO_SINO else y)) for op in ALL_OPS)


@pytest.mark.parametrize("plant", ["radius", "col_pitch"])
def test_source_shift_radius_equals_cone_arc_with_moved_source(plant):
    """SourceShiftRadius c (SpiralDisCUDA3d.cu:72-75) moves source and detector
    out by c and keeps the detector's absolute column pitch: cone-arc with
    SID + c, SDD + c and column pitch and shift scaled by SDD / (SDD + c)."""
    c = 64.0                                         # SDD / (SDD + c) = 0.75
    gs = _g3("spiral")
    ga = _g3("cone_arc", sid=gs.sid + c, sdd=gs.sdd + c, col=gs.col * 0.75, csh=gs.csh * 0.75)
    ang = _dyadic_views(16)
    spos = np.zeros((4, 16), dtype=np.float32)
    spos[1] = c
    x = _rand((1, gs.Dz, gs.H, gs.W), 64)
    y = _rand((1, 16, gs.R, gs.C), 65)
    _assert_identity(lambda: _all_ops_rel(lambda op, t: O.op3d("spiral", op, t, ang, gs, spos),
                                          lambda op, t: O.op3d("cone_arc", op, t, ang, ga), x, y),
                     plant)


@pytest.mark.parametrize("plant", ["angle_source", "angle_columns"])
def test_source_shift_angle_equals_cone_arc_at_the_shifted_view(plant):
    """SourceShiftAngle d (flying focal spot, SpiralDisCUDA3d.cu:75-78)
    rotates the source by d while the detector stays: cone-arc at view
    angle + d with column shift - d (SDD - SID) / SDD - d."""
    d = 2.0 ** -6
    gs = _g3("spiral")                               # (SDD - SID) / SDD = 0.5
    ga = _g3("cone_arc", csh=gs.csh - 0.5 * d - d)
    ang = _dyadic_views(16)
    spos = np.zeros((4, 16), dtype=np.float32)
    spos[2] = d
    ang_a = (ang + np.float32(d)).astype(np.float32)
    x = _rand((1, gs.Dz, gs.H, gs.W), 66)
    y = _rand((1, 16, gs.R, gs.C), 67)
    _assert_identity(lambda: _all_ops_rel(lambda op, t: O.op3d("spiral", op, t, ang, gs, spos),
                                          lambda op, t: O.op3d("cone_arc", op, t, ang_a, ga), x, y),
                     plant)


def test_source_shift_axial_trades_with_axial_position_and_row_shift():
    """SourceShiftAxial s moves only the source (SpiralDisCUDA3d.cu:81-82):
    (axial a, shift s, BinRowShift r) equals (axial a + s, shift 0,
    BinRowShift r - s)."""
    s, r = 0.25, 0.5
    g1 = _g3("spiral", rsh=r)
    g2 = _g3("spiral", rsh=r - s)
    ang = _dyadic_views(16)
    axial = (np.arange(16) * 0.25 - 2.0).astype(np.float32)
    sp1 = np.stack([axial, np.zeros(16), np.zeros(16), np.full(16, s)]).astype(np.float32)
    sp2 = np.stack([axial + np.float32(s), np.zeros(16), np.zeros(16), np.zeros(16)]).astype(np.float32)
    x = _rand((1, g1.Dz, g1.H, g1.W), 68)
    y = _rand((1, 16, g1.R, g1.C), 69)
    _assert_identity(lambda: _all_ops_rel(lambda op, t: O.op3d("spiral", op, t, ang, g1, sp1),
                                          lambda op, t: O.op3d("spiral", op, t, ang, g2, sp2), x, y),
                     "axial")
