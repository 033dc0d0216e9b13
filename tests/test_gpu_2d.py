"""GPU parity of the 2D HIP path against the CPU oracle (needs an MI355X).

Tolerances (FP32 results):
* vs the FP64 oracle (the same algorithm, same axis decisions, all FP64):
  relative L2 <= 1e-5;
* vs the reference's own CPU code (golden vectors from oracle/_ref):
  relative L2 <= max(1e-5, 2 x the reference's own distance to the FP64
  oracle), because the reference accumulates pixel-edge positions in FP32
  and carries that error itself;
* adjoint dot-product test <Ax, y> = <x, A^T y>: 1e-6 relative, inner
  products accumulated in FP64.
Max-abs errors are printed with every comparison.
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pridediff_amd.phantom import ellipses2d

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
OPNAMES = {"proj": "Proj", "proj_t": "ProjTrans", "bproj": "BackProj", "bproj_t": "BackProjTrans",
           "bproj_w": "BackProjWeighted", "bproj_tw": "BackProjTransWeighted"}
GEOMNAMES = {"para": "ParaDis2d", "flat": "FanFlatDis2d", "arc": "FanArcDis2d"}


@pytest.fixture(scope="module")
def rt():
    from pridediff_amd import load_radontorch
    return load_radontorch()


def _ops(geom):
    return ["proj", "proj_t", "bproj", "bproj_t"] + ([] if geom == "para" else ["bproj_w", "bproj_tw"])


def gpu_op(rt, geom, op, x, angles, g):
    fn = getattr(rt, OPNAMES[op] + GEOMNAMES[geom])
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))[:, None].cuda()
    at = torch.from_numpy(np.ascontiguousarray(angles, dtype=np.float32)).cuda()
    opt = torch.tensor(g.options(), dtype=torch.float32)
    out = fn(xt, opt, at)
    torch.cuda.synchronize()
    return out[:, 0].cpu().numpy()


def report(tag, got, want):
    r, m = O.rel_l2(got, want), O.max_abs(got, want)
    print(f"{tag}: rel_l2={r:.3e} max_abs={m:.3e}", flush=True)
    return r


def _golden(geom):
    z = np.load(os.path.join(GOLDEN, f"dd2d_{geom}.npz"))
    opt = z["options"]
    return z, O.Geom2D(int(opt[0]), int(opt[1]), int(opt[2]), *[float(v) for v in opt[3:]])


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
def test_all_ops_match_oracle_and_reference(rt, geom):
    z, g = _golden(geom)
    ang = z["angles"]
    for op in _ops(geom):
        inp = z["image"] if op in O.TO_SINO else z["sino"]
        got = gpu_op(rt, geom, op, inp, ang, g)
        f64 = O.op2d(geom, op, inp, ang, g, "f64")
        floor = O.rel_l2(z[op], f64)
        assert report(f"{geom}/{op} vs f64", got, f64) <= 1e-5
        assert report(f"{geom}/{op} vs reference (floor {floor:.2e})", got, z[op]) <= max(1e-5, 2 * floor)


def _c1_geom(n=512, d=736):
    return O.Geom2D(n, n, d, 0.7, 0.7, 1.2858 / 1085.6, 595.0, 1085.6)


def _angles(v, full=2 * np.pi):
    return (np.arange(v, dtype=np.float32) * np.float32(full / v)).astype(np.float32)


def test_fan_arc_c1_projection_parity(rt):
    """512^2, 720 views, 736 bins (the graded fan-arc geometry)."""
    g, ang = _c1_geom(), _angles(720)
    x = ellipses2d(512, 2, seed=3, scale=0.04)
    got = gpu_op(rt, "arc", "proj", x, ang, g)
    f64 = O.op2d("arc", "proj", x[:1], ang, g, "f64")
    assert report("fan-arc 512^2 proj vs f64", got[:1], f64) <= 1e-5
    if O.ref_available():
        ref = O.op2d("arc", "proj", x[:1], ang, g, "ref")
        floor = O.rel_l2(ref, f64)
        assert report(f"fan-arc 512^2 proj vs reference (floor {floor:.2e})", got[:1], ref) <= max(1e-5, 2 * floor)


def test_fan_arc_c1_adjoint_parity(rt):
    g, ang = _c1_geom(), _angles(720)
    y = np.random.default_rng(4).uniform(0, 1, (1, 720, 736)).astype(np.float32)
    got = gpu_op(rt, "arc", "proj_t", y, ang, g)
    f64 = O.op2d("arc", "proj_t", y, ang, g, "f64")
    assert report("fan-arc 512^2 proj_t vs f64", got, f64) <= 1e-5
    if O.ref_available():
        ref = O.op2d("arc", "proj_t", y, ang, g, "ref")
        floor = O.rel_l2(ref, f64)
        assert report(f"fan-arc 512^2 proj_t vs reference (floor {floor:.2e})", got, ref) <= max(1e-5, 2 * floor)


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
def test_dot_product_adjoint(rt, geom):
    g = {"para": O.Geom2D(96, 80, 141, 0.9, 1.1, 1.0, 1.0, 1.0, 0.4, -0.3, 0.2),
         "flat": O.Geom2D(96, 80, 141, 0.9, 1.1, 1.6, 190.0, 350.0, 0.4, -0.3, 0.5),
         "arc": O.Geom2D(96, 80, 141, 0.9, 1.1, 1.6 / 350.0, 190.0, 350.0, 0.4, -0.3, 0.5 / 350.0)}[geom]
    ang = _angles(150)
    rng = np.random.default_rng(10)
    x = rng.uniform(0, 1, (3, g.H, g.W)).astype(np.float32)
    y = rng.uniform(0, 1, (3, 150, g.D)).astype(np.float32)
    pairs = [("proj", "proj_t"), ("bproj_t", "bproj")] + ([] if geom == "para" else [("bproj_tw", "bproj_w")])
    for fwd, adj in pairs:
        ax = gpu_op(rt, geom, fwd, x, ang, g).astype(np.float64)
        aty = gpu_op(rt, geom, adj, y, ang, g).astype(np.float64)
        lhs, rhs = float(np.sum(ax * y)), float(np.sum(x * aty))
        rel = abs(lhs - rhs) / abs(lhs)
        print(f"{geom} <{fwd} x, y> = {lhs:.10e}  <x, {adj} y> = {rhs:.10e}  rel={rel:.2e}", flush=True)
        assert rel <= 1e-6


def test_batch_tiles_are_consistent(rt):
    """A batch of 16 in one call equals 16 single-instance calls bit for bit."""
    g, ang = O.Geom2D(128, 128, 185, 0.7, 0.7, 1.2858 / 1085.6, 595.0, 1085.6), _angles(90)
    x = ellipses2d(128, 16, seed=12)
    y = np.random.default_rng(13).uniform(0, 1, (16, 90, 185)).astype(np.float32)
    for op, inp in (("proj", x), ("proj_t", y)):
        whole = gpu_op(rt, "arc", op, inp, ang, g)
        single = np.concatenate([gpu_op(rt, "arc", op, inp[i:i + 1], ang, g) for i in range(16)])
        assert np.array_equal(whole, single), op


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
def test_batch_tiles_on_partial_pixel_tiles(rt, geom):
    """A batch of 16 (quad-major copies, built through 16 x 16-pixel LDS
    tiles) and a batch of 3 (tile 4) equal single-instance calls bit for bit
    on a rectangular image whose sides are not multiples of 16."""
    if geom == "para":
        g = O.Geom2D(70, 53, 97, 1.1, 1.1, 1.3, 1.0, 1.0, 0.4, -0.6, 0.2)
    elif geom == "flat":
        g = O.Geom2D(70, 53, 97, 1.1, 1.1, 1.3, 595.0, 1085.6, 0.4, -0.6, 0.2)
    else:
        g = O.Geom2D(70, 53, 97, 1.1, 1.1, 1.3 / 1085.6, 595.0, 1085.6, 0.4, -0.6, 0.2 / 1085.6)
    ang = _angles(40)
    rng = np.random.default_rng(31)
    ops = ("proj", "bproj_t") if geom == "para" else ("proj", "bproj_t", "bproj_tw")
    for batch in (16, 3):
        x = rng.random((batch, 53, 70)).astype(np.float32)
        for op in ops:
            whole = gpu_op(rt, geom, op, x, ang, g)
            single = np.concatenate([gpu_op(rt, geom, op, x[i:i + 1], ang, g)
                                     for i in range(batch)])
            assert np.array_equal(whole, single), (geom, batch, op)


def test_autograd_gradient_and_finite_differences(rt):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    n, d, v = 64, 92, 60
    P = projector2d(n, d, 0.7 * 8, 1.2858 * 8, 595.0, 1085.6, scan_type="arc").cuda()
    ang = torch.from_numpy(_angles(v)).cuda()
    xt = torch.from_numpy(ellipses2d(n, 2, seed=14, scale=0.04))[:, None].cuda()
    y = P.projection(xt, ang).detach()
    y = y + 0.01 * torch.randn_like(y)
    x = (xt + 0.005 * torch.randn_like(xt)).requires_grad_(True)
    loss = ((P.projection(x, ang) - y) ** 2).sum()
    loss.backward()
    grad = x.grad.detach().double().cpu().numpy()
    # gradient = 2 A^T (A x - y): compare with the FP64 oracle
    g = O.Geom2D(n, n, d, 0.7 * 8, 0.7 * 8, 1.2858 * 8 / 1085.6, 595.0, 1085.6)
    xd = x.detach()[:, 0].double().cpu().numpy()
    r = O.op2d("arc", "proj", xd, _angles(v), g, "f64") - y[:, 0].double().cpu().numpy()
    want = 2 * O.op2d("arc", "proj_t", r, _angles(v), g, "f64")
    assert report("grad vs 2 A^T(Ax-y) FP64", grad[:, 0], want) <= 1e-4
    # central finite difference along a random direction (exact for a quadratic)
    dirn = torch.randn_like(x)
    eps = 1e-2
    with torch.no_grad():
        lp = ((P.projection(x + eps * dirn, ang) - y) ** 2).double().sum()
        lm = ((P.projection(x - eps * dirn, ang) - y) ** 2).double().sum()
    fd = float((lp - lm) / (2 * eps))
    an = float((x.grad.double() * dirn.double()).sum())
    print(f"finite difference {fd:.8e} vs analytic {an:.8e}", flush=True)
    assert abs(fd - an) <= 1e-3 * abs(an)


def test_autocast_inputs_are_cast_to_float32(rt):
    """Under autocast the autograd Functions cast bf16 inputs to FP32
    (custom_fwd(cast_inputs=float32), function.py:4-9 of the reference): the
    result is FP32 and equals the FP32 call on the rounded input, and the
    backward runs."""
    from pridediff_amd.deeprecon.torch.projector import projector2d
    n, d, v = 64, 92, 60
    P = projector2d(n, d, 0.7 * 8, 1.2858 * 8, 595.0, 1085.6, scan_type="arc").cuda()
    ang = torch.from_numpy(_angles(v)).cuda()
    x = torch.from_numpy(ellipses2d(n, 2, seed=15, scale=0.04))[:, None].cuda()
    xb = x.to(torch.bfloat16).requires_grad_(True)
    with torch.autocast("cuda", dtype=torch.bfloat16):
        out = P.projection(xb, ang)
    assert out.dtype == torch.float32
    want = P.projection(xb.detach().float(), ang)
    assert torch.equal(out, want)
    out.sum().backward()
    assert xb.grad is not None and xb.grad.dtype == torch.bfloat16


@pytest.mark.parametrize("geom", ["para", "flat", "arc"])
def test_system_matrix_matches_oracle(rt, geom):
    z, g = _golden(geom)
    ang = torch.from_numpy(z["angles"][:8].copy())
    opt = torch.tensor(g.options(), dtype=torch.float32)
    fn = getattr(rt, "GetSysMatrix" + {"para": "Para2d", "flat": "FanFlat2d", "arc": "FanArc2d"}[geom])
    m = fn(opt, ang).coalesce().to_dense().double().numpy()
    want = np.zeros_like(m)
    np.add.at(want, (z["sys_rows"], z["sys_cols"]), z["sys_vals"].astype(np.float64))
    assert report(f"{geom} system matrix vs reference", m, want) <= 2e-5


def test_errors_are_reported(rt):
    g = O.Geom2D(32, 32, 40, 1.0, 1.0, 1.0, 1.0, 1.0)
    opt = torch.tensor(g.options(), dtype=torch.float32)
    ang = torch.zeros(4).cuda()
    with pytest.raises(RuntimeError, match="HIP device only"):
        rt.ProjParaDis2d(torch.zeros(1, 1, 32, 32), opt, ang.cpu())
    with pytest.raises(RuntimeError, match="geometry"):
        rt.ProjParaDis2d(torch.zeros(1, 1, 31, 32).cuda(), opt, ang)
    with pytest.raises(RuntimeError, match="float32"):
        rt.ProjParaDis2d(torch.zeros(1, 1, 32, 32, dtype=torch.float64).cuda(), opt, ang)
    with pytest.raises(RuntimeError, match="sinogram"):
        rt.ProjTransParaDis2d(torch.zeros(1, 1, 5, 40).cuda(), opt, ang)


def test_parallel_c0_workload(rt):
    """The reference's own CPU case: parallel beam, 256x256 image of 10 seeded
    ellipses, pixel 1.0 mm, 180 views evenly over [0, pi), 367 bins of
    1.0 mm, batch 1; forward projection and backprojection (and ProjTrans)
    against the FP64 oracle and, where oracle/_ref is built, the reference's
    own code."""
    g = O.Geom2D(256, 256, 367, 1.0, 1.0, 1.0, 1.0, 1.0)
    ang = (np.arange(180, dtype=np.float32) * np.float32(np.pi / 180)).astype(np.float32)
    x = ellipses2d(256, 1, count=10, seed=0)
    sino = O.op2d("para", "proj", x, ang, g, "f64").astype(np.float32)
    for op, inp in (("proj", x), ("bproj", sino), ("proj_t", sino)):
        got = gpu_op(rt, "para", op, inp, ang, g)
        want = O.op2d("para", op, inp, ang, g, "f64")
        assert report(f"para 256^2 {op} vs FP64", got, want) <= 1e-5
        if O.ref_available():
            ref = O.op2d("para", op, inp, ang, g, "ref")
            floor = O.rel_l2(ref, want)
            r = report(f"para 256^2 {op} vs reference (its floor {floor:.2e})", got, ref)
            assert r <= max(1e-5, 2 * floor)
