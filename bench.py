"""Benchmark: batched autograd FP + adjoint of the fan-arc DD projector.

Workload (one "step"): 16 seeded 512x512 ellipse phantoms (0.7 mm pixels,
attenuation scaled to 0-0.04 /mm), equiangular fan beam, 720 views over 2*pi,
736 detector bins of 1.2858 mm at SDD 1085.6 mm, SID 595 mm; y = A x_true with
Poisson noise at 1e5 incident photons; one step evaluates the gradient of
||A x - y||^2 at x = x_true + noise through autograd: forward projection A x
for the whole batch, residual, and the matched adjoint A^T in backward.

* batched path: the step runs on one radontorch.Plan2d (geometry built once,
  the batch shares every geometry evaluation) and is replayed from a
  captured HIP graph;
* plain path: the step issues one radontorch op call per slice and per
  direction (16 forward + 16 adjoint calls), eagerly.

Prints one JSON line (metric/value/unit/... plus roofline and cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP+BP detector samples/s & %HBM roofline, 512² fan batch16; batched vs plain"
N, D, V, B = 512, 736, 720, 16
PIX, DET, SID, SDD = 0.7, 1.2858, 595.0, 1085.6
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 oracle check")
    ap.add_argument("--no-c2", action="store_true", help="skip the uneven-batch measurement")
    ap.add_argument("--no-3d", action="store_true",
                    help="skip the cone-flat and helical legs of the default run")
    ap.add_argument("--no-c0", action="store_true", help="skip the parallel-beam CPU-case leg")
    ap.add_argument("--no-fidelity", action="store_true",
                    help="skip the data-consistency step leg of the default run")
    ap.add_argument("--no-dc-loop", action="store_true",
                    help="skip the helical workload's 50-step descent loop (profiling runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner sweep (automatic tile, default kernel variant); "
                         "for profiling runs")
    ap.add_argument("--variant", type=int, default=None,
                    help="uneven workload: run this kernel variant (DD_ADJ_* | DD_FWD_PREFIX) "
                         "at one workgroup per item instead of tuning (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 512 MiB device copy (reference kernel for HBM counters)")
    ap.add_argument("--counters-from", metavar="DIR",
                    help="summarize rocprofv3 --pmc outputs under DIR into "
                         "profiles/kernel_counters.json and exit (no GPU needed)")
    ap.add_argument("--workload", default="fanarc",
                    choices=["fanarc", "uneven", "cone_flat", "helical", "fidelity"],
                    help="fanarc: 512^2 fan-arc batch 16 (headline); cone_flat: 256^3 cone-flat "
                         "batch 2; helical: 512x512x64 helical batch 4; fidelity: PrideDiff's "
                         "data-consistency step x - w FBP(A x - y) on the fan-arc batch; "
                         "uneven: 64 fan-arc instances of uneven size in one launch per op")
    return ap.parse_args()


def workload(torch, dev):
    from pridediff_amd.deeprecon.torch.projector import projector2d
    from pridediff_amd.phantom import ellipses2d

    P = projector2d(N, D, PIX, DET, SID, SDD, scan_type="arc").to(dev)
    ang_np = (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V)).astype(np.float32)
    ang = torch.from_numpy(ang_np).to(dev)
    x_true = torch.from_numpy(ellipses2d(N, B, seed=2024, scale=0.04))[:, None].to(dev)
    gen = torch.Generator(device=dev).manual_seed(7)
    with torch.no_grad():
        line = P.projection(x_true, ang)
        counts = torch.poisson(1e5 * torch.exp(-line), generator=gen)
        y = -torch.log(torch.clamp(counts, min=1.0) / 1e5)
        x0 = (x_true + 0.002 * torch.randn(x_true.shape, device=dev, generator=gen)).contiguous()
    return P, ang, ang_np, x_true, x0, y.contiguous()


COUNTERS = os.path.join(ROOT, "profiles", "kernel_counters.json")


def _counters():
    try:
        with open(COUNTERS) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def roofline_for(key, kernel, ms, b_alg, step_ms=None, n_ops=2):
    """Roofline of one gather kernel.

    The gathers are bound by the per-CU vector-memory path (texture
    addresser -> L1, 64 B per clock per CU), not by HBM: every (pixel, view)
    value passes through it, while the caches keep most re-reads off HBM.
    ``achieved`` = the bytes the kernel moves through that path per launch
    (rocprofv3 counters of the same launch, committed beside this script) /
    the launch time measured here with HIP events; ``peak`` = CUs x 64 B/clk
    x the 2.4 GHz peak engine clock.  HBM is reported as the real fraction of
    8 TB/s (FETCH_SIZE + WRITE_SIZE per launch / time), and ``effective_GBps``
    is the algorithmic view-pass bytes / time (it exceeds the HBM peak
    because caches serve the per-view re-reads).
    """
    c = _counters()
    k = c.get("kernels", {}).get(key, {})
    t = ms * 1e-3
    out = {"bound": "vmem", "achieved": None, "peak": None, "unit": "GB/s", "frac": None,
           "traffic": k.get("hbm_bytes"), "kernel": kernel,
           "effective_GBps": b_alg / t / 1e9, "effective_frac_of_hbm_peak": b_alg / t / HBM_PEAK}
    if step_ms:
        out["step_effective_frac_of_hbm_peak"] = n_ops * b_alg / (step_ms * 1e-3) / HBM_PEAK
    if k.get("vmem_bytes") and c.get("clock_hz") and c.get("cus"):
        peak = c["cus"] * c.get("l1_bytes_per_clk", 64) * c["clock_hz"]
        out.update(achieved=k["vmem_bytes"] / t / 1e9, peak=peak / 1e9,
                   frac=k["vmem_bytes"] / t / peak)
    if k.get("hbm_bytes"):
        out["hbm"] = {"achieved": k["hbm_bytes"] / t / 1e9, "peak": HBM_PEAK / 1e9,
                      "frac": k["hbm_bytes"] / t / HBM_PEAK}
    if k.get("ta_busy_frac") is not None:
        # busy fractions of the two halves of the vector-memory path in the
        # counter runs: the addresser (TA) and the data return (TD)
        out["busy_in_counter_run"] = {"ta": k["ta_busy_frac"], "td": k.get("td_busy_frac")}
        if k.get("lds_busy_frac") is not None:
            # kernels that read staged data from LDS: its indexed-access
            # busy share and the part of that stalled by bank conflicts
            out["busy_in_counter_run"].update(lds=k["lds_busy_frac"],
                                              lds_bank_conflict=k["lds_bank_conflict_frac"])
        if k.get("valu_busy") is not None:
            # rocprofv3's VALUBusy and VALUUtilization (share of lanes active)
            out["busy_in_counter_run"].update(valu=k["valu_busy"],
                                              valu_lanes_active=k["valu_lanes_active"])
    if k:
        out["counts_from"] = ("rocprofv3 --pmc runs of this workload, "
                              "profiles/kernel_counters.md; times measured in this run")
    return out


# Counter groups per workload: key -> kernels whose dispatches make up one
# HIP-event-timed launch of that key (a forward is gather + chunk reduction).
COUNTER_KEYS = {
    "fanarc": {"fanarc_b16:k_pixels<16,0>": ["k_pixels<16, 0, "],
               "fanarc_b16:k_bins_split<16,0>": ["k_bins_split<16, 0>", "k_bins_reduce<0>"],
               "fanarc_plain:k_pixels<1,0>": ["k_pixels<1, 0, "],
               "fanarc_plain:k_bins_split<1,0>": ["k_bins_split<1, 0>", "k_bins_reduce<0>"],
               "fanarc_plain:k_pixels_prefix<0>": ["k_prefix_sino<true>", "k_prefix_ranges",
                                                   "k_pixels_prefix<0>", "k_sum_chunks"],
               "fanarc_plain:k_bins_prefix<0>": ["k_prefix_lines", "k_bins_prefix<0>",
                                                 "k_bins_reduce<0>"]},
    "uneven": {"uneven64:k_batch_quads_window": ["k_batch_quads<0, 8>"],
               "uneven64:k_batch_quads": ["k_batch_quads<0, 0>"],
               "uneven64:k_batch_quads_prefix": ["k_batch_prefix_sino<true>",
                                                 "k_batch_quads_prefix<0>"],
               "uneven64:k_batch_bins": ["k_batch_bins<0, false>", "k_batch_reduce<0>"],
               "uneven64:k_batch_bins_prefix": ["k_batch_prefix", "k_batch_bins<0, true>",
                                                "k_batch_reduce<0>"]},
    "cone_flat": {"cone_flat:k_voxels3d": ["k_voxel_ranges", "k_voxels3d<2, 0, true, true>"],
                  "cone_flat:k_bins3d": ["k_bins3d<2, 0, false>"]},
    "helical": {"helical:k_voxels3d": ["k_voxels3d<4, 0, false, false>"],
                "helical:k_bins3d": ["k_bins3d<4, 0, true>"]},
}


def summarize_counters(base, out_path):
    """Write the per-kernel counter file ``roofline_for`` reads.

    ``base/<workload>_<pass>/run_counter_collection.csv`` are rocprofv3
    ``--pmc`` outputs of ``bench.py --workload <workload>`` for the passes
    A (TA/GRBM/SQ), B (TCP/TD), FETCH (FETCH_SIZE) and WRITE (WRITE_SIZE);
    the fanarc FETCH/WRITE passes add ``--copy-probe`` (a 512 MiB device
    copy), whose counters calibrate FETCH_SIZE and WRITE_SIZE to bytes.
    An optional pass LDS (SQ_LDS_IDX_ACTIVE, SQ_LDS_BANK_CONFLICT,
    SQ_INSTS_LDS, GRBM_GUI_ACTIVE) adds the share of cycles the LDS serves
    indexed reads and writes, and the share of those stalled by bank
    conflicts.  An optional pass VALU (SQ_ACTIVE_INST_VALU,
    SQ_THREAD_CYCLES_VALU, SQ_INSTS_VALU, SQ_INSTS_VALU_ADD_F64,
    SQ_INSTS_VALU_FMA_F64, GRBM_GUI_ACTIVE) adds rocprofv3's VALUBusy and
    VALUUtilization (the share of a wave's lanes active in its VALU
    instructions) and the FP64 share of the VALU instructions.  Values are
    means over the dispatches of each kernel.
    """
    import collections
    import csv

    def load(path):
        vals = collections.defaultdict(lambda: collections.defaultdict(list))
        dur = collections.defaultdict(dict)
        with open(path) as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"].split("(dd::")[0].split("(char")[0]
                name = name.replace("void ", "").replace("dd::(anonymous namespace)::", "")
                vals[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
                dur[name][r["Dispatch_Id"]] = (int(r["End_Timestamp"])
                                               - int(r["Start_Timestamp"])) * 1e-9
        return vals, dur

    def find(vals, pat):
        ks = [k for k in vals if k == pat] or [k for k in vals if k.startswith(pat)]
        if len(ks) > 1:
            raise SystemExit(f"counter summary: {pat!r} matches {ks}")
        return ks[0] if ks else None

    def mean(vals, k, c):
        return statistics.mean(vals[k][c])

    cus, clk = 256, 2.4e9
    out = {"cus": cus, "l1_bytes_per_clk": 64, "clock_hz": clk, "kernels": {}}
    F, _ = load(os.path.join(base, "fanarc_FETCH", "run_counter_collection.csv"))
    W, _ = load(os.path.join(base, "fanarc_WRITE", "run_counter_collection.csv"))
    cf, cw = find(F, "__amd_rocclr_copyBuffer"), find(W, "__amd_rocclr_copyBuffer")
    copy_kib = 512 * 1024
    fmax, wmax = max(F[cf]["FETCH_SIZE"]), max(W[cw]["WRITE_SIZE"])
    cal = {"copy_kib": copy_kib, "fetch_kib": fmax, "write_kib": wmax,
           "fetch_scale": copy_kib / fmax, "write_scale": copy_kib / wmax}
    for wl, keys in COUNTER_KEYS.items():
        runs = {p: load(os.path.join(base, f"{wl}_{p}", "run_counter_collection.csv"))
                for p in ("A", "B", "FETCH", "WRITE")}
        lds_csv = os.path.join(base, f"{wl}_LDS", "run_counter_collection.csv")
        Lv = load(lds_csv)[0] if os.path.exists(lds_csv) else None
        valu_csv = os.path.join(base, f"{wl}_VALU", "run_counter_collection.csv")
        Vv = load(valu_csv)[0] if os.path.exists(valu_csv) else None
        for key, pats in keys.items():
            if any(find(runs[p][0], pat) is None for pat in pats for p in runs):
                continue  # a kernel this run's configuration does not launch
            e = {"kernels": [], "pmc_ms": 0.0, "vmem_instructions": 0.0, "waves": 0.0,
                 "valu_instructions": 0.0, "vmem_bytes": 0.0, "l1_sector_bytes": 0.0,
                 "fetch_kib_raw": 0.0, "write_kib_raw": 0.0, "_cyc": 0.0, "_ta": 0.0, "_td": 0.0,
                 "_lcyc": 0.0, "_lidx": 0.0, "_lbank": 0.0, "_linst": 0.0,
                 "_vcyc": 0.0, "_vact": 0.0, "_vthr": 0.0, "_vins": 0.0, "_vf64": 0.0}
            for pat in pats:
                (A, tA), (Bv, _), (Fv, _), (Wv, _) = (runs[p] for p in ("A", "B", "FETCH", "WRITE"))
                ka, kb, kf, kw = find(A, pat), find(Bv, pat), find(Fv, pat), find(Wv, pat)
                e["kernels"].append(ka)
                e["pmc_ms"] += statistics.mean(tA[ka].values()) * 1e3
                e["_cyc"] += mean(A, ka, "GRBM_GUI_ACTIVE") / 8   # summed over 8 XCDs
                e["_ta"] += mean(A, ka, "TA_TA_BUSY_sum") / cus
                e["_td"] += mean(Bv, kb, "TD_TD_BUSY_sum") / cus
                e["vmem_instructions"] += mean(A, ka, "SQ_INSTS_VMEM_RD")
                e["valu_instructions"] += mean(A, ka, "SQ_INSTS_VALU")
                e["waves"] += mean(A, ka, "SQ_WAVES")
                e["vmem_bytes"] += mean(A, ka, "TA_TA_BUSY_sum") * out["l1_bytes_per_clk"]
                e["l1_sector_bytes"] += mean(Bv, kb, "TCP_TOTAL_CACHE_ACCESSES_sum") * 32
                e["fetch_kib_raw"] += mean(Fv, kf, "FETCH_SIZE")
                e["write_kib_raw"] += mean(Wv, kw, "WRITE_SIZE")
                kl = find(Lv, pat) if Lv is not None else None
                if kl is not None:
                    e["_lcyc"] += mean(Lv, kl, "GRBM_GUI_ACTIVE") / 8
                    e["_lidx"] += mean(Lv, kl, "SQ_LDS_IDX_ACTIVE")
                    e["_lbank"] += mean(Lv, kl, "SQ_LDS_BANK_CONFLICT")
                    e["_linst"] += mean(Lv, kl, "SQ_INSTS_LDS")
                kv = find(Vv, pat) if Vv is not None else None
                if kv is not None:
                    e["_vcyc"] += mean(Vv, kv, "GRBM_GUI_ACTIVE") / 8
                    e["_vact"] += mean(Vv, kv, "SQ_ACTIVE_INST_VALU")
                    e["_vthr"] += mean(Vv, kv, "SQ_THREAD_CYCLES_VALU")
                    e["_vins"] += mean(Vv, kv, "SQ_INSTS_VALU")
                    e["_vf64"] += (mean(Vv, kv, "SQ_INSTS_VALU_ADD_F64")
                                   + mean(Vv, kv, "SQ_INSTS_VALU_FMA_F64"))
            lcyc, lidx, lbank, linst = (e.pop(k) for k in ("_lcyc", "_lidx", "_lbank", "_linst"))
            vcyc, vact, vthr, vins, vf64 = (e.pop(k) for k in
                                            ("_vcyc", "_vact", "_vthr", "_vins", "_vf64"))
            if vcyc > 0 and vact > 0:
                e["valu_busy"] = vact / cus / vcyc
                e["valu_lanes_active"] = vthr / (vact * 64)
                e["valu_f64_share"] = vf64 / vins
            if lcyc > 0 and linst > 0:
                e["lds_busy_frac"] = lidx / cus / lcyc
                e["lds_bank_conflict_frac"] = lbank / lidx
                e["lds_instructions"] = linst
            cyc, ta, td = e.pop("_cyc"), e.pop("_ta"), e.pop("_td")
            e["measured_clock_hz"] = cyc / (e["pmc_ms"] * 1e-3)
            e["ta_busy_frac"], e["td_busy_frac"] = ta / cyc, td / cyc
            e["hbm_bytes"] = (e["fetch_kib_raw"] * cal["fetch_scale"]
                              + e["write_kib_raw"] * cal["write_scale"]) * 1024
            out["kernels"][key] = e
    out["calibration"] = cal
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    for k, v in out["kernels"].items():
        t = v["pmc_ms"] * 1e-3
        print(f"{k:34s} {v['pmc_ms']:8.3f} ms  clk {v['measured_clock_hz'] / 1e9:.2f} GHz  "
              f"TA {v['ta_busy_frac']:.2f}  vmem {v['vmem_bytes'] / t / (cus * 64 * clk):.2f}  "
              f"HBM {v['hbm_bytes'] / 1e9:7.3f} GB = {v['hbm_bytes'] / t / HBM_PEAK:.3f} of peak"
              + (f"  LDS {v['lds_busy_frac']:.2f} (bank conflicts {v['lds_bank_conflict_frac']:.2f})"
                 if "lds_busy_frac" in v else ""))
    print(json.dumps(cal))


def roofline_c1(kernel, ms, b_alg, step_s):
    return roofline_for("fanarc_b16:" + kernel, kernel, ms, b_alg, step_s * 1e3)


def timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def timed_in_turn(torch, fns, steps, warmup, rounds=3):
    """Median over ``rounds`` of each function's time per call, the functions
    timed one after another in every round, so that a clock still ramping up
    during the first measurements does not favour the ones timed last."""
    for fn in fns:
        for _ in range(warmup):
            fn()
    res = [[] for _ in fns]
    for _ in range(rounds):
        for i, fn in enumerate(fns):
            res[i].append(timed(torch, fn, steps, 0))
    return [statistics.median(r) for r in res]


def uneven_c2(torch, dev, steps, warmup, check, tune_grid=True, cpu=False, variant=None):
    """64 fan-arc instances, N in {256,384,512}, V in [96,720], D = ceil(1.44 N):
    one launch per op and phase inside one HIP graph vs 64 plain calls."""
    from pridediff_amd.autotune import tune_uneven
    from pridediff_amd.batched import GraphStep, UnevenBatch2d, data_consistency_grad
    from pridediff_amd.deeprecon.torch.projector import projector2d
    from pridediff_amd.phantom import ellipses2d

    rng = np.random.default_rng(2025)
    projs, angs, angs_np, imgs = [], [], [], []
    for i in range(64):
        n = int(rng.choice([256, 384, 512]))
        v = int(rng.integers(96, 721))
        d = int(math.ceil(1.44 * n))
        projs.append(projector2d(n, d, PIX, DET, SID, SDD, scan_type="arc").to(dev))
        a = (np.arange(v, dtype=np.float32) * np.float32(2 * np.pi / v)).astype(np.float32)
        angs_np.append(a)
        angs.append(torch.from_numpy(a).to(dev))
        imgs.append(torch.from_numpy(ellipses2d(n, 1, seed=3000 + i, scale=0.04))[0].to(dev))
    ub = UnevenBatch2d(projs, angs)
    tune = {}
    if variant is not None:
        ub.plan.set_variant(variant)
        tune = {"grid": 0, "variant": variant}
    elif tune_grid:
        tune_uneven(ub, report=tune)
    x = ub.pack_images(imgs)
    with torch.no_grad():
        y = (ub.projection(x) * 1.001).contiguous()
    step = GraphStep(lambda a, b: data_consistency_grad(ub.projection, a, b),
                     x.clone().requires_grad_(True), y)
    t_b = timed(torch, lambda: step.replay(x, y), steps, warmup)
    ys = ub.unpack_sinos(y)

    def plain():
        for P, a, im, yy in zip(projs, angs, imgs, ys):
            xi = im[None, None].detach().clone().requires_grad_(True)
            data_consistency_grad(lambda t: P.projection(t, a), xi, yy[None, None])
    t_p = timed(torch, plain, 2, 1)
    from pridediff_amd import load_radontorch
    rt = load_radontorch()
    rt.profile_enable(True)
    rt.profile_read(0), rt.profile_read(1)
    for _ in range(3):
        xg = x.detach().clone().requires_grad_(True)
        data_consistency_grad(ub.projection, xg, y)
    torch.cuda.synchronize()
    (fp_ms, fp_n), (bp_ms, bp_n) = rt.profile_read(0), rt.profile_read(1)
    rt.profile_enable(False)
    shapes = ub.shapes
    b_alg = sum(8.0 * v * (h * w + d) for h, w, v, d in shapes)
    out = {"instances": 64, "batched_ms": t_b * 1e3, "plain_ms": t_p * 1e3,
           "detector_samples": sum(v * d for _, _, v, d in shapes),
           "speedup": t_p / t_b, "step_effective_frac_of_hbm_peak": b_alg / t_b / HBM_PEAK,
           "algorithmic_GB": b_alg / 1e9, "grid": tune.get("grid"),
           "variant": tune.get("variant"),
           "forward_ms": fp_ms / max(fp_n, 1), "adjoint_ms": bp_ms / max(bp_n, 1)}
    var = tune.get("variant", rt.ADJ_QUAD)
    out["forward_kernel"] = "running sums" if var & rt.FWD_PREFIX else "pixel walk"
    adj_name = {rt.ADJ_QUAD: "k_batch_quads", rt.ADJ_BLOCK_WINDOW: "k_batch_quads_window",
                rt.ADJ_BLOCK_PREFIX: "k_batch_quads_prefix",
                rt.ADJ_TILE16: "k_batch_tile"}.get(var & ~rt.FWD_PREFIX, "k_batch_pixels")
    fwd_name = "k_batch_bins_prefix" if var & rt.FWD_PREFIX else "k_batch_bins"
    dom = ((fwd_name, out["forward_ms"]) if out["forward_ms"] >= out["adjoint_ms"]
           else (adj_name, out["adjoint_ms"]))
    out["roofline"] = roofline_for("uneven64:" + dom[0], dom[0], dom[1], b_alg / 2)
    if check:
        from oracle import oracle as O
        j = min(range(64), key=lambda i: shapes[i][0] * shapes[i][2])
        h, w, v, d = shapes[j]
        g = O.Geom2D(w, h, d, PIX, PIX, DET / SDD, SID, SDD)
        with torch.no_grad():
            got = ub.unpack_sinos(ub.projection(x))[j].cpu().numpy()[None]
        want = O.op2d("arc", "proj", imgs[j].cpu().numpy()[None], angs_np[j], g, "f64")
        out["smallest_instance_fp_rel_l2_vs_fp64"] = O.rel_l2(got, want)
    if cpu:
        # the lightest, a middle and the heaviest instance by V * N^2
        order = sorted(range(64), key=lambda i: shapes[i][2] * shapes[i][0] * shapes[i][1])
        pick = [order[0], order[32], order[-1]]
        out["cpu_sample"] = c2_cpu_sample([(i, shapes[i]) for i in pick])
        per_inst_gpu = out["batched_ms"] * 1e-3 / 64
        out["cpu_sample"]["note"] = (f"batched GPU step {per_inst_gpu * 1e3:.3f} ms per instance "
                                     "on average over the 64")
    return out


def bench3d(torch, kind, steps, warmup, check, dc_loop=True):
    """3D autograd FP + adjoint step, replayed from a captured HIP graph.

    cone_flat: 256^3 volume of 20 ellipsoids (1 mm voxels), flat detector
    512 x 384 of 1 mm at SDD 1500, SID 1000, 360 views, batch 2.
    helical: 512 x 512 x 64 volume (0.7, 0.7, 1.0 mm), curved detector 736 x 64
    (1.2858 mm x 1.82454 mm at SDD 1085.6, SID 595), 3 turns of 1152 views,
    source z linear over [-96, 96] mm, batch 4.
    """
    from pridediff_amd import load_radontorch
    from pridediff_amd.batched import GraphStep, data_consistency_grad
    from pridediff_amd.deeprecon.torch.projector import projector3d, projectorsp
    from pridediff_amd.phantom import ellipsoids3d

    rt = load_radontorch()
    dev = torch.device("cuda:0")
    if kind == "cone_flat":
        W = H = Dz = 256
        C, R, V, Bn = 512, 384, 360, 2
        P = projector3d((W, H, Dz), (C, R), 1.0, 1.0, 1000.0, 1500.0, scan_type="flat").to(dev)
        ang_np = (np.arange(V) * (2 * np.pi / V)).astype(np.float32)
        spos_np = None
        vox = (1.0, 1.0, 1.0)
        z_src = np.zeros(V)
        half_rows = R * 1.0 / 2 * 1000.0 / 1500.0  # row extent at the isocentre
        sid, r_fov = 1000.0, 0.5 * math.hypot(W, H)
        desc = "cone-flat 256^3, 360 views, 512x384 flat detector, batch 2, autograd FP + adjoint"
    else:
        W, H, Dz = 512, 512, 64
        C, R, Bn = 736, 64, 4
        V = 3 * 1152
        P = projectorsp((W, H, Dz), (C, R), (0.7, 0.7, 1.0), (1.2858, 1.82454), 595.0, 1085.6,
                        scan_type="arc").to(dev)
        ang_np = (np.arange(V) * (2 * np.pi / 1152)).astype(np.float32)
        spos_np = np.zeros((4, V), dtype=np.float32)
        spos_np[0] = np.linspace(-96.0, 96.0, V, dtype=np.float32)
        vox = (0.7, 0.7, 1.0)
        z_src = spos_np[0].astype(np.float64)
        half_rows = R * 1.82454 / 2 * 595.0 / 1085.6
        sid, r_fov = 595.0, 0.5 * math.hypot(W * 0.7, H * 0.7)
        desc = ("helical 512x512x64, 3456 views (3 turns), 736x64 curved detector, batch 4, "
                "autograd FP + adjoint")
    ang = torch.from_numpy(ang_np).to(dev)
    spos = None if spos_np is None else torch.from_numpy(spos_np).to(dev)
    x_true = torch.from_numpy(ellipsoids3d(W, H, Dz, Bn, count=20, seed=77))[:, None].to(dev)
    x_true = (x_true * 0.02).contiguous()
    fwd = (lambda t: P.projection(t, ang)) if spos is None else (lambda t: P.projection(t, ang, spos))
    with torch.no_grad():
        y = (fwd(x_true) * 1.001).contiguous()
    x0 = (x_true + 0.0005).contiguous()
    step = GraphStep(lambda a, b: data_consistency_grad(fwd, a, b), x0.clone().requires_grad_(True), y)
    t = timed(torch, lambda: step.replay(x0, y), steps, warmup)

    loop_out = None
    if kind == "helical" and dc_loop:
        # 50 descent steps on ||A x - y||^2, each one replay of a captured
        # graph holding A x, the residual, A^T through autograd and the update;
        # fixed step from an exact line search at the start point.
        with torch.no_grad():
            xs = torch.zeros_like(x_true)
        xg = xs.clone().requires_grad_(True)
        g0 = data_consistency_grad(fwd, xg, y).detach()
        with torch.no_grad():
            ag = fwd(g0)
            alpha = float((g0 * g0).sum() / (2.0 * (ag * ag).sum()))

        def update(a, c):
            grad = data_consistency_grad(fwd, a, c)
            return (a - alpha * grad).detach()
        loop = GraphStep(update, xs.clone().requires_grad_(True), y)
        losses = []
        with torch.no_grad():
            losses.append(float(((fwd(xs) - y) ** 2).sum()))
        xk = xs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(50):
            xk = loop.replay(xk, y)
        torch.cuda.synchronize()
        t_loop = time.perf_counter() - t0
        with torch.no_grad():
            losses.append(float(((fwd(xk) - y) ** 2).sum()))
        loop_out = {"steps": 50, "seconds": t_loop, "ms_per_step": t_loop / 50 * 1e3,
                   "loss_start": losses[0], "loss_end": losses[1], "step_size": alpha,
                   "start": "zero volume"}

    rt.profile_enable(True)
    rt.profile_read(0), rt.profile_read(1)
    for _ in range(2):
        xg = x0.detach().clone().requires_grad_(True)
        data_consistency_grad(fwd, xg, y)
    torch.cuda.synchronize()
    fp_ms, fp_n = rt.profile_read(0)
    bp_ms, bp_n = rt.profile_read(1)
    rt.profile_enable(False)
    fp_ms, bp_ms = fp_ms / max(fp_n, 1), bp_ms / max(bp_n, 1)

    # in-cone voxels per view: slices that meet the cone's z range over the
    # field of view (source distance sid +- r_fov)
    zmag = (sid + r_fov) / sid
    z_edges = (np.arange(Dz + 1) - Dz / 2.0) * vox[2]
    in_cone = 0
    for zs in z_src:
        lo, hi = zs - half_rows * zmag, zs + half_rows * zmag
        sl = np.count_nonzero((z_edges[1:] > lo) & (z_edges[:-1] < hi))
        in_cone += sl * W * H
    b_alg = 4.0 * (in_cone + V * R * C) * Bn  # bytes per operator launch
    dom_ms, dom = (fp_ms, "k_bins3d") if fp_ms >= bp_ms else (bp_ms, "k_voxels3d")
    samples = Bn * V * R * C
    out = {
        "metric": "FP+BP detector samples/s & %HBM roofline (3D)",
        "value": samples / t, "unit": "detector samples/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms, y = 1.001 A x_true)",
        "config": {"workload": desc, "volume": [W, H, Dz], "views": V, "detector": [C, R],
                   "batch": Bn},
        "roofline": dict(roofline_for(f"{kind}:{dom}", dom, dom_ms, b_alg, t * 1e3),
                         in_cone_voxel_views=in_cone),
        "kernels": {"k_bins3d (A, volume->sino)": {"ms": fp_ms, "effective_GBps": b_alg / fp_ms / 1e6},
                    "k_voxels3d (A^T, sino->volume)": {"ms": bp_ms,
                                                       "effective_GBps": b_alg / bp_ms / 1e6}},
        "cpu_baseline": None,
    }
    if loop_out is not None:
        out["dc_loop"] = loop_out
    if check:
        from oracle import oracle as O
        if kind == "cone_flat":
            g = O.Geom3D(W, H, Dz, C, R, 1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 1500.0)
            scan = "cone_flat"
        else:
            g = O.Geom3D(W, H, Dz, C, R, 0.7, 0.7, 1.0, 1.2858 / 1085.6, 1.82454, 595.0, 1085.6)
            scan = "spiral"
        # full-size launches, checked on sampled views: the forward's views
        # against the oracle on those views, the adjoint of a sinogram that
        # is zero outside them against the oracle's adjoint of those views
        sel = np.linspace(0, V - 1, 4).round().astype(np.int64)
        a_sub = ang_np[sel]
        s_sub = None if spos_np is None else np.ascontiguousarray(spos_np[:, sel])
        sp_args = () if spos is None else (spos,)
        with torch.no_grad():
            got = P.projection(x_true[:1], ang, *sp_args)[:, 0, sel].cpu().numpy()
        want = O.op3d(scan, "proj", x_true[:1, 0].cpu().numpy(), a_sub, g, s_sub)
        # the batch of the workload (so the launch uses its instance tile);
        # instance 0 is compared
        ysp = torch.zeros(Bn, 1, V, R, C, device=dev)
        ysp[:, :, torch.from_numpy(sel).to(dev)] = torch.rand(
            Bn, 1, len(sel), R, C, device=dev, generator=torch.Generator(device=dev).manual_seed(11))
        want_t = O.op3d(scan, "proj_t", ysp[:1, 0, torch.from_numpy(sel).to(dev)].cpu().numpy(),
                        a_sub, g, s_sub)
        acc = {"views_checked": sel.tolist(),
               "fp_rel_l2_vs_fp64": O.rel_l2(got, want), "fp_max_abs_vs_fp64": O.max_abs(got, want)}
        # the adjoint with compensated running sums (the default) and with
        # running sums rounded to FP32, accuracy and kernel time of each
        for comp in (True, False):
            rt.set_compensated3d(1 if comp else 0)
            with torch.no_grad():
                got_t = P.projection_t(ysp, ang, *sp_args)[:1, 0].cpu().numpy()
                rt.profile_enable(True)
                rt.profile_read(1)
                for _ in range(2):
                    P.projection_t(y, ang, *sp_args)
                torch.cuda.synchronize()
                ms, n = rt.profile_read(1)
                rt.profile_enable(False)
            tag = "compensated" if comp else "rounded"
            acc[f"adjoint_{tag}"] = {"rel_l2_vs_fp64": O.rel_l2(got_t, want_t),
                                     "max_abs_vs_fp64": O.max_abs(got_t, want_t),
                                     "k_voxels3d_ms": ms / max(n, 1)}
        rt.set_compensated3d(2)
        acc["default_mode"] = ("compensated" if Bn <= 2 else "rounded") + \
            " (automatic: compensated at instance tiles 1-2)"
        acc["lds_staging"] = ("on" if Bn <= 2 else "on for compensated sums only") + \
            " (automatic: compensated sums or instance tiles 1-2; same bits either way)"
        out["accuracy"] = acc
    return out


def bench_fidelity(torch, steps, warmup, ctx=None):
    """PrideDiff's data-consistency step x - w * FBP(A x - y)
    (model/pridediff.py:56-68) on the 512^2 fan-arc batch of 16: the fused
    planned step (residual and update in the projector kernels' writes) in
    one HIP graph, vs the unfused planned step in a graph, vs the drop-in
    module called on the batch (GEMM filter), vs the same with the
    reference's conv2d filter.  ``ctx`` = (projector, angles, x0, y, tuned
    batched projector) of the headline run: then only the fused, unfused and
    module steps are timed (the compact leg of the default run)."""
    from pridediff_amd.autotune import tune_uniform
    from pridediff_amd.batched import GraphStep
    from pridediff_amd.fbp import PlannedFBP2d, fidelity_step, fidelity_step_fused

    dev = torch.device("cuda:0")
    compact = ctx is not None
    if compact:
        P, ang, x0, y, BP = ctx
    else:
        P, ang, ang_np, x_true, x0, y = workload(torch, dev)
        BP = tune_uniform(P, ang, B)
    fbp = PlannedFBP2d(P, ang, B, plan=BP)
    w = torch.tensor(0.5, device=dev)
    with torch.no_grad():
        # fused: residual in the forward's reduction, update in the BP's write
        fused = GraphStep(lambda a, c: fidelity_step_fused(fbp, a, c, w.reshape(1)), x0, y)
        step = GraphStep(lambda a, c: fidelity_step(BP.projection, fbp, a, c, w), x0, y)

        def module():
            return x0 - w * P.filtered_backprojection(P.projection(x0, ang) - y, ang)
        t_graph, t_unfused, t_module = timed_in_turn(
            torch, (lambda: fused.replay(x0, y), lambda: step.replay(x0, y), module),
            steps, warmup)
        out_graph = fused.replay(x0, y).clone()
        out_unfused = step.replay(x0, y).clone()
        out_module = module()
    rel = lambda a, b: float((a - b).double().norm() / b.double().norm())
    if compact:
        return {"workload": "x - w FBP(A x - y) on the same batch",
                "fused_graph_ms": t_graph * 1e3, "unfused_graph_ms": t_unfused * 1e3,
                "module_call_ms": t_module * 1e3,
                "value": B * V * D / t_graph, "unit": "detector samples/s",
                "agreement": {"fused_vs_unfused_rel_l2": rel(out_graph, out_unfused),
                              "graph_vs_module_rel_l2": rel(out_graph, out_module)}}
    with torch.no_grad():
        def conv_fbp(sino):
            f = torch.nn.functional.conv2d(sino * P.weight, P.filter, padding=(0, D - 1))
            r = P.backprojection_w(f, ang) * (ang[-1] - ang[0]).abs() / (V - 1)
            return r / 2
        t_conv = timed(torch, lambda: x0 - w * conv_fbp(P.projection(x0, ang) - y), 3, 1)
        out_conv = x0 - w * conv_fbp(P.projection(x0, ang) - y)
        # filter stage alone: GEMM vs conv2d
        r = (P.projection(x0, ang) - y).contiguous()
        t_gemm_only = timed(torch, lambda: fbp.filter(r), 10, 2)
        t_conv_only = timed(torch, lambda: torch.nn.functional.conv2d(
            r * P.weight, P.filter, padding=(0, D - 1)), 3, 1)
    samples = B * V * D
    out = {
        "metric": "PrideDiff data-consistency steps: detector samples/s",
        "value": samples / t_graph, "unit": "detector samples/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": t_graph * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipse phantoms, Poisson noise at 1e5 photons)",
        "config": {"workload": "x - w FBP(A x - y), fan-arc 512x512, 720 views, 736 bins, batch 16",
                   "image": [N, N], "views": V, "bins": D, "batch": B},
        "unfused_graph_ms": t_unfused * 1e3,
        "module_call_ms": t_module * 1e3,
        "module_call_conv2d_filter_ms": t_conv * 1e3,
        "filter_stage_ms": {"gemm": t_gemm_only * 1e3, "conv2d": t_conv_only * 1e3},
        "agreement": {"fused_vs_unfused_rel_l2": rel(out_graph, out_unfused),
                      "graph_vs_module_rel_l2": rel(out_graph, out_module),
                      "graph_vs_conv2d_rel_l2": rel(out_graph, out_conv)},
        "cpu_baseline": None,
    }
    return out


def _ref_threads(views, mode):
    """deeprecon's CPU thread policy min(ceil(V/128), hardware_concurrency)
    (FanArcDisCPU2d.cpp:820-822); the FP32 port is single-threaded."""
    hw = os.cpu_count() or 1
    return (min(math.ceil(views / 128), hw) if mode == "ref" else 1), hw


def _cpu_time(fn, runs=3):
    fn()  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_baseline(ang_np, x_np, y_np):
    """Reference CPU code (oracle/_ref) on 2 slices: FP + ProjTrans, median of 3."""
    from oracle import oracle as O

    g = O.Geom2D(N, N, D, PIX, PIX, DET / SDD, SID, SDD)
    mode = "ref" if O.ref_available() else "f32"
    xs, ys = x_np[:2], y_np[:2]
    t_fp = _cpu_time(lambda: O.op2d("arc", "proj", xs, ang_np, g, mode))
    t_pt = _cpu_time(lambda: O.op2d("arc", "proj_t", ys, ang_np, g, mode))
    t = t_fp + t_pt
    threads, hw = _ref_threads(V, mode)
    return {
        "value": 2 * V * D / t, "unit": "detector samples/s", "cores": threads,
        "hardware_concurrency": hw,
        "kind": "reference" if mode == "ref" else "port",
        "per_slice_s": {"proj": t_fp / 2, "proj_t": t_pt / 2, "proj_plus_proj_t": t / 2},
        "sample": (f"2 slices of the 512^2 fan-arc workload (720 views x 736 bins), "
                   f"Proj and ProjTrans each the median of 3 runs ({t_fp:.2f} s + {t_pt:.2f} s); "
                   f"deeprecon's CPU thread policy min(ceil(V/128), cores) = {threads} threads "
                   f"on a {hw}-CPU host"),
    }


def parallel_c0(torch, dev, check):
    """The reference's own CPU case: parallel beam, 256x256 image of 10 seeded
    ellipses (1.0 mm pixels), 180 views over [0, pi), 367 bins of 1.0 mm,
    batch 1; Proj + ProjTrans by the reference CPU code (its thread policy:
    ceil(180/128) = 2 threads) and by one plain GPU call each."""
    from oracle import oracle as O
    from pridediff_amd import load_radontorch
    from pridediff_amd.phantom import ellipses2d

    rt = load_radontorch()
    n, d, v = 256, 367, 180
    g = O.Geom2D(n, n, d, 1.0, 1.0, 1.0, 1.0, 1.0)
    ang_np = (np.arange(v, dtype=np.float32) * np.float32(np.pi / v)).astype(np.float32)
    x_np = ellipses2d(n, 1, count=10, seed=0)
    opt = torch.tensor(g.options(), dtype=torch.float32)
    ang = torch.from_numpy(ang_np).to(dev)
    x = torch.from_numpy(x_np)[:, None].to(dev)

    def gpu():
        return rt.ProjTransParaDis2d(rt.ProjParaDis2d(x, opt, ang), opt, ang)
    t_gpu = timed(torch, gpu, 50, 5)
    mode = "ref" if O.ref_available() else "f32"
    sino_np = O.op2d("para", "proj", x_np, ang_np, g, mode)
    t_cpu = _cpu_time(lambda: (O.op2d("para", "proj", x_np, ang_np, g, mode),
                               O.op2d("para", "proj_t", sino_np, ang_np, g, mode)), runs=5)
    threads, hw = _ref_threads(v, mode)
    out = {"workload": "parallel 256x256, 180 views over [0, pi), 367 bins, batch 1, Proj + ProjTrans",
           "gpu_ms": t_gpu * 1e3, "gpu_samples_per_s": v * d / t_gpu,
           "cpu_ms": t_cpu * 1e3, "cpu_samples_per_s": v * d / t_cpu,
           "cpu_kind": "reference" if mode == "ref" else "port", "cpu_threads": threads,
           "hardware_concurrency": hw}
    if check:
        with torch.no_grad():
            fp = rt.ProjParaDis2d(x, opt, ang)[:, 0].cpu().numpy()
            y = torch.from_numpy(sino_np.astype(np.float32))[:, None].to(dev)
            pt = rt.ProjTransParaDis2d(y, opt, ang)[:, 0].cpu().numpy()
        out["accuracy"] = _accuracy_vs(O, "para", "proj", x_np, fp, ang_np, g)
        out["accuracy"].update(_accuracy_vs(O, "para", "proj_t", sino_np, pt, ang_np, g))
    return out


def _accuracy_vs(O, geom, op, inp, got, ang_np, g):
    """Distance of a GPU result from the FP64 oracle and from the reference's
    own CPU output, with the reference's own distance from FP64 (its floor)."""
    f64 = O.op2d(geom, op, inp, ang_np, g, "f64")
    out = {f"{op}_rel_l2_vs_fp64": O.rel_l2(got, f64), f"{op}_max_abs_vs_fp64": O.max_abs(got, f64)}
    if O.ref_available():
        ref = O.op2d(geom, op, inp, ang_np, g, "ref")
        out[f"{op}_rel_l2_vs_reference"] = O.rel_l2(got, ref)
        out[f"{op}_max_abs_vs_reference"] = O.max_abs(got, ref)
        out[f"{op}_reference_floor_rel_l2_vs_fp64"] = O.rel_l2(ref, f64)
    return out


def c2_cpu_sample(shapes_sel):
    """Reference CPU code on a few of the uneven instances: Proj + ProjTrans
    per instance with its thread policy; returns per-instance samples/s."""
    from oracle import oracle as O
    from pridediff_amd.phantom import ellipses2d

    mode = "ref" if O.ref_available() else "f32"
    rows = []
    for i, (h, w, v, d) in shapes_sel:
        g = O.Geom2D(w, h, d, PIX, PIX, DET / SDD, SID, SDD)
        a = (np.arange(v, dtype=np.float32) * np.float32(2 * np.pi / v)).astype(np.float32)
        x = ellipses2d(w, 1, seed=3000 + i, scale=0.04)
        y = O.op2d("arc", "proj", x, a, g, mode)
        t = _cpu_time(lambda: (O.op2d("arc", "proj", x, a, g, mode),
                               O.op2d("arc", "proj_t", y, a, g, mode)))
        threads, hw = _ref_threads(v, mode)
        rows.append({"instance": i, "image": [h, w], "views": v, "bins": d, "seconds": t,
                     "samples_per_s": v * d / t, "threads": threads})
    tot_s = sum(r["seconds"] for r in rows)
    tot_samples = sum(r["views"] * r["bins"] for r in rows)
    return {"value": tot_samples / tot_s, "unit": "detector samples/s",
            "cores": max(r["threads"] for r in rows),
            "kind": "reference" if mode == "ref" else "port",
            "sample": (f"{len(rows)} of the 64 instances (lightest, middle, heaviest by V*N^2), "
                       "Proj + ProjTrans each, median of 3; deeprecon's thread policy per "
                       "instance"),
            "instances": rows, "hardware_concurrency": os.cpu_count() or 1}


def main():
    args = parse()
    if args.counters_from:
        summarize_counters(args.counters_from, COUNTERS)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    if args.workload == "uneven":
        dev = torch.device("cuda:0")
        c2 = uneven_c2(torch, dev, args.steps, args.warmup, not args.no_accuracy, not args.no_tune,
                       cpu=not args.no_cpu, variant=args.variant)
        samples = c2.pop("detector_samples")
        print(json.dumps({
            "metric": "FP+BP detector samples/s & %HBM roofline, 64 uneven fan-arc instances",
            "value": samples / (c2["batched_ms"] * 1e-3), "unit": "detector samples/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": c2["batched_ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded ellipse phantoms)",
            "config": {"workload": "64 fan-arc instances, N in {256,384,512}, V in [96,720], "
                                   "D = ceil(1.44 N), one launch per op"},
            "roofline": c2.pop("roofline"),
            "uneven_c2": c2, "cpu_baseline": c2.pop("cpu_sample", None)}), flush=True)
        return
    if args.workload == "fidelity":
        print(json.dumps(bench_fidelity(torch, args.steps, args.warmup)), flush=True)
        return
    if args.workload != "fanarc":
        print(json.dumps(bench3d(torch, args.workload, args.steps, args.warmup,
                                 not args.no_accuracy, dc_loop=not args.no_dc_loop)), flush=True)
        return
    from pridediff_amd import load_radontorch
    from pridediff_amd.autotune import tune_uniform
    from pridediff_amd.batched import GraphStep, data_consistency_grad

    rt = load_radontorch()
    dev = torch.device("cuda:0")
    P, ang, ang_np, x_true, x0, y = workload(torch, dev)
    tune = {}
    if args.no_tune:
        from pridediff_amd.batched import BatchedProjector2d
        BP = BatchedProjector2d(P, ang, B)
        tune = {"skipped": True, "tile": BP.tile, "variant": BP.variant}
    else:
        BP = tune_uniform(P, ang, B, report=tune)

    # ---- batched path: graph-captured autograd step ----------------------
    step = GraphStep(lambda x, yy: data_consistency_grad(BP.projection, x, yy),
                     x0.clone().requires_grad_(True), y)
    grad_batched = step.replay(x0, y).clone()
    t_batched = timed(torch, lambda: step.replay(x0, y), args.steps, args.warmup)

    # ---- plain path: one op call per slice and direction, eager ----------
    def plain_step():
        grads = []
        for i in range(B):
            xi = x0[i:i + 1].detach().clone().requires_grad_(True)
            grads.append(data_consistency_grad(lambda t: P.projection(t, ang), xi, y[i:i + 1]))
        return grads

    grad_plain = torch.cat(plain_step())
    t_plain = timed(torch, plain_step, args.plain_steps, 1)
    # the same calls through single-call running sums in both directions
    # (opt-in variant: the default keeps the walks, bit-identical to the
    # batched path)
    rt.set_adjoint_variant(rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX)
    try:
        grad_plain_rs = torch.cat(plain_step())
        t_plain_rs = timed(torch, plain_step, args.plain_steps, 1)
    finally:
        rt.set_adjoint_variant(-1)

    # ---- per-kernel timing (HIP events on the launch stream) -------------
    rt.profile_enable(True)
    rt.profile_read(0), rt.profile_read(1)
    for _ in range(5):
        xg = x0.detach().clone().requires_grad_(True)
        data_consistency_grad(BP.projection, xg, y)
    torch.cuda.synchronize()
    fp_ms, fp_n = rt.profile_read(0)
    bp_ms, bp_n = rt.profile_read(1)
    rt.profile_enable(False)
    fp_ms, bp_ms = fp_ms / max(fp_n, 1), bp_ms / max(bp_n, 1)
    b_alg = 4.0 * V * (N * N + D) * B  # bytes per operator launch over the batch
    # the adjoint kernel of the chosen variant
    adj = {rt.ADJ_QUAD: "k_pixels_quad", rt.ADJ_TILE16: "k_pixels_tile",
           rt.ADJ_TILE8: "k_pixels_tile"}.get(BP.variant, "k_pixels")
    kernels = {
        "k_bins_split + k_bins_reduce (A, image->sino)":
            {"ms": fp_ms, "effective_GBps": b_alg / (fp_ms * 1e-3) / 1e9},
        f"{adj} (A^T, sino->image)": {"ms": bp_ms, "effective_GBps": b_alg / (bp_ms * 1e-3) / 1e9},
    }
    dom_ms = max(fp_ms, bp_ms)
    dom_name = (f"k_bins_split<{BP.tile},0>" if fp_ms >= bp_ms else f"{adj}<{BP.tile},0>")
    roofline = roofline_c1(dom_name, dom_ms, b_alg, t_batched)

    if args.copy_probe:
        src = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()

    # ---- accuracy --------------------------------------------------------
    acc = {"batched_vs_plain_grad_rel_l2": float(
        (grad_batched - grad_plain).double().norm() / grad_plain.double().norm())}
    x_np = x_true[:, 0].cpu().numpy()
    y_np = y[:, 0].cpu().numpy()
    if not args.no_accuracy:
        from oracle import oracle as O
        g = O.Geom2D(N, N, D, PIX, PIX, DET / SDD, SID, SDD)
        with torch.no_grad():
            fp = BP.projection(x0)[:1, 0].cpu().numpy()
            yt = torch.rand(1, 1, V, D, device=dev, generator=torch.Generator(device=dev).manual_seed(3))
            yt = yt.expand(B, 1, V, D).contiguous()
            pt = BP.projection_t(yt)[:1, 0].cpu().numpy()
        acc.update(_accuracy_vs(O, "arc", "proj", x0[:1, 0].cpu().numpy(), fp, ang_np, g))
        acc.update(_accuracy_vs(O, "arc", "proj_t", yt[:1, 0].cpu().numpy(), pt, ang_np, g))
        acc["note"] = ("one slice of the batch-16 planned path; the reference's own FP32 CPU "
                       "code is *_reference_floor from FP64 (sequential FP32 pixel-edge "
                       "accumulation, FanArcDisCPU2d.cpp:131)")

    c2 = None if args.no_c2 else uneven_c2(torch, dev, max(3, args.steps // 2), 2,
                                           not args.no_accuracy, not args.no_tune,
                                           cpu=not args.no_cpu)
    if c2 is not None:
        c2.pop("detector_samples")
    legs = {}
    if not args.no_3d:
        for kind in ("cone_flat", "helical"):
            r = bench3d(torch, kind, 3, 1, not args.no_accuracy,
                        dc_loop=(kind == "helical" and not args.no_dc_loop))
            legs[kind] = {k: r[k] for k in ("config", "ms_per_step", "value", "roofline", "kernels",
                                            "accuracy", "dc_loop") if k in r}
    c0 = None if args.no_c0 else parallel_c0(torch, dev, not args.no_accuracy)
    fid = None if args.no_fidelity else bench_fidelity(torch, max(5, args.steps // 2), 2,
                                                       ctx=(P, ang, x0, y, BP))
    cpu = None if args.no_cpu else cpu_baseline(ang_np, x_np, y_np)
    samples = B * V * D
    out = {
        "metric": METRIC,
        "value": samples / t_batched,
        "unit": "detector samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batched * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipse phantoms, Poisson noise at 1e5 photons)",
        "config": {"workload": "fan-arc 512x512, 720 views, 736 bins, batch 16, autograd FP + adjoint",
                   "image": [N, N], "views": V, "bins": D, "batch": B,
                   "pixel_mm": PIX, "det_mm": DET, "sid_mm": SID, "sdd_mm": SDD},
        "plain": {"ms_per_step": t_plain * 1e3, "value": samples / t_plain,
                  "batched_speedup_per_slice": t_plain / t_batched,
                  "running_sums": {
                      "ms_per_step": t_plain_rs * 1e3,
                      "batched_speedup_per_slice": t_plain_rs / t_batched,
                      "grad_rel_l2_vs_batched": float(
                          (grad_batched - grad_plain_rs).double().norm()
                          / grad_batched.double().norm())}},
        "autotune": tune,
        "roofline": roofline,
        "kernels": kernels,
        "accuracy": acc,
        "uneven_c2": c2,
        "cone_flat_c3": legs.get("cone_flat"),
        "helical_c4": legs.get("helical"),
        "parallel_c0": c0,
        "fidelity_step": fid,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
