"""Autotuner for the batched projector paths.

Sweeps the launch parameters of a planned operator on the problem it will
run — the instance tile (how many instances share one geometry evaluation,
1-16), the sinogram -> image kernel variant (per pixel, 16x16 or 8x8 tiles
with LDS edge positions, or one thread per 2x2 pixels) and the workgroup
tiles of both gathers (forward: 63-bin waves x views per workgroup over one
per-XCD line chunk; per-pixel adjoint: 16x16, 32x8 or 64x4 pixels) of a
uniform batch; the grid size and kernel variant of an uneven batch —
times one forward projection plus one adjoint per candidate with HIP events,
keeps the fastest, and remembers it in a JSON cache keyed by the device and
the geometry, so later runs skip the sweep.

Cache file: ``$PRIDEDIFF_AMD_TUNE_CACHE`` or
``~/.cache/pridediff_amd/autotune.json``.
"""
from __future__ import annotations

import hashlib
import json
import os
import statistics

import torch

from pridediff_amd.batched import BatchedProjector2d, UnevenBatch2d, rt

CACHE = os.environ.get("PRIDEDIFF_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "pridediff_amd", "autotune.json"))


SWEEP_VERSION = 9  # bumped when the set of swept parameters changes

_LIB_DIGEST = None


def _lib_digest() -> str:
    """Digest of the built libdd.so: any rebuild of the kernels invalidates
    every cached choice, with no version to remember to bump."""
    global _LIB_DIGEST
    if _LIB_DIGEST is None:
        lib = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdd.so")
        h = hashlib.sha1()
        with open(lib, "rb") as f:
            for block in iter(lambda: f.read(1 << 20), b""):
                h.update(block)
        _LIB_DIGEST = h.hexdigest()[:16]
    return _LIB_DIGEST


def _device_key() -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(p, "gcnArchName", "") or p.name
    return f"s{SWEEP_VERSION}|lib:{_lib_digest()}|{arch}|{p.multi_processor_count}cu"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass  # a read-only home only costs a re-sweep next time


def time_ms(fn, reps: int = 5, warmup: int = 2) -> float:
    """Median milliseconds of ``fn()`` on the current stream (HIP events)."""
    for _ in range(warmup):
        fn()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    return statistics.median(times)


FWD_SHAPES = ((4, 1), (2, 1), (1, 1), (4, 2), (2, 2), (1, 4))  # (waves, views) per workgroup
PIX_WIDTHS = (16, 32, 64)                                        # per-pixel adjoint workgroups


def tune_uniform(projector, angles: torch.Tensor, batch: int, tiles=(1, 2, 4, 8, 16),
                 variants=None, use_cache: bool = True, shapes: bool = True,
                 report: dict | None = None, running_sums: bool = True) -> BatchedProjector2d:
    """Fastest ``BatchedProjector2d`` for this projector, angle set and batch:
    first the instance tile x adjoint kernel variant, then, for the winner,
    the workgroup tile of each gather — forward (63-bin waves x views per
    workgroup; its lines are one of the 8 per-XCD chunks) and per-pixel
    adjoint (pixels per workgroup row).  The walk variants all give the same
    bits; ``running_sums`` also sweeps, at tile 1, the running-sum adjoint
    alone and both directions, which round differently (within 1e-6 of the
    walks)."""
    if variants is None:
        variants = (rt.ADJ_PER_PIXEL, rt.ADJ_TILE16, rt.ADJ_TILE8, rt.ADJ_QUAD)
    opts = [round(float(v), 6) for v in projector.host_options().tolist()]
    key = (f"{_device_key()}|uniform|{projector.scan_type}|{opts}|V={angles.numel()}|B={batch}"
           f"|rs{int(running_sums)}")
    table = _load() if use_cache else {}
    if key in table:
        if report is not None:
            report.update(table[key])
        return BatchedProjector2d(projector, angles, batch, tile=table[key]["tile"],
                                  variant=table[key]["variant"],
                                  shape=tuple(table[key].get("shape", (0, 0, 0))))
    h, w = projector.image_size[1], projector.image_size[0]
    x = torch.rand(batch, 1, h, w, device=angles.device)
    y = torch.rand(batch, 1, angles.numel(), projector.num_det, device=angles.device)
    results = {}
    for t in tiles:
        if t > 1 and t // 2 >= batch:
            continue
        # the single-call running sums run at tile 1 only
        extra = ((rt.ADJ_BLOCK_PREFIX, rt.ADJ_BLOCK_PREFIX | rt.FWD_PREFIX)
                 if t == 1 and running_sums else ())
        for var in tuple(variants) + extra:
            bp = BatchedProjector2d(projector, angles, batch, tile=t, variant=var)
            with torch.no_grad():
                results[(t, var)] = time_ms(lambda: (bp.plan.run(rt.PROJ, x),
                                                     bp.plan.run(rt.PROJ_T, y)))
            del bp
    best = min(results, key=results.get)
    swept = {f"tile{k[0]}/var{k[1]}": v for k, v in results.items()}
    shape, best_ms = (0, 0, 0), results[best]
    if shapes:
        def timed_shape(sh):
            bp = BatchedProjector2d(projector, angles, batch, tile=best[0], variant=best[1],
                                    shape=sh)
            with torch.no_grad():
                ms = time_ms(lambda: (bp.plan.run(rt.PROJ, x), bp.plan.run(rt.PROJ_T, y)))
            del bp
            return ms
        fwd = {}
        for wv in FWD_SHAPES:
            fwd[wv] = timed_shape((wv[0], wv[1], 0))
            swept[f"tile{best[0]}/var{best[1]}/fwd{wv[0]}x{wv[1]}"] = fwd[wv]
        wv = min(fwd, key=fwd.get)
        pix = {0: fwd[wv]}
        if best[1] == rt.ADJ_PER_PIXEL:
            for pw in PIX_WIDTHS:
                pix[pw] = timed_shape((wv[0], wv[1], pw))
                swept[f"tile{best[0]}/var{best[1]}/fwd{wv[0]}x{wv[1]}/pix{pw}"] = pix[pw]
        pw = min(pix, key=pix.get)
        shape, best_ms = (wv[0], wv[1], pw), pix[pw]
    entry = {"tile": best[0], "variant": best[1], "shape": list(shape), "ms": best_ms,
             "swept": swept}
    if use_cache:
        table = _load()
        table[key] = entry
        _store(table)
    if report is not None:
        report.update(entry)
    return BatchedProjector2d(projector, angles, batch, tile=best[0], variant=best[1], shape=shape)


def tune_uneven(batch: UnevenBatch2d, grids_per_cu=(0, 2, 8, 32, 128, 512), variants=None,
                use_cache: bool = True, report: dict | None = None,
                prefix: bool = True) -> UnevenBatch2d:
    """Sets the fastest grid size and kernel variant on ``batch``
    (in place).  Grid 0 is one workgroup per work item (the default).
    ``prefix``: also sweep the running-sum adjoint (``ADJ_BLOCK_PREFIX``)
    and try the winner with the running-sum forward (``FWD_PREFIX``); both
    round differently from the walks (same operator, errors of the same
    size against FP64)."""
    if variants is None:
        variants = (rt.ADJ_PER_PIXEL, rt.ADJ_TILE16, rt.ADJ_QUAD, rt.ADJ_BLOCK_WINDOW)
        if prefix:
            variants += (rt.ADJ_BLOCK_PREFIX,)
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    sig = sorted(batch.shapes)
    key = f"{_device_key()}|uneven|{batch.scan_type}|{sig}|prefix{int(prefix)}"
    table = _load() if use_cache else {}
    if key in table:
        batch.plan.set_grid(table[key]["grid"])
        batch.plan.set_variant(table[key]["variant"])
        if report is not None:
            report.update(table[key])
        return batch
    dev = torch.device("cuda", torch.cuda.current_device())
    x = torch.rand(batch.plan.image_size, device=dev)
    y = torch.rand(batch.plan.sino_size, device=dev)
    results = {}
    for gpc in grids_per_cu:
        for var in variants:
            batch.plan.set_grid(gpc * cus)
            batch.plan.set_variant(var)
            with torch.no_grad():
                results[(gpc * cus, var)] = time_ms(lambda: (batch.plan.run(rt.PROJ, x),
                                                             batch.plan.run(rt.PROJ_T, y)))
    best = min(results, key=results.get)
    if prefix:
        # the running-sum forward on the winner (not bit-equal to the walk:
        # same operator, its own rounding)
        var = best[1] | rt.FWD_PREFIX
        batch.plan.set_grid(best[0])
        batch.plan.set_variant(var)
        with torch.no_grad():
            results[(best[0], var)] = time_ms(lambda: (batch.plan.run(rt.PROJ, x),
                                                       batch.plan.run(rt.PROJ_T, y)))
        best = min(results, key=results.get)
    batch.plan.set_grid(best[0])
    batch.plan.set_variant(best[1])
    entry = {"grid": best[0], "variant": best[1], "ms": results[best],
             "swept": {f"grid{k[0]}/var{k[1]}": v for k, v in results.items()}}
    if use_cache:
        table = _load()
        table[key] = entry
        _store(table)
    if report is not None:
        report.update(entry)
    return batch
