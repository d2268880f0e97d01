"""Autotuner for the launch shape of the rendering path (sgpu_set_tuning).

Two parameters decide how well the wavefront kernels fill an MI355X:

* vcm_wavefront_paths: paths per VCM wavefront, i.e. the capacity of every path, ray, hit, light
  vertex, photon-key and connection-slot queue of a render. Larger wavefronts put more rays into
  each traversal launch (hiding the latency of dependent node fetches) and more queries into each
  merge launch, at the cost of device memory (the server's footprint grows linearly).
* trace_blocks_per_cu: the traversal grid. Rays beyond the grid are taken grid-stride, so more
  blocks than fit on a CU at once balance the load between waves of uneven traversal length.

The search is coordinate-wise: wavefront sizes at the default grid, then grids at the best size,
then sizes again at the best grid; a size whose server footprint does not fit the memory budget is
skipped. `probe(ctx)` runs the workload to tune for once, synchronously; the best of `repeats`
timed runs (after one untimed run) is its cost. Results can be kept in a JSON file keyed by device
name and workload, so a service tunes once per device.
"""
from __future__ import annotations

import json
import os
import time

WAVEFRONT_PATHS = [1 << k for k in range(18, 25)]
BLOCKS_PER_CU = [4, 8, 16, 24, 32]


def _timed(probe, ctx, repeats, timer):
    probe(ctx)
    best = float("inf")
    for _ in range(repeats):
        t0 = timer()
        probe(ctx)
        best = min(best, timer() - t0)
    return best


def tune(ctx, probe, wavefront_paths=None, blocks_per_cu=None, memory_budget=None, max_path_length=10,
         repeats=2, log=None, footprint=None, timer=time.perf_counter) -> dict:
    """Searches the launch shape for `probe` on `ctx` and leaves the best one set on the context.
    Returns {"best": {...}, "seconds": s, "default_seconds": s0, "trials": [...]}. memory_budget
    (bytes, default: 80 % of the free device memory) bounds the VCM server footprint of a size
    (`footprint(ctx, paths, max_path_length)`, by default sgpu_vcm_server_footprint)."""
    if footprint is None:
        from . import native as N
        footprint = N.VcmServer.footprint

    sizes = sorted(wavefront_paths or WAVEFRONT_PATHS)
    grids = sorted(blocks_per_cu or BLOCKS_PER_CU)
    if memory_budget is None:
        memory_budget = int(0.8 * ctx.device_info()["free_bytes"])
    fits = [s for s in sizes if footprint(ctx, s, max_path_length) <= memory_budget]
    if not fits:
        raise RuntimeError("autotune: no wavefront size fits the memory budget")
    start = ctx.tuning()
    cache = {}
    trials = []

    def cost(size, grid):
        key = (size, grid)
        if key not in cache:
            ctx.set_tuning(vcm_wavefront_paths=size, trace_blocks_per_cu=grid)
            cache[key] = _timed(probe, ctx, repeats, timer)
            trials.append({"vcm_wavefront_paths": size, "trace_blocks_per_cu": grid, "seconds": cache[key]})
            if log:
                log(f"autotune: wavefront {size}, {grid} blocks/CU: {cache[key] * 1e3:.1f} ms")
        return cache[key]

    default_grid = start["trace_blocks_per_cu"] if start["trace_blocks_per_cu"] in grids else grids[len(grids) // 2]
    default_size = start["vcm_wavefront_paths"]
    default_seconds = cost(default_size, default_grid) if default_size in fits else None
    size = min(fits, key=lambda s: cost(s, default_grid))
    grid = min(grids, key=lambda g: cost(size, g))
    size = min(fits, key=lambda s: cost(s, grid))
    best = {"vcm_wavefront_paths": size, "trace_blocks_per_cu": grid}
    seconds = cost(size, grid)
    if "trace_lds_nodes" in start:
        # the top of the 8-wide tree in LDS, tried at the best shape
        for lds in sorted({0, 9} - {start["trace_lds_nodes"]}):
            ctx.set_tuning(trace_lds_nodes=lds, **best)
            t = _timed(probe, ctx, repeats, timer)
            trials.append(dict(best, trace_lds_nodes=lds, seconds=t))
            if log:
                log(f"autotune: {lds} top nodes in LDS: {t * 1e3:.1f} ms")
            if t < seconds:
                seconds = t
                best["trace_lds_nodes"] = lds
        best.setdefault("trace_lds_nodes", start["trace_lds_nodes"])
    ctx.set_tuning(**best)
    return {"best": dict(ctx.tuning(), **best), "seconds": seconds, "default_seconds": default_seconds,
            "trials": trials}


def cache_key(ctx, workload: str) -> str:
    return f"{ctx.device_info()['name']}|{workload}"


def load(path: str, key: str):
    """The tuning stored for `key` in the JSON file at `path`, or None."""
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def save(path: str, key: str, best: dict):
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[key] = best
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
