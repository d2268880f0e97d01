"""Benchmark of the ray-cast hot path on one MI355X.

Workload (selas_amd/workloads.py): 2,097,152 triangles in 256 Gaussian clusters and 16,777,216
rays (half primary camera rays, half cosine bounce rays). One step = one closest-hit pass
(sgpu_intersect) plus one any-hit pass (sgpu_occluded) over all rays, inputs resident in HBM.
value = rays traced per second over both passes, in Mrays/s.

Prints one JSON line. The roofline object describes the dominant kernel (closest hit):
algorithmic bytes per ray = 48 B of compulsory ray/hit traffic + nodes visited x 64 B + triangles
tested x 48 B (counted by the instrumented traversal on all rays), divided by its average duration
measured with HIP events on the stream it runs on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "VCM samples/s + Mrays/s at 1M tris; p50/p99 job latency vs plain GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic():
    """HBM bytes per launch of the closest-hit kernel measured with rocprofv3 --pmc on this
    workload (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if recorded."""
    p = os.path.join(ROOT, "profiles", "traffic_bvh_intersect_2Mtri_16Mray.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def cpu_baseline(t9, rays, sample_rays, threads):
    import oracle as O

    log(f"cpu baseline: building the oracle SAH tree over {t9.shape[0]} triangles")
    bvh = O.CpuBvh(t9)
    n = rays["org_x"].shape[0]
    # same mix as the GPU workload: equal parts of primary and bounce rays
    half = n // 2
    k = sample_rays // 2
    sel = np.concatenate([np.arange(0, half, max(1, half // k))[:k], half + np.arange(0, half, max(1, half // k))[:k]])
    sub = {kk: np.ascontiguousarray(v[sel]) for kk, v in rays.items()}
    log(f"cpu baseline: {sel.size} rays closest + any hit on {threads} threads")
    t0 = time.perf_counter()
    h = bvh.closest(sub, threads)
    t1 = time.perf_counter()
    bvh.occluded(sub, threads)
    t2 = time.perf_counter()
    secs = t2 - t0
    return {
        "value": round(2 * sel.size / secs / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sel.size} rays (every {max(1, half // k)}th primary and bounce ray), closest hit "
                  f"{t1 - t0:.2f} s + any hit {t2 - t1:.2f} s, oracle SAH tree (scalar C, not Embree)",
        "hit_fraction": round(float((h['prim'] != 0xFFFFFFFF).mean()), 4),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rays", type=int, default=4194304, help="rays in the CPU baseline sample")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the triangles (1.0 = the workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run sgpu_device_copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        from selas_amd import native as N
        ctx = N.Context(0)
        a = ctx.empty(args.copy_calibration)
        b = ctx.empty(args.copy_calibration)
        for _ in range(3):
            N.check(N.lib.sgpu_device_copy(b.ptr, a.ptr, args.copy_calibration, None))
        N.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration, "launches": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the ray-cast "
              "workload shards by ray ranges with a replicated BVH, which is not measured here.", file=sys.stderr)
        sys.exit(2)

    from selas_amd import native as N
    from selas_amd import workloads as W

    log("building workload")
    t9 = W.traversal_scene(args.scale)
    rays = W.traversal_rays(t9)
    n = int(rays["org_x"].shape[0])

    ctx = N.Context(0)
    stream = None
    scene = N.Scene(ctx)
    scene.add_soup(t9)
    tb = time.perf_counter()
    scene.commit()
    build_ms = (time.perf_counter() - tb) * 1e3
    info = scene.info()
    log(f"scene: {info}, build {build_ms:.1f} ms")

    dr = N.DeviceRays(ctx, rays)
    dh = N.DeviceHits(ctx, n, with_ng=False)
    scene.intersect(dr, dh)
    h = dh.to_host()
    tfar, _ = W.anyhit_tfar(h["t"], h["prim_id"])
    rays_any = dict(rays)
    rays_any["tfar"] = tfar
    dra = N.DeviceRays(ctx, rays_any)
    occ = ctx.empty(4 * n)

    stats = scene.intersect_stats(dr)
    nodes_per_ray = stats["nodes_visited"] / n
    tris_per_ray = stats["tris_tested"] / n
    bytes_per_ray = 48 + nodes_per_ray * stats["bytes_node"] + tris_per_ray * stats["bytes_tri"]
    log(f"closest hit: {nodes_per_ray:.2f} nodes/ray, {tris_per_ray:.2f} tris/ray, {bytes_per_ray:.1f} B/ray")

    for _ in range(args.warmup):
        scene.intersect(dr, dh, stream)
        scene.occluded(dra, occ, stream)
    N.synchronize()

    ev = [(N.Event(), N.Event(), N.Event()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b, c in ev:
        a.record(stream)
        scene.intersect(dr, dh, stream)
        b.record(stream)
        scene.occluded(dra, occ, stream)
        c.record(stream)
    N.synchronize()
    wall = time.perf_counter() - t0
    scene.check()
    ms_isect = [a.elapsed_ms(b) for a, b, _ in ev]
    ms_occ = [b.elapsed_ms(c) for _, b, c in ev]
    ms_step = [a.elapsed_ms(c) for a, _, c in ev]
    total_ms = sum(ms_step)
    avg_isect = sum(ms_isect) / len(ms_isect)
    avg_occ = sum(ms_occ) / len(ms_occ)
    value = 2 * n * args.steps / (total_ms / 1e3) / 1e6
    achieved = n * bytes_per_ray / (avg_isect / 1e3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(t9, rays, args.cpu_rays, min(16, os.cpu_count() or 1))

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "bvh_traversal_2Mtri_16Mray",
            "triangles": int(t9.shape[0]),
            "rays": n,
            "passes": "closest hit + any hit",
            "bvh": "device LBVH, binary nodes (64 B), 1 triangle per leaf (48 B)",
        },
        "closest_hit_mrays_s": round(n / (avg_isect / 1e3) / 1e6, 2),
        "any_hit_mrays_s": round(n / (avg_occ / 1e3) / 1e6, 2),
        "kernel_ms": {"intersect": round(avg_isect, 3), "occluded": round(avg_occ, 3)},
        "bvh_build_ms": round(build_ms, 1),
        "wall_s_timed_region": round(wall, 3),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(),
            "kernel": "k_trace<false,false> (closest hit)",
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
