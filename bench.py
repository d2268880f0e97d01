"""Benchmarks of the rendering hot path on one MI355X. Prints one JSON line.

--workload vcm (default): vertex connection and merging on the VCM workload scene
(selas_amd/workloads.py vcm_scene: 64 displaced-grid quad meshes, 1,048,576 triangles, per-face
16 x 16 RGB8 textures, one quad and one point light), 512 x 512, max path length 10, merge radius
0.3 % of the scene diagonal shrinking
with alpha 0.75. One step = one VCM iteration (a light path and a camera path per pixel); the timed
region is one sgpu_render_vcm call of `steps` iterations (32 = the workload's job), with the scene
and the BVH resident in HBM. value = camera samples per second; Mrays/s counts every closest-hit
and shadow ray. The roofline object describes the closest-hit traversal kernel inside the render:
algorithmic bytes per ray = 48 B of compulsory ray/hit traffic + nodes visited x 64 B + triangles
tested x 48 B (counted over all closest-hit rays of an untimed iteration), divided by the kernel's
device time summed from HIP events recorded around each launch on the render's stream.

--workload traversal: the BVH traversal microbenchmark (2,097,152 triangles in 256 Gaussian
clusters, 16,777,216 rays, half primary and half bounce rays); one step = one closest-hit pass plus
one any-hit pass over all rays; value in Mrays/s.
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


def load_traffic(name="traffic_bvh_intersect_2Mtri_16Mray.json"):
    """HBM bytes per launch of the closest-hit kernel measured with rocprofv3 --pmc on a workload
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if recorded."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def cpu_baseline_traversal(t9, rays, sample_rays, threads):
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


def cpu_baseline_vcm(sc, width, iterations, threads):
    import oracle as O

    log("cpu baseline: building the oracle SAH tree")
    rs = O.RenderScene(sc)
    log(f"cpu baseline: VCM {width}x{width}, {iterations} iterations on {threads} threads")
    t0 = time.perf_counter()
    fb, st = O.render_vcm(rs, sc["camera"], width, width, iterations, sc["max_path_length"], sc["merge_radius"],
                          sc["radius_alpha"], 1, threads=threads)
    secs = time.perf_counter() - t0
    return {
        "value": round(width * width * iterations / secs, 1),
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{iterations} VCM iterations at {width}x{width} on the same scene ({secs:.2f} s, "
                  f"{st['rays'] / secs / 1e6:.2f} Mrays/s), oracle restatement (scalar C, oracle SAH tree, not Embree)",
        "image_mean": round(float(fb.mean()), 5),
    }


def bench_vcm(args):
    from selas_amd import native as N
    from selas_amd import workloads as W

    steps = 32 if args.steps is None else args.steps
    w = args.width
    log("building VCM workload scene")
    sc = W.vcm_scene()
    n_tris = sum(W.geom_tri9(g).shape[0] for g in sc["geoms"])
    ctx = N.Context(0)
    tb = time.perf_counter()
    scene = N.build_render_scene(ctx, sc)
    build_ms = (time.perf_counter() - tb) * 1e3
    info = scene.info()
    log(f"scene: {info}, upload + build {build_ms:.1f} ms")
    L, r, alpha = sc["max_path_length"], sc["merge_radius"], sc["radius_alpha"]
    cam = sc["camera"]
    fb = ctx.empty(3 * w * w * 4)

    # traversal counts over the closest-hit rays of one untimed iteration
    _, cnt = N.render_vcm(scene, cam, w, w, 1, L, r, alpha, 1, iteration_base=0, return_stats=True,
                          flags=N.VCM_COUNT_TRAVERSAL, framebuffer=fb)
    nodes_per_ray = cnt["nodes_visited"] / cnt["closest_rays"]
    tris_per_ray = cnt["tris_tested"] / cnt["closest_rays"]
    bytes_per_ray = 48 + nodes_per_ray * 64 + tris_per_ray * 48
    log(f"closest hit in VCM: {nodes_per_ray:.2f} nodes/ray, {tris_per_ray:.2f} tris/ray, {bytes_per_ray:.1f} B/ray")
    if args.warmup:
        N.render_vcm(scene, cam, w, w, args.warmup, L, r, alpha, 1, framebuffer=fb)
    N.synchronize()

    e0, e1 = N.Event(), N.Event()
    t0 = time.perf_counter()
    e0.record()
    _, st = N.render_vcm(scene, cam, w, w, steps, L, r, alpha, 1, return_stats=True, flags=N.VCM_KERNEL_TIMES,
                         framebuffer=fb)
    e1.record()
    N.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    scene.check()
    img = fb.download(np.float32, 3 * w * w)
    fb.free()
    secs = ms / 1e3
    samples = w * w * steps
    value = samples / secs
    closest_ms = st["closest_ms"]
    achieved = st["closest_rays"] * bytes_per_ray / (closest_ms / 1e3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_vcm(sc, w, args.cpu_iterations, min(16, os.cpu_count() or 1))

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"vcm_{w}x{w}_{steps}it_L{L}_1Mtri",
            "triangles": int(n_tris),
            "width": w,
            "height": w,
            "iterations": steps,
            "max_path_length": L,
            "merge_radius": round(float(r), 6),
            "radius_alpha": alpha,
            "lights": "one quad light + one point light",
            "textures": "per-face 16x16 RGB8 value noise (Ptex f_bspline), 403 MB",
        },
        "mrays_s": round(st["rays"] / secs / 1e6, 2),
        "rays_per_sample": round(st["rays"] / samples, 3),
        "closest_rays": st["closest_rays"],
        "shadow_rays": st["shadow_rays"],
        "light_vertices_per_iteration": round(st["light_vertices"] / steps, 1),
        "phase_ms": {k: round(st[k], 2) for k in ("closest_ms", "shadow_ms", "light_ms", "camera_ms", "merge_ms",
                                                  "grid_ms")},
        "bvh_build_ms": round(build_ms, 1),
        "wall_s_timed_region": round(wall, 3),
        "image_mean": round(float(img.mean()), 5),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic("traffic_bvh_intersect_vcm_1Mtri.json"),
            "kernel": "k_trace<false,false> (closest hit, all launches of the render)",
            "launches": st["closest_launches"],
            "avg_launch_ms": round(closest_ms / max(1, st["closest_launches"]), 4),
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("vcm", "traversal"), default="vcm")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (vcm: 32 iterations, traversal: 10)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=512, help="vcm: image width and height")
    ap.add_argument("--cpu-iterations", type=int, default=2, help="vcm: iterations in the CPU baseline sample")
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
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}). VCM would shard "
              "by iterations (each GPU its own light vertices and photon grid, one sum of the frame "
              "buffers at the end) and ray casts by ray ranges with a replicated BVH; neither is "
              "measured here.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "vcm":
        bench_vcm(args)
    else:
        if args.steps is None:
            args.steps = 10
        bench_traversal(args)


def bench_traversal(args):
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
        cpu = cpu_baseline_traversal(t9, rays, args.cpu_rays, min(16, os.cpu_count() or 1))

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
