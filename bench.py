"""Benchmarks of the rendering hot path on one MI355X. Prints one JSON line.

--workload vcm (default): vertex connection and merging on the VCM workload scene
(selas_amd/workloads.py vcm_scene: 64 displaced-grid quad meshes, 1,048,576 triangles, per-face
16 x 16 RGB8 textures, one quad and one point light), 512 x 512, 32 iterations, max path length 10,
merge radius 0.3 % of the scene diagonal shrinking with alpha 0.75. One step = one whole job (the
512 x 512 x 32-iteration render, from scratch); the timed region is `steps` such sgpu_render_vcm
calls with the scene and the BVH resident in HBM. value = camera samples per second; Mrays/s
counts every closest-hit and shadow ray. HIP events recorded around every launch on the render's
stream time the kernels. The roofline object describes the vertex-merging kernel (k_cam_merge, the
largest kernel of the render): its algorithmic bytes are counted over an untimed run of the same
job (camera vertices read, range searches, photons visited, photons within the radius; see
merge_bytes); roofline_closest_hit does the same for the closest-hit traversal (48 B of ray/hit
traffic + nodes visited x node bytes + triangles tested x 48 B per ray). cpu_baseline renders the
same job with the oracle on the host's cores and compares the images pixel by pixel.

The default run adds two legs to the same line (skip them with --no-batch / --no-serve):
"batched": 64 uneven jobs (square images with side log-uniform in [64, 1024], 1..64 iterations,
each on a scene of its own of 10K..2M triangles), rendered as one batch and, for comparison, one job
at a time with its iterations packed (sgpu_render_vcm, the solo schedule) and one job at a time
with one wavefront per iteration (the plain GPU schedule); "serving": a Poisson stream of such
requests over 16 resident scenes at --serve-rate requests/s (more than the GPU serves, so requests
queue) through the request loop (selas_amd/serving.py), p50/p99 latency from arrival to the image
on the host, beside the plain schedule on the same stream. The top-level p50_ms/p99_ms,
vs_plain_gpu and vs_solo_gpu come from these legs.

--workload batch / serve: the batched leg alone (with --solo, also one job at a time with its
iterations packed), and the serving leg swept over --rates.

--workload pt: unidirectional path tracing of the closed room (24 + 20,480 triangles), 256 x 256,
16 samples per pixel, path length 8; one step = the whole image.

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


def cpu_threads():
    """Threads for the CPU baselines: the host CPUs this process may run on, at most 16 (the CPU
    share of one GPU on an 8-GPU host), with what the host reports beside it."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail)), {"host_cpus": os.cpu_count(), "affinity_cpus": avail,
                                    "share_rule": "min(16, CPUs this process may run on): one GPU's share of the host"}


def cpu_baseline_vcm(sc, width, iterations, threads, seed=1):
    import oracle as O

    log("cpu baseline: building the oracle SAH tree")
    rs = O.RenderScene(sc)
    log(f"cpu baseline: VCM {width}x{width}, {iterations} iterations on {threads} threads")
    t0 = time.perf_counter()
    fb, st = O.render_vcm(rs, sc["camera"], width, width, iterations, sc["max_path_length"], sc["merge_radius"],
                          sc["radius_alpha"], seed, threads=threads)
    secs = time.perf_counter() - t0
    return {
        "value": round(width * width * iterations / secs, 1),
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the whole job ({iterations} VCM iterations at {width}x{width}, seed {seed}) on the same scene "
                  f"({secs:.2f} s, {st['rays'] / secs / 1e6:.2f} Mrays/s), oracle restatement (scalar C, oracle "
                  f"SAH tree, not Embree)",
        "image_mean": round(float(fb.mean()), 5),
    }, fb


# The VCM workload's job: 512 x 512, 32 iterations (one step of the vcm workload).
C3_ITERATIONS = 32

# Algorithmic bytes of the vertex-merging kernel (k_cam_merge): per camera vertex read 80 B (its
# CamVertex record); per range search 4 B (its path's slice) + 8 cells x 8 B (two cell ends) + 16 B
# (the contribution slot and its flag); per photon visited 16 B (its position); per photon within
# the radius 80 B (its light-vertex record).
MERGE_B_ENTRY, MERGE_B_QUERY, MERGE_B_VISIT, MERGE_B_NEAR = 80, 84, 16, 80


def merge_bytes(st):
    return (st["merge_entries"] * MERGE_B_ENTRY + st["merge_queries"] * MERGE_B_QUERY +
            st["merge_visited"] * MERGE_B_VISIT + st["merge_near"] * MERGE_B_NEAR)


def bench_vcm(args):
    from selas_amd import native as N
    from selas_amd import workloads as W

    steps = 5 if args.steps is None else args.steps
    its = C3_ITERATIONS
    w = args.width
    log("building VCM workload scene")
    sc = W.vcm_scene()
    n_tris = sum(W.geom_tri9(g).shape[0] for g in sc["geoms"])
    ctx = new_context(N, args)
    tb = time.perf_counter()
    scene = N.build_render_scene(ctx, sc)
    build_ms = (time.perf_counter() - tb) * 1e3
    info = scene.info()
    log(f"scene: {info}, upload + build {build_ms:.1f} ms")
    L, r, alpha = sc["max_path_length"], sc["merge_radius"], sc["radius_alpha"]
    cam = sc["camera"]
    fb = ctx.empty(3 * w * w * 4)

    def job(flags=0, stats=False):
        return N.render_vcm(scene, cam, w, w, its, L, r, alpha, 1, return_stats=stats, flags=flags, framebuffer=fb)

    tuning = apply_tuning(args, ctx, lambda c: job(), workload=f"vcm_{w}x{w}_{its}it_L{L}_{n_tris}tri")
    # work counters of the same job (untimed): traversal of every closest-hit ray and the vertex
    # merging's range searches, on the tree in use
    _, cnt = job(N.VCM_COUNT_TRAVERSAL | N.VCM_COUNT_MERGE, stats=True)
    node_bytes = 128 if ctx.tuning()["trace_tree"] == 1 else 64
    nodes_per_ray = cnt["nodes_visited"] / cnt["closest_rays"]
    tris_per_ray = cnt["tris_tested"] / cnt["closest_rays"]
    bytes_per_ray = 48 + nodes_per_ray * node_bytes + tris_per_ray * 48
    m_bytes_job = merge_bytes(cnt)
    log(f"closest hit in VCM: {nodes_per_ray:.2f} nodes/ray ({node_bytes} B), {tris_per_ray:.2f} tris/ray, "
        f"{bytes_per_ray:.1f} B/ray")
    log(f"merging: {cnt['merge_entries']} camera vertices, {cnt['merge_queries']} range searches, "
        f"{cnt['merge_visited']} photons visited, {cnt['merge_near']} within the radius, "
        f"{cnt['merge_launches']} launches, {m_bytes_job / 1e6:.1f} MB per job")
    # warm-up: `warmup` whole jobs, untimed
    for _ in range(args.warmup):
        job()
    N.synchronize()

    # timed: `steps` whole jobs (512 x 512, 32 iterations each), each rendered from scratch; events
    # around every launch time the kernels on the render's stream
    flags = 0 if args.no_phase_times else N.VCM_KERNEL_TIMES
    tot = None
    e0, e1 = N.Event(), N.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        _, st = job(flags, stats=True)
        tot = st if tot is None else {k: tot[k] + st[k] for k in tot}
    e1.record()
    N.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    st = tot
    scene.check()
    img = fb.download(np.float32, 3 * w * w).reshape(w, w, 3)
    # the same job as captured graphs replay it (no per-launch events)
    replay = None
    if not args.no_phase_times and ctx.tuning()["vcm_graphs"]:
        ga, gb = N.Event(), N.Event()
        ga.record()
        for _ in range(max(1, min(steps, 3))):
            job()
        gb.record()
        N.synchronize()
        g_ms = ga.elapsed_ms(gb) / max(1, min(steps, 3))
        replay = {"samples_per_s": round(w * w * its / (g_ms / 1e3), 1), "ms_per_job": round(g_ms, 2),
                  "note": "the light and camera passes of each wavefront replayed as captured hipGraphs"}
    if args.no_phase_times:
        # per-launch device times from one more (untimed) job with events
        _, st_t = job(N.VCM_KERNEL_TIMES, stats=True)
        for k in ("closest_ms", "shadow_ms", "light_ms", "camera_ms", "merge_ms", "grid_ms", "merge_launches"):
            st[k] = st_t[k] * steps
    fb.free()
    secs = ms / 1e3
    samples = w * w * its * steps
    value = samples / secs
    closest_ms = st["closest_ms"]
    achieved_closest = st["closest_rays"] * bytes_per_ray / (closest_ms / 1e3) / 1e9
    merge_ms = st["merge_ms"]
    m_launches = st["merge_launches"]
    achieved_merge = m_bytes_job * steps / (merge_ms / 1e3) / 1e9
    phases = {k: round(st[k] / steps, 2) for k in ("closest_ms", "shadow_ms", "light_ms", "camera_ms", "merge_ms",
                                                     "grid_ms")}

    cpu = None
    if not args.no_cpu:
        threads, host = cpu_threads()
        cpu, cpu_img = cpu_baseline_vcm(sc, w, its, threads)
        cpu.update(host)
        cpu_img = cpu_img.reshape(w, w, 3)
        cpu["same_job_as_gpu"] = True
        cpu["pixels_within_1e-3"] = round(_frac_close(img, cpu_img), 5)
        cpu["rmse_gpu_vs_cpu_rel"] = round(float(np.sqrt(((img - cpu_img) ** 2).mean()) / max(1e-12, cpu_img.mean())), 6)
    scene.release()
    del sc
    batch = serving = None
    if not args.no_batch:
        batch = batch_summary(run_batch(N, W, ctx, args.batch_jobs, args.batch_seed, not args.no_solo))
    if not args.no_serve:
        serving = serving_summary(serve_sweep(N, W, ctx, [args.serve_rate], args.serve_jobs or 32,
                                              args.batch_seed, wavefront_paths=args.serve_paths)[0])

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
            "workload": f"vcm_{w}x{w}_{its}it_L{L}_1Mtri",
            "step": f"one whole job: {w}x{w}, {its} iterations, rendered from scratch",
            "triangles": int(n_tris),
            "width": w,
            "height": w,
            "iterations": its,
            "max_path_length": L,
            "merge_radius": round(float(r), 6),
            "radius_alpha": alpha,
            "lights": "one quad light + one point light",
            "textures": "per-face 16x16 RGB8 value noise (Ptex f_bspline), 403 MB",
        },
        "mrays_s": round(st["rays"] / secs / 1e6, 2),
        "rays_per_sample": round(st["rays"] / samples, 3),
        "closest_rays_per_job": st["closest_rays"] // steps,
        "shadow_rays_per_job": st["shadow_rays"] // steps,
        "light_vertices_per_iteration": round(st["light_vertices"] / (its * steps), 1),
        "phase_ms_per_job": phases,
        "bvh_build_ms": round(build_ms, 1),
        "wall_s_timed_region": round(wall, 3),
        "image_mean": round(float(img.mean()), 5),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved_merge, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_merge / HBM_PEAK_GBS, 4),
            "traffic": load_traffic("traffic_cam_merge_vcm_512x512_32it_1Mtri.json"),
            "kernel": "sgpu_vcm::k_cam_merge (vertex merging, all launches of the timed jobs)",
            "launches": m_launches,
            "avg_launch_ms": round(merge_ms / max(1, m_launches), 4),
            "bytes_per_launch": round(m_bytes_job * steps / max(1, m_launches)),
            "per_job": {"camera_vertices": cnt["merge_entries"], "range_searches": cnt["merge_queries"],
                        "photons_visited": cnt["merge_visited"], "photons_near": cnt["merge_near"],
                        "bytes": m_bytes_job},
            "bytes_rule": f"{MERGE_B_ENTRY} B per camera vertex + {MERGE_B_QUERY} B per range search + "
                          f"{MERGE_B_VISIT} B per photon visited + {MERGE_B_NEAR} B per photon within the radius",
        },
        "roofline_closest_hit": {
            "bound": "hbm",
            "achieved": round(achieved_closest, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_closest / HBM_PEAK_GBS, 4),
            "traffic": load_traffic("traffic_bvh_intersect_vcm_1Mtri.json"),
            "kernel": ("k_trace8<false,false,false> (closest hit, 8-wide nodes, all launches of the timed jobs)"
                       if node_bytes == 128 else "k_trace<false,false,false> (closest hit, all launches)"),
            "node_bytes": node_bytes,
            "launches": st["closest_launches"],
            "avg_launch_ms": round(closest_ms / max(1, st["closest_launches"]), 4),
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
        },
        "cpu_baseline": cpu,
        "tuning": tuning,
        "graph_replay": replay,
        "batched": batch,
        "serving": serving,
    }
    if serving:
        out["p50_ms"] = serving["served"]["p50_ms"]
        out["p99_ms"] = serving["served"]["p99_ms"]
    if batch:
        out["vs_plain_gpu"] = batch["vs_plain_gpu"]
        out["vs_solo_gpu"] = batch.get("vs_solo_gpu")
    print(json.dumps(out), flush=True)


TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selas_amd", "tuning.json")


def new_context(N, args):
    """A context on device 0, with the tree builder of --bvh-builder (scenes are built with the
    context's builder, so it is set before any scene)."""
    ctx = N.Context(0)
    if getattr(args, "bvh_builder", None) is not None:
        ctx.set_tuning(bvh_builder=args.bvh_builder)
    return ctx


def apply_tuning(args, ctx, probe, workload=None):
    """Launch shape for the timed runs: given on the command line, the defaults (--no-tune), the
    shape the autotuner stored for this GPU architecture and workload (selas_amd/tuning.json, unless
    --retune), or found by the autotuner (selas_amd/autotune.py) with `probe` (the timed workload,
    run once)."""
    from selas_amd import autotune as A

    if args.trace_tree is not None:
        ctx.set_tuning(trace_tree=args.trace_tree)
    if args.order is not None:
        ctx.set_tuning(vcm_order=args.order)
    if args.merge_lanes is not None:
        ctx.set_tuning(vcm_merge_lanes=args.merge_lanes)
    if args.wavefront_paths or args.trace_blocks:
        ctx.set_tuning(vcm_wavefront_paths=args.wavefront_paths, trace_blocks_per_cu=args.trace_blocks)
        return dict(ctx.tuning(), source="command line")
    if args.no_tune:
        return dict(ctx.tuning(), source="defaults")
    arch = ctx.device_info()["name"].split(":")[0]
    key = f"{arch}|{workload}"
    stored = None if (args.retune or workload is None) else A.load(TUNING_FILE, key)
    if stored:
        ctx.set_tuning(**{k: v for k, v in stored.items() if k in ctx.tuning()})
        if args.order is not None:
            ctx.set_tuning(vcm_order=args.order)
        if args.merge_lanes is not None:
            ctx.set_tuning(vcm_merge_lanes=args.merge_lanes)
        given = [k for k in ("order", "merge_lanes") if getattr(args, k) is not None]
        return dict(ctx.tuning(), source=f"stored autotuner result ({os.path.basename(TUNING_FILE)}: {key})"
                    + (f", {' and '.join(given)} from the command line" if given else ""))
    t0 = time.perf_counter()
    r = A.tune(ctx, probe, log=log)
    return dict(r["best"], source="autotuned", tuned_ms=round(r["seconds"] * 1e3, 1),
                default_ms=None if r["default_seconds"] is None else round(r["default_seconds"] * 1e3, 1),
                trials=len(r["trials"]), search_s=round(time.perf_counter() - t0, 1))


def _build_scene(N, W, ctx, seed, tris):
    """A sized VCM workload scene on the device; the host arrays are dropped once it is built."""
    spec = W.sized_vcm_scene(seed, tris)
    n = sum(W.geom_tri9(g).shape[0] for g in spec["geoms"])
    scene = N.build_render_scene(ctx, spec)
    return scene, spec["camera"], float(spec["merge_radius"]), float(spec["radius_alpha"]), n


def _frac_close(a, b):
    return float((np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5).mean())


def run_batch(N, W, ctx, n_jobs, seed, solo, tune_args=None):
    """n_jobs uneven VCM jobs, each on a scene of its own, rendered as one batch (sgpu_render_vcm_batch)
    and, for comparison, one job at a time with one wavefront per iteration (the plain GPU
    schedule) and, with `solo`, one job at a time with its iterations packed (sgpu_render_vcm)."""
    mix = W.job_mix(seed, n_jobs)
    t0 = time.perf_counter()
    jobs, tris = [], 0
    for k, m in enumerate(mix):
        scene, cam, radius, alpha, n = _build_scene(N, W, ctx, 10000 + k, m["scene_tris"])
        tris += n
        jobs.append({"scene": scene, "camera": cam, "width": m["width"], "height": m["height"],
                     "iterations": m["iterations"], "max_path_length": m["max_path_length"], "base_radius": radius,
                     "radius_alpha": alpha, "seed": m["seed"]})
    log(f"batch: {len(jobs)} scenes, {tris} triangles, built in {time.perf_counter() - t0:.1f} s")
    samples = sum(j["width"] * j["height"] * j["iterations"] for j in jobs)
    tuning = None
    if tune_args is not None:
        # probe: the first 16 jobs with at most 8 iterations each
        sub = [dict(j, iterations=min(j["iterations"], 8)) for j in jobs[:16]]
        sub_fb = [ctx.empty(12 * j["width"] * j["height"]) for j in sub]
        tuning = apply_tuning(tune_args, ctx, lambda c: N.render_vcm_batch(sub, framebuffers=sub_fb))
        for b in sub_fb:
            b.free()
    fb_batch = [ctx.empty(12 * j["width"] * j["height"]) for j in jobs]
    fb_plain = [ctx.empty(12 * j["width"] * j["height"]) for j in jobs]
    # warm-up (untimed): one iteration of the first jobs
    nw = min(len(jobs), 8)
    N.render_vcm_batch([dict(j, iterations=1) for j in jobs[:nw]], framebuffers=fb_batch[:nw])
    N.synchronize()

    e0, e1 = N.Event(), N.Event()
    e0.record()
    st = N.render_vcm_batch(jobs, framebuffers=fb_batch, return_stats=True)[1]
    e1.record()
    N.synchronize()
    batch_ms = e0.elapsed_ms(e1)
    log(f"batch: {samples} samples in {batch_ms:.1f} ms, {st['wavefronts']} wavefronts")

    def one_at_a_time(flags, fbs):
        a, b = N.Event(), N.Event()
        a.record()
        rays = 0
        for j, fb in zip(jobs, fbs):
            _, s1 = N.render_vcm(j["scene"], j["camera"], j["width"], j["height"], j["iterations"],
                                 j["max_path_length"], j["base_radius"], j["radius_alpha"], j["seed"],
                                 return_stats=True, flags=flags, framebuffer=fb)
            rays += s1["rays"]
        b.record()
        N.synchronize()
        return a.elapsed_ms(b), rays

    plain_ms, plain_rays = one_at_a_time(N.VCM_ONE_SLICE_PER_WAVEFRONT, fb_plain)
    log(f"plain: {plain_ms:.1f} ms")
    solo_ms = None
    if solo:
        solo_ms, _ = one_at_a_time(0, fb_plain)
        log(f"one job at a time, iterations packed: {solo_ms:.1f} ms")
    close = [_frac_close(a.download(np.float32, 3 * j["width"] * j["height"]),
                         b.download(np.float32, 3 * j["width"] * j["height"]))
             for a, b, j in zip(fb_batch, fb_plain, jobs)]
    for b in fb_batch + fb_plain:
        b.free()
    for j in jobs:
        j["scene"].release()
    return {
        "mix": mix,
        "jobs": len(jobs),
        "triangles_total": int(tris),
        "samples": int(samples),
        "samples_per_s": samples / (batch_ms / 1e3),
        "ms": batch_ms,
        "mrays_s": st["rays"] / (batch_ms / 1e3) / 1e6,
        "wavefronts": st["wavefronts"],
        "plain_ms": plain_ms,
        "plain_samples_per_s": samples / (plain_ms / 1e3),
        "plain_mrays_s": plain_rays / (plain_ms / 1e3) / 1e6,
        "solo_ms": solo_ms,
        "same_images_min_frac": min(close),
        "tuning": tuning,
    }


def batch_summary(b):
    return {
        "jobs": b["jobs"],
        "samples": b["samples"],
        "triangles_total": b["triangles_total"],
        "samples_per_s": round(b["samples_per_s"], 1),
        "mrays_s": round(b["mrays_s"], 2),
        "ms": round(b["ms"], 1),
        "wavefronts": b["wavefronts"],
        "plain_gpu": {"samples_per_s": round(b["plain_samples_per_s"], 1), "ms": round(b["plain_ms"], 1),
                      "mrays_s": round(b["plain_mrays_s"], 2),
                      "schedule": "one job at a time, one wavefront per iteration"},
        "vs_plain_gpu": round(b["plain_ms"] / b["ms"], 3),
        "vs_solo_gpu": None if b["solo_ms"] is None else round(b["solo_ms"] / b["ms"], 3),
        "solo_gpu": None if b["solo_ms"] is None else {
            "samples_per_s": round(b["samples"] / (b["solo_ms"] / 1e3), 1), "ms": round(b["solo_ms"], 1),
            "schedule": "one job at a time, its iterations packed into wavefronts (sgpu_render_vcm)"},
        "same_images_min_frac": round(b["same_images_min_frac"], 5),
    }


def bench_batch(args):
    from selas_amd import native as N
    from selas_amd import workloads as W

    ctx = new_context(N, args)
    b = run_batch(N, W, ctx, args.batch_jobs, args.batch_seed, args.solo, tune_args=args)
    cpu = None
    if not args.no_cpu:
        threads, host = cpu_threads()
        cpu = dict(cpu_baseline_batch(b["mix"], args.cpu_iterations, threads), **host)
    summ = batch_summary(b)
    out = {
        "metric": METRIC,
        "value": round(b["samples_per_s"], 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": args.warmup,
        "ms_per_step": round(b["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"vcm_batch_{b['jobs']}jobs",
            "jobs": b["jobs"],
            "resolution": "square, side log-uniform in [64, 1024]",
            "iterations": "uniform in 1..64",
            "scenes": "one per job, 64 displaced-grid quad meshes, triangles log-uniform in 10K..2M, textured",
            "triangles_total": b["triangles_total"],
            "samples": b["samples"],
            "max_path_length": 10,
        },
        **{k: v for k, v in summ.items() if k not in ("jobs", "samples", "triangles_total")},
        "tuning": b["tuning"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_batch(mix, iterations, threads):
    """The oracle on one job of the mix (the median scene size), `iterations` iterations."""
    import oracle as O
    from selas_amd import workloads as W

    k = int(np.argsort([m["scene_tris"] for m in mix])[len(mix) // 2])
    m = mix[k]
    spec = W.sized_vcm_scene(10000 + k, m["scene_tris"])
    rs = O.RenderScene(spec)
    w, h = m["width"], m["height"]
    t0 = time.perf_counter()
    O.render_vcm(rs, spec["camera"], w, h, iterations, m["max_path_length"], spec["merge_radius"],
                 spec["radius_alpha"], m["seed"], threads=threads)
    secs = time.perf_counter() - t0
    return {"value": round(w * h * iterations / secs, 1), "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"job {k} of the mix ({w}x{h}, {iterations} iterations, {secs:.2f} s), oracle "
                      f"restatement (scalar C, oracle SAH tree, not Embree)"}


def serve_sweep(N, W, ctx, rates, n_jobs, seed, n_scenes=16, plain=True, budget=200 << 30, wavefront_paths=1 << 22):
    """Poisson request streams at each rate, served by the request loop over resident scenes,
    and by the plain one-job-at-a-time GPU schedule on the same requests."""
    from selas_amd import serving as S

    sizes = W.scene_sizes(seed, n_scenes)
    t0 = time.perf_counter()
    scenes = {}
    for sid in range(n_scenes):
        sc, cam, radius, alpha, _ = _build_scene(N, W, ctx, seed * 1000 + sid, sizes[sid])
        scenes[sid] = (sc, cam, radius, alpha)
    log(f"serve: {n_scenes} resident scenes ({int(sizes.sum())} triangles) built in {time.perf_counter() - t0:.1f} s")
    def no_build(sid):
        raise RuntimeError(f"scene {sid} is not resident")

    served = S.GpuEngine(ctx, no_build, scenes=scenes, wavefront_paths=wavefront_paths)
    base = S.GpuEngine(ctx, no_build, scenes=scenes, plain=True) if plain else None
    out = []
    try:
        # warm-up: one small request through each engine
        warm = [S.Request(0, 0.0, 0, 64, 64, 1, 1)]
        S.RequestLoop(served, budget).run(warm)
        if base:
            S.RequestLoop(base, budget, max_slots=1).run([S.Request(0, 0.0, 0, 64, 64, 1, 1)])
        for rate in rates:
            n = n_jobs(rate) if callable(n_jobs) else n_jobs
            mix = W.job_mix(seed * 7919 + int(rate * 10), n, n_scenes=n_scenes)
            reqs = S.poisson_requests(seed, rate, mix)
            loop = S.RequestLoop(served, budget, max_slots=64)
            loop.run(reqs)
            res = {"rate": rate, "served": S.summarize(reqs, loop.makespan)}
            res["served"].update(steps=loop.steps, peak_in_flight=loop.peak_in_flight,
                                 peak_hbm_bytes=int(loop.peak_used + served.fixed_bytes))
            if base:
                reqs_b = S.poisson_requests(seed, rate, mix)
                lb = S.RequestLoop(base, budget, max_slots=1)
                lb.run(reqs_b)
                res["plain"] = S.summarize(reqs_b, lb.makespan)
            log(f"serve rate {rate}/s: " + json.dumps({k: v for k, v in res.items() if k != "rate"}))
            out.append(res)
    finally:
        served.close()
        if base:
            base.close()
        for sc, *_ in scenes.values():
            sc.release()
    return out


def serving_summary(r):
    out = {"rate_per_s": r["rate"], "served": _round(r["served"])}
    if "plain" in r:
        out["plain_gpu"] = _round(r["plain"])
        out["vs_plain_gpu"] = round(r["served"]["samples_per_s"] / r["plain"]["samples_per_s"], 3)
    return out


def _round(d):
    return {k: (round(v, 2) if isinstance(v, float) else v) for k, v in d.items()}


def bench_serve(args):
    from selas_amd import native as N
    from selas_amd import workloads as W

    rates = [float(x) for x in args.rates.split(",")]
    ctx = new_context(N, args)
    # requests per rate: about 8 s of arrivals, 16..64
    n_jobs = args.serve_jobs or (lambda r: int(min(64, max(16, round(8 * r)))))
    if args.trace_blocks:
        ctx.set_tuning(trace_blocks_per_cu=args.trace_blocks)
    res = serve_sweep(N, W, ctx, rates, n_jobs, args.batch_seed, plain=not args.no_plain,
                      wavefront_paths=args.serve_paths)
    top = max(res, key=lambda r: r["served"]["samples_per_s"])
    out = {
        "metric": METRIC,
        "value": round(top["served"]["samples_per_s"], 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": round(1e3 * sum(r["served"]["makespan_s"] for r in res) / len(res), 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"vcm_serve_poisson_{len(rates)}rates",
            "rates_per_s": rates,
            "jobs_per_rate": [r["served"]["jobs"] + r["served"]["rejected"] for r in res],
            "scenes": "16 resident, 64 displaced-grid quad meshes each, triangles log-uniform in 10K..2M, textured",
            "resolution": "square, side log-uniform in [64, 1024]",
            "iterations": "uniform in 1..64",
            "admission": "arrival order, 64 slots, HBM budget 200 GiB (scenes + framebuffers)",
            "schedule": f"continuous batching, wavefronts of {args.serve_paths} paths, least remaining work first",
        },
        "sweep": [{"rate": r["rate"], "served": _round(r["served"]),
                   "plain": _round(r["plain"]) if "plain" in r else None} for r in res],
        "p50_ms": round(top["served"]["p50_ms"], 1),
        "p99_ms": round(top["served"]["p99_ms"], 1),
        "vs_plain_gpu": (round(top["served"]["samples_per_s"] / top["plain"]["samples_per_s"], 3)
                         if "plain" in top else None),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("vcm", "pt", "traversal", "batch", "serve", "instances", "deferred",
                                           "curves"),
                    default="vcm")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (vcm: whole 32-iteration jobs, default 5; traversal: passes, default 10)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=512, help="vcm: image width and height")
    ap.add_argument("--cpu-iterations", type=int, default=32,
                    help="vcm: iterations in the CPU baseline sample (32: the whole job)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rays", type=int, default=4194304, help="rays in the CPU baseline sample")
    ap.add_argument("--batch-jobs", type=int, default=64, help="batch: jobs in the mix; serve: -")
    ap.add_argument("--batch-seed", type=int, default=7, help="batch/serve: seed of the job mix and scenes")
    ap.add_argument("--solo", action="store_true", help="batch workload: also time the packed one-job-at-a-time schedule")
    ap.add_argument("--no-solo", action="store_true", help="vcm: skip the packed one-job-at-a-time schedule")
    ap.add_argument("--rates", default="1,2,5,10,20,50,100,200", help="serve: Poisson arrival rates (jobs/s)")
    ap.add_argument("--serve-rate", type=float, default=50.0,
                    help="vcm: arrival rate of the serving leg (requests/s; 50 offers more than the GPU serves)")
    ap.add_argument("--serve-jobs", type=int, default=None,
                    help="requests per rate (default: 32 in the serving leg; 8 s of arrivals, 16..64, in serve)")
    ap.add_argument("--no-plain", action="store_true", help="serve: skip the plain one-job-at-a-time runs")
    ap.add_argument("--serve-paths", type=int, default=1 << 22, help="serving: paths per wavefront of the server")
    ap.add_argument("--no-batch", action="store_true", help="vcm: skip the batched leg")
    ap.add_argument("--no-serve", action="store_true", help="vcm: skip the serving leg")
    ap.add_argument("--no-tune", action="store_true", help="use the default launch shape instead of autotuning")
    ap.add_argument("--retune", action="store_true",
                    help="run the autotuner even where selas_amd/tuning.json holds a result")
    ap.add_argument("--no-phase-times", action="store_true",
                    help="vcm: time the render without per-launch events (lets it replay captured graphs)")
    ap.add_argument("--wavefront-paths", type=int, default=None, help="VCM paths per wavefront (skips autotuning)")
    ap.add_argument("--trace-blocks", type=int, default=None, help="traversal blocks per CU (skips autotuning)")
    ap.add_argument("--order", type=int, default=None, choices=(0, 1, 2, 3),
                    help="VCM work order (sgpu_tuning.vcm_order bits: 1 merge in spatial bins, 2 shade by material)")
    ap.add_argument("--bvh-builder", type=int, default=None, choices=(0, 1),
                    help="binary tree under the 8-wide one: 0 linear BVH (default), 1 locally-ordered clustering")
    ap.add_argument("--merge-lanes", type=int, default=None, choices=(8, 16, 32),
                    help="VCM lanes per vertex-merging range search (sgpu_tuning.vcm_merge_lanes)")
    ap.add_argument("--trace-tree", type=int, default=None, choices=(0, 1),
                    help="traversal tree: 0 binary, 1 8-wide compressed (default: the context's)")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the triangles (1.0 = the workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run sgpu_device_copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        from selas_amd import native as N
        ctx = new_context(N, args)
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
    elif args.workload == "pt":
        bench_pt(args)
    elif args.workload == "batch":
        bench_batch(args)
    elif args.workload == "serve":
        bench_serve(args)
    elif args.workload == "deferred":
        bench_deferred(args)
    elif args.workload == "instances":
        if args.steps is None:
            args.steps = 10
        bench_instances(args)
    elif args.workload == "curves":
        if args.steps is None:
            args.steps = 10
        bench_curves(args)
    else:
        if args.steps is None:
            args.steps = 10
        bench_traversal(args)


# Algorithmic bytes of the path tracer's shading kernel (k_pt_shade): per closest-hit ray read 60 B
# (queue entry 4, ray origin and direction 24, hit geometry, distance and normal 20, throughput 12);
# per hit 36 B more (random state read and written 16, path flags read and written 8, throughput
# written 12); per shadow ray written 48 B (ray 32, path 4, contribution 12); per continuation ray
# written 36 B (ray 32, queue entry 4); per miss 24 B (radiance read and written).
PT_B_RAY, PT_B_HIT, PT_B_SHADOW, PT_B_CONT, PT_B_MISS = 60, 36, 48, 36, 24


def bench_pt(args):
    from selas_amd import native as N
    from selas_amd import workloads as W

    steps = 10 if args.steps is None else args.steps
    w, spp, L, seed = 256, 16, 8, 1
    sc = W.room_scene(1)
    n_tris = sum(W.geom_tri9(g).shape[0] for g in sc["geoms"])
    ctx = new_context(N, args)
    scene = N.build_render_scene(ctx, sc)
    cam = sc["camera"]
    fb = ctx.empty(2 * 3 * w * w * 4)

    def image(flags=0, stats=True):
        return N.render_pt(scene, cam, w, w, spp, L, seed, flags=flags, return_stats=stats, framebuffer=fb)

    _, cnt = image(N.PT_COUNT_TRAVERSAL)
    node_bytes = 128 if ctx.tuning()["trace_tree"] == 1 else 64
    nodes_per_ray = cnt["nodes_visited"] / cnt["closest_rays"]
    tris_per_ray = cnt["tris_tested"] / cnt["closest_rays"]
    bytes_per_ray = 48 + nodes_per_ray * node_bytes + tris_per_ray * 48
    cont = cnt["closest_rays"] - w * w * spp
    shade_bytes = (cnt["closest_rays"] * PT_B_RAY + cnt["hits"] * PT_B_HIT + cnt["shadow_rays"] * PT_B_SHADOW +
                   cont * PT_B_CONT + (cnt["closest_rays"] - cnt["hits"]) * PT_B_MISS)
    for _ in range(args.warmup):
        image(stats=False)
    N.synchronize()
    tot = None
    e0, e1 = N.Event(), N.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        _, st = image()
        tot = st if tot is None else {k: tot[k] + st[k] for k in tot}
    e1.record()
    N.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    img = fb.download(np.float32, 2 * 3 * w * w).reshape(2, w, w, 3)
    fb.free()
    scene.check()
    samples = w * w * spp * steps
    value = samples / (ms / 1e3)
    b = max(1, tot["bounces"])
    shade = {"bound": "hbm", "achieved": round(shade_bytes * steps / (tot["shade_ms"] / 1e3) / 1e9, 1),
             "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": load_traffic("traffic_pt_shade_room_256x256_16spp.json"),
             "kernel": "k_pt_shade<false> (all launches of the timed images)", "launches": tot["bounces"],
             "avg_launch_ms": round(tot["shade_ms"] / b, 4), "bytes_per_launch": round(shade_bytes * steps / b),
             "bytes_rule": f"{PT_B_RAY} B per closest-hit ray + {PT_B_HIT} B per hit + {PT_B_SHADOW} B per shadow "
                           f"ray + {PT_B_CONT} B per continuation ray + {PT_B_MISS} B per miss"}
    closest = {"bound": "hbm", "achieved": round(tot["closest_rays"] * bytes_per_ray / (tot["closest_ms"] / 1e3) / 1e9, 1),
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
               "kernel": "k_trace8<false,false,false> (closest hit, all launches of the timed images)",
               "launches": tot["bounces"], "avg_launch_ms": round(tot["closest_ms"] / b, 4),
               "bytes_per_ray": round(bytes_per_ray, 1), "nodes_per_ray": round(nodes_per_ray, 3),
               "tris_per_ray": round(tris_per_ray, 3)}
    for r in (shade, closest):
        r["frac"] = round(r["achieved"] / HBM_PEAK_GBS, 4)
    dominant, other = (shade, closest) if tot["shade_ms"] >= tot["closest_ms"] else (closest, shade)
    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads, host = cpu_threads()
        rs = O.RenderScene(sc)
        t1 = time.perf_counter()
        cimg = O.render_pt(rs, cam, w, w, spp, L, seed, threads=threads)
        secs = time.perf_counter() - t1
        cpu = {"value": round(w * w * spp / secs, 1), "unit": "samples/s", "cores": threads, "kind": "port",
               "sample": f"the whole image ({w}x{w} pixels x {spp} paths, seed {seed}; {secs:.2f} s), oracle "
                         "restatement of EvaluatePath (scalar C, oracle SAH tree, not Embree)",
               "pixels_within_1e-3": round(_frac_close(img, cimg), 5),
               "rmse_gpu_vs_cpu_rel": round(float(np.sqrt(((img[1] - cimg[1]) ** 2).mean()) /
                                                  max(1e-12, cimg[1].mean())), 6), **host}
    scene.release()
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
        "config": {"workload": f"pt_room_{w}x{w}_{spp}spp_L{L}", "step": "one whole image, rendered from scratch",
                   "triangles": int(n_tris), "width": w, "height": w, "spp": spp, "max_path_length": L,
                   "scene": "closed room of 12 quads + icosphere (20,480 tris), one quad light of radiance 15"},
        "mrays_s": round((tot["closest_rays"] + tot["shadow_rays"]) / (ms / 1e3) / 1e6, 2),
        "closest_rays_per_image": tot["closest_rays"] // steps,
        "shadow_rays_per_image": tot["shadow_rays"] // steps,
        "phase_ms_per_image": {k: round(tot[k] / steps, 3) for k in ("closest_ms", "shadow_ms", "shade_ms", "other_ms")},
        "wall_s_timed_region": round(wall, 3),
        "image_mean": round(float(img[1].mean()), 5),
        "roofline": dominant,
        "roofline_other": other,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_traversal(args):
    from selas_amd import native as N
    from selas_amd import workloads as W

    log("building workload")
    t9 = W.traversal_scene(args.scale)
    rays = W.traversal_rays(t9)
    n = int(rays["org_x"].shape[0])

    ctx = new_context(N, args)
    if args.trace_tree is not None:
        ctx.set_tuning(trace_tree=args.trace_tree)
    if args.trace_blocks:
        ctx.set_tuning(trace_blocks_per_cu=args.trace_blocks)
    wide = ctx.tuning()["trace_tree"] == 1
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
        threads, host = cpu_threads()
        cpu = dict(cpu_baseline_traversal(t9, rays, args.cpu_rays, threads), **host)

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
            "bvh": ("device LBVH collapsed to 8-wide compressed nodes (128 B), triangles as children (48 B)"
                    if wide else "device LBVH, binary nodes (64 B), 1 triangle per leaf (48 B)"),
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
            "kernel": "k_trace8<false,false,false> (closest hit)" if wide else "k_trace<false,false,false> (closest hit)",
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def bench_curves(args):
    """Ray casts against round B-spline curves: a field of 147,456 grass blades of 10 control points
    (7 segments each, 1,032,192 segments; radius 3 mm tapering to 1 mm over blades 10-25 cm tall, on
    a 2 m x 2 m ground of two triangles), rays from a camera above the field (1024 x 1024, 2 samples
    per pixel) plus as many cosine rays up from the ground: closest hit (triangle pass, then the curve
    pass) and any hit."""
    from selas_amd import native as N
    from selas_amd import workloads as W

    log("building curve workload")
    rng = np.random.default_rng(77)
    ns, ncp = 147456, 10
    root = np.stack([rng.uniform(-1, 1, ns), np.zeros(ns), rng.uniform(-1, 1, ns)], 1)
    height = rng.uniform(0.10, 0.25, ns)
    d = np.tile([0.0, 1.0, 0.0], (ns, 1)) + 0.3 * rng.normal(size=(ns, 3)) * [1, 0, 1]
    pts = [root]
    for k in range(ncp - 1):
        d = d + 0.15 * rng.normal(size=(ns, 3)) * [1, 0.2, 1]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts.append(pts[-1] + d * (height / (ncp - 1))[:, None])
    pts = np.stack(pts, 1)
    rad = np.linspace(0.003, 0.001, ncp)[None, :].repeat(ns, 0)
    verts = np.concatenate([pts, rad[:, :, None]], 2).reshape(-1, 4).astype(np.float32)
    idx = (np.arange(ns)[:, None] * ncp + np.arange(ncp - 3)[None, :]).reshape(-1).astype(np.uint32)
    ground = np.array([[-1.2, 0, -1.2, 1.2, 0, -1.2, 1.2, 0, 1.2], [-1.2, 0, -1.2, 1.2, 0, 1.2, -1.2, 0, 1.2]],
                      np.float32)
    cam = N.gen_camera_rays((0.0, 1.2, -1.9), (0.0, 0.05, 0.0), (0.0, 1.0, 0.0), 50.0, 1024, 1024, 2)
    nb = cam["org_x"].shape[0]
    o = np.stack([rng.uniform(-1, 1, nb), np.full(nb, 1e-3), rng.uniform(-1, 1, nb)], 1)
    u1, u2 = rng.random(nb), rng.random(nb)
    r_, phi = np.sqrt(u1), 2 * np.pi * u2
    dd = np.stack([r_ * np.cos(phi), np.sqrt(1 - u1), r_ * np.sin(phi)], 1)
    bounce = {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": dd[:, 0], "dir_y": dd[:, 1],
              "dir_z": dd[:, 2], "tnear": np.zeros(nb), "tfar": np.full(nb, 1e30)}
    rays = {k: np.concatenate([cam[k], bounce[k]]).astype(np.float32) for k in N.RAY_FIELDS}
    n = int(rays["org_x"].shape[0])

    ctx = new_context(N, args)
    scene = N.Scene(ctx)
    scene.add_soup(ground)
    tb = time.perf_counter()
    scene.add_curves(verts, idx)
    scene.commit()
    build_ms = (time.perf_counter() - tb) * 1e3
    dr = N.DeviceRays(ctx, rays)
    dh = N.DeviceHits(ctx, n, with_ng=True)
    scene.intersect(dr, dh)
    h = dh.to_host()
    curve_hit = float((h["geom_id"] == 1).mean())
    tfar, _ = W.anyhit_tfar(h["t"], h["prim_id"])
    rays_any = dict(rays)
    rays_any["tfar"] = tfar
    dra = N.DeviceRays(ctx, rays_any)
    occ = ctx.empty(4 * n)
    st = scene.intersect_stats(dr)
    for _ in range(args.warmup):
        scene.intersect(dr, dh)
        scene.occluded(dra, occ)
    N.synchronize()
    ev = [(N.Event(), N.Event(), N.Event()) for _ in range(args.steps)]
    for a, b, c in ev:
        a.record(None)
        scene.intersect(dr, dh)
        b.record(None)
        scene.occluded(dra, occ)
        c.record(None)
    N.synchronize()
    scene.check()
    mi = sum(a.elapsed_ms(b) for a, b, _ in ev) / args.steps
    mo = sum(b.elapsed_ms(c) for _, b, c in ev) / args.steps
    nodes, segs = st["curve_nodes_visited"] / n, st["curve_segments_tested"] / n
    hit_rate = float((h["prim_id"] != N.INVALID_ID).mean())
    # per ray: ray 32 B + hit distance 4 B read, nodes x 64 B, segments x 80 B, 32 B hit record written
    bytes_per_ray = 36 + nodes * 64 + segs * 80 + 32 * curve_hit
    achieved = n * bytes_per_ray / (mi / 1e3) / 1e9
    out = {
        "metric": "curve closest + any-hit Mrays/s",
        "value": round(2 * n / ((mi + mo) / 1e3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mi + mo, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "curves_grass_1Mseg_4Mray", "segments": int(idx.shape[0]), "strands": ns,
                   "control_points": int(verts.shape[0]), "rays": n},
        "closest_hit_mrays_s": round(n / (mi / 1e3) / 1e6, 2),
        "any_hit_mrays_s": round(n / (mo / 1e3) / 1e6, 2),
        "hit_rate": round(hit_rate, 4),
        "curve_hit_rate": round(curve_hit, 4),
        "curve_nodes_per_ray": round(nodes, 3),
        "curve_segments_per_ray": round(segs, 3),
        "build_ms": round(build_ms, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "closest hit: k_trace8 over the ground, then k_curves<false,false,false>",
                     "bytes_per_ray": round(bytes_per_ray, 1)},
        "cpu_baseline": None,
    }
    occ.free()
    scene.release()
    print(json.dumps(out), flush=True)


def bench_instances(args):
    """Ray casts through two instance levels: one 8192-triangle cluster placed 16 times in a subscene,
    the subscene placed 16 times in the scene (256 composed instances, 2,097,152 triangles in world
    space), against the same triangles moved to world space as one scene of geometries, on the
    traversal workload's rays (8 camera samples per pixel at 1024 x 1024 plus as many cosine
    bounce rays from the world triangles)."""
    from selas_amd import native as N
    from selas_amd import workloads as W

    log("building instanced workload")
    rng = np.random.default_rng(2024)
    model = N.gen_cluster_soup(2024, 1, 8192, 0.02, 1e-4, 1e-2)
    model = (model.reshape(-1, 3) - model.reshape(-1, 3).mean(0)).reshape(-1, 9).astype(np.float32)
    sub = [np.vstack([np.hstack([_rotation(rng), rng.uniform(-0.15, 0.15, (3, 1))]), [0, 0, 0, 1]]) for _ in range(16)]
    top = [np.vstack([np.hstack([_rotation(rng), rng.uniform(0.2, 0.8, (3, 1))]), [0, 0, 0, 1]]) for _ in range(16)]
    world = []
    p = model.reshape(-1, 3).astype(np.float64)
    for a in top:
        for b in sub:
            m = a @ b
            world.append((p @ m[:3, :3].T + m[:3, 3]).astype(np.float32).reshape(-1, 9))
    world = np.concatenate(world)
    rays = W.traversal_rays(world)
    n = int(rays["org_x"].shape[0])

    ctx = new_context(N, args)
    if args.trace_blocks:
        ctx.set_tuning(trace_blocks_per_cu=args.trace_blocks)
    dm = N.Scene(ctx)
    dm.add_soup(model)
    dm.commit()
    ds = N.Scene(ctx)
    for b in sub:
        ds.add_instance(dm, b)
    ds.commit()
    inst = N.Scene(ctx)
    tb = time.perf_counter()
    for a in top:
        inst.add_instance(ds, a)
    inst.commit()
    build_ms = (time.perf_counter() - tb) * 1e3
    flat = N.Scene(ctx)
    flat.add_soup(world)
    flat.commit()
    log(f"instanced: {inst.info()}, device bytes {inst.device_bytes() + ds.device_bytes() + dm.device_bytes()}; "
        f"flat: {flat.info()}, device bytes {flat.device_bytes()}")

    dr = N.DeviceRays(ctx, rays)
    dh = N.DeviceHits(ctx, n, with_ng=False)
    res = {}
    for name, scene in (("instanced", inst), ("flat", flat)):
        scene.intersect(dr, dh)
        h = dh.to_host()
        tfar, _ = W.anyhit_tfar(h["t"], h["prim_id"])
        rays_any = dict(rays)
        rays_any["tfar"] = tfar
        dra = N.DeviceRays(ctx, rays_any)
        occ = ctx.empty(4 * n)
        st = scene.intersect_stats(dr)
        for _ in range(args.warmup):
            scene.intersect(dr, dh)
            scene.occluded(dra, occ)
        N.synchronize()
        ev = [(N.Event(), N.Event(), N.Event()) for _ in range(args.steps)]
        for a, b, c in ev:
            a.record(None)
            scene.intersect(dr, dh)
            b.record(None)
            scene.occluded(dra, occ)
            c.record(None)
        N.synchronize()
        scene.check()
        mi = sum(a.elapsed_ms(b) for a, b, _ in ev) / args.steps
        mo = sum(b.elapsed_ms(c) for _, b, c in ev) / args.steps
        res[name] = {"closest_hit_mrays_s": round(n / (mi / 1e3) / 1e6, 2), "any_hit_mrays_s": round(n / (mo / 1e3) / 1e6, 2),
                     "ms_per_step": round(mi + mo, 3), "nodes_per_ray": round(st["nodes_visited"] / n, 3),
                     "tris_per_ray": round(st["tris_tested"] / n, 3), "hit_rate": round(float((h["prim_id"] != N.INVALID_ID).mean()), 4)}
        log(f"{name}: {res[name]}")
        occ.free()
    r = res["instanced"]
    bytes_per_ray = 48 + r["nodes_per_ray"] * 128 + r["tris_per_ray"] * 48
    ms_i = 1e3 * n / (r["closest_hit_mrays_s"] * 1e6)
    achieved = n * bytes_per_ray / (ms_i / 1e3) / 1e9
    out = {
        "metric": "instanced closest + any-hit Mrays/s",
        "value": round(2 * n / (r["ms_per_step"] / 1e3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "instanced_traversal_16x16x8k_16Mray", "composed_instances": 256, "levels": 2,
                   "triangles_stored": int(model.shape[0]), "triangles_world": int(world.shape[0]), "rays": n},
        "instanced": res["instanced"],
        "flat_same_triangles": res["flat"],
        "top_tree_build_ms": round(build_ms, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace8i<false,false,false> (closest hit; 64-B instance records not counted)",
                     "bytes_per_ray": round(bytes_per_ray, 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_deferred(args):
    """The deferred path tracer (sgpu_render_deferred_pt, DeferredPathTracer::GenerateImage) on the
    VCM workload's 1M-triangle textured scene with an image-based sky (256 x 512): 512 x 512 pixels,
    4 x 4 CMJ samples per pixel (SamplesPerPixelX_ x SamplesPerPixelY_), paths bounded at 10
    segments; one step = one whole image (4,194,304 paths). Timed without the hit sort (the
    library's default, `value`) and with it (Selas's ordering of shading); the two images are
    compared bit for bit."""
    from selas_amd import native as N
    from selas_amd import workloads as W

    w, sx, sy, L = args.width, 4, 4, 10
    steps = 5 if args.steps is None else args.steps
    log("building deferred workload scene")
    sc = W.vcm_scene()
    sc["ibl"] = {"rgb": W.sky_map(256, 512), "rotation": 0.0, "exposure": 1.0}
    n_tris = sum(W.geom_tri9(g).shape[0] for g in sc["geoms"])
    ctx = new_context(N, args)
    if args.trace_blocks:
        ctx.set_tuning(trace_blocks_per_cu=args.trace_blocks)
    scene = N.build_render_scene(ctx, sc)
    cam = sc["camera"]
    fb = ctx.empty(3 * w * w * 4)
    res, imgs = {}, {}
    modes = (("unsorted", 0), ("sorted", N.DPT_SORT_HITS), ("rays_sorted", N.DPT_SORT_RAYS),
             ("hits_and_rays_sorted", N.DPT_SORT_HITS | N.DPT_SORT_RAYS))
    for name, flags in modes:
        for _ in range(args.warmup):
            N.render_deferred_pt(scene, cam, w, w, sx, sy, L, 1, flags=flags, framebuffer=fb)
        N.synchronize()
        e0, e1 = N.Event(), N.Event()
        e0.record()
        for _ in range(steps):
            N.render_deferred_pt(scene, cam, w, w, sx, sy, L, 1, flags=flags, framebuffer=fb)
        e1.record()
        N.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        imgs[name] = fb.download(np.float32, 3 * w * w)
        res[name] = {"samples_per_s": round(w * w * sx * sy / (ms / 1e3), 1), "ms_per_image": round(ms, 3)}
        log(f"deferred {name}: {res[name]}")
    scene.check()
    fb.free()
    identical = all(np.array_equal(imgs[m].view(np.uint32), imgs["unsorted"].view(np.uint32)) for m, _ in modes)
    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads, host = cpu_threads()
        log("cpu baseline: building the oracle SAH tree")
        rs = O.RenderScene(sc)
        t0 = time.perf_counter()
        img = O.render_dpt(rs, cam, w, w, sx, sy, L, 1, threads=threads)
        secs = time.perf_counter() - t0
        cpu = {"value": round(w * w * sx * sy / secs, 1), "unit": "samples/s", "cores": threads, "kind": "port",
               "sample": f"the whole image ({w}x{w} pixels x {sx}x{sy} samples, same seed; {secs:.2f} s), oracle "
                         "restatement (scalar C, oracle SAH tree, not Embree)",
               "image_mean": round(float(img.mean()), 5),
               "mean_rel_diff_vs_gpu": round(float(abs(img.mean() / imgs["unsorted"].mean() - 1)), 6), **host}
    scene.release()
    out = {
        "metric": "deferred path tracer samples/s at 1M tris",
        "value": res["unsorted"]["samples_per_s"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": res["unsorted"]["ms_per_image"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"deferred_pt_{w}x{w}_{sx}x{sy}spp_L{L}_1Mtri_sky", "triangles": int(n_tris),
                   "width": w, "height": w, "spp": [sx, sy], "max_path_length": L,
                   "sky": "256 x 512 equirectangular image-based light"},
        "unsorted": res["unsorted"],
        "sorted": res["sorted"],
        "rays_sorted": res["rays_sorted"],
        "hits_and_rays_sorted": res["hits_and_rays_sorted"],
        "sorted_equals_unsorted": identical,
        "image_mean": round(float(imgs["unsorted"].mean()), 5),
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
