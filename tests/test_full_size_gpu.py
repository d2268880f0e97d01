"""Parity at the full sizes the benchmarks run, through the C ABI.

* The VCM workload's job exactly as bench.py times it: 1,048,576 textured triangles, 512 x 512,
  32 iterations, path length 10, with the launch shape stored for gfx950 (selas_amd/tuning.json),
  against the oracle's CPU restatement of the same job (same seeds), per pixel.
* The batched mix: 64 uneven jobs (64^2..1024^2 pixels, 1..64 iterations, each on its own scene of
  10K..2M triangles) rendered as one batch; every image equals the job rendered alone
  (sgpu_render_vcm), and the three jobs of least work also equal the oracle.
* The served stream above capacity: 32 Poisson requests at 50 per second over 16 resident scenes
  (more work arrives than the GPU finishes, so requests queue and overlap); every result equals the
  job rendered alone.

Bars, as in the other VCM tests: per pixel |a - b| <= 1e-3 |b| + 1e-5 for at least 99 % of the
values against the oracle (99.9 % between two device schedules, whose only order-dependent sums
are the light-tracing splat atomics); relative RMSE against the oracle <= 2e-3 of the mean.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from selas_amd import native as N
from selas_amd import serving as S
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _frac(a, b):
    return float((np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5).mean())


def _rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(float(b.mean()), 1e-12))


def _threads():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _stored_tuning(ctx, workload):
    arch = ctx.device_info()["name"].split(":")[0]
    with open(os.path.join(ROOT, "selas_amd", "tuning.json")) as f:
        t = json.load(f).get(f"{arch}|{workload}")
    if t:
        ctx.set_tuning(**{k: v for k, v in t.items() if k in ctx.tuning()})
    return t


def test_vcm_workload_job_matches_oracle():
    sc = W.vcm_scene()
    n_tris = sum(W.geom_tri9(g).shape[0] for g in sc["geoms"])
    assert n_tris == 1048576
    w, its, L = 512, 32, sc["max_path_length"]
    ctx = N.Context(0)
    scene = None
    try:
        tuning = _stored_tuning(ctx, f"vcm_{w}x{w}_{its}it_L{L}_{n_tris}tri")
        print(f"vcm workload job {w}x{w} x {its} iterations, tuning {tuning}", flush=True)
        scene = N.build_render_scene(ctx, sc)
        gpu, st = N.render_vcm(scene, sc["camera"], w, w, its, L, sc["merge_radius"], sc["radius_alpha"], 1,
                               return_stats=True)
        scene.check()
        assert st["samples"] == w * w * its
        print(f"gpu done ({st['wavefronts']} wavefronts); oracle on {_threads()} threads", flush=True)
        rs = O.RenderScene(sc)
        cpu, cst = O.render_vcm(rs, sc["camera"], w, w, its, L, sc["merge_radius"], sc["radius_alpha"], 1,
                                threads=_threads())
        assert np.isfinite(gpu).all() and cpu.mean() > 0
        frac, rel = _frac(gpu, cpu), _rel_rmse(gpu, cpu)
        print(f"pixels within 1e-3: {frac:.5f}, relative RMSE {rel:.2e}", flush=True)
        assert frac >= 0.99, frac
        assert rel <= 2e-3, rel
        assert abs(st["light_vertices"] / its / cst["light_vertices_per_iteration"] - 1) < 1e-3
    finally:
        if scene is not None:
            scene.release()
        ctx.close()


def test_batched_mix_equals_solo_and_oracle(gpu_ctx):
    mix = W.job_mix(7, 64)
    jobs = []
    try:
        for k, m in enumerate(mix):
            spec = W.sized_vcm_scene(10000 + k, m["scene_tris"])
            scene = N.build_render_scene(gpu_ctx, spec)
            jobs.append({"scene": scene, "camera": spec["camera"], "width": m["width"], "height": m["height"],
                         "iterations": m["iterations"], "max_path_length": m["max_path_length"],
                         "base_radius": float(spec["merge_radius"]), "radius_alpha": float(spec["radius_alpha"]),
                         "seed": m["seed"]})
            del spec  # host arrays dropped; the oracle's jobs rebuild theirs below
        print(f"batch: 64 scenes built, {sum(m['width'] * m['height'] * m['iterations'] for m in mix)} samples",
              flush=True)
        fbs = [gpu_ctx.empty(12 * j["width"] * j["height"]) for j in jobs]
        try:
            st = N.render_vcm_batch(jobs, framebuffers=fbs, return_stats=True)[1]
            assert st["samples"] == sum(j["width"] * j["height"] * j["iterations"] for j in jobs)
            print(f"batch rendered in {st['wavefronts']} wavefronts; comparing with solo renders", flush=True)
            worst = 1.0
            for k, (j, fb) in enumerate(zip(jobs, fbs)):
                img = fb.download(np.float32, 3 * j["width"] * j["height"]).reshape(j["height"], j["width"], 3)
                alone = N.render_vcm(j["scene"], j["camera"], j["width"], j["height"], j["iterations"],
                                     j["max_path_length"], j["base_radius"], j["radius_alpha"], j["seed"])
                assert np.isfinite(img).all()
                f = _frac(img, alone)
                worst = min(worst, f)
                assert f >= 0.999, (k, f)
            print(f"solo: worst job {worst:.5f}", flush=True)
        finally:
            for b in fbs:
                b.free()
        # the jobs of least work against the oracle
        work = [m["width"] * m["height"] * m["iterations"] * m["scene_tris"] ** 0.3 for m in mix]
        for k in np.argsort(work)[:3]:
            k = int(k)
            m, j = mix[k], jobs[k]
            spec = W.sized_vcm_scene(10000 + k, m["scene_tris"])
            print(f"oracle: job {k} ({m['width']}^2 x {m['iterations']}, {m['scene_tris']:.0f} tris)", flush=True)
            rs = O.RenderScene(spec)
            cpu, _ = O.render_vcm(rs, spec["camera"], m["width"], m["height"], m["iterations"], m["max_path_length"],
                                  spec["merge_radius"], spec["radius_alpha"], m["seed"], threads=_threads())
            alone = N.render_vcm(j["scene"], j["camera"], j["width"], j["height"], j["iterations"],
                                 j["max_path_length"], j["base_radius"], j["radius_alpha"], j["seed"])
            assert _frac(alone, cpu) >= 0.99, k
            assert _rel_rmse(alone, cpu) <= 2e-3, k
    finally:
        for j in jobs:
            j["scene"].release()


def test_served_stream_above_capacity_equals_solo(gpu_ctx):
    seed, n_scenes = 7, 16
    sizes = W.scene_sizes(seed, n_scenes)
    scenes = {}
    try:
        for sid in range(n_scenes):
            spec = W.sized_vcm_scene(seed * 1000 + sid, sizes[sid])
            scenes[sid] = (N.build_render_scene(gpu_ctx, spec), spec["camera"], float(spec["merge_radius"]),
                           float(spec["radius_alpha"]))
        print(f"serve: {n_scenes} resident scenes built", flush=True)

        def no_build(sid):
            raise RuntimeError(f"scene {sid} is not resident")

        eng = S.GpuEngine(gpu_ctx, no_build, scenes=scenes, wavefront_paths=1 << 22)
        try:
            mix = W.job_mix(seed * 7919 + 500, 32, n_scenes=n_scenes)
            reqs = S.poisson_requests(seed, 50.0, mix)
            loop = S.RequestLoop(eng, 200 << 30, max_slots=64, keep_images=True)
            loop.run(reqs)
            summ = S.summarize(reqs, loop.makespan)
            print(f"served: {json.dumps({k: round(v, 2) if isinstance(v, float) else v for k, v in summ.items()})}",
                  flush=True)
            assert all(r.status == "done" for r in reqs)
            assert loop.peak_in_flight > 1
            # above capacity: the stream took longer than its arrivals, so requests waited
            assert loop.makespan > reqs[-1].arrival
            assert summ["wait_p99_ms"] > 0
            for r in reqs:
                sc, cam, radius, alpha = scenes[r.scene]
                alone = N.render_vcm(sc, cam, r.width, r.height, r.iterations, r.max_path_length, radius, alpha,
                                     r.seed)
                assert _frac(r.image, alone) >= 0.999, r.id
        finally:
            eng.close()
    finally:
        for sc, *_ in scenes.values():
            sc.release()
