"""Batched VCM (sgpu_render_vcm_batch): many jobs of uneven size on several scenes in one schedule.

Every slice (one iteration of one job) carries its own camera, radius schedule, random streams,
photon grid and frame, so a job's image does not depend on what it is batched with: each batched
image must match the same job rendered alone, per pixel, to the bar of the other VCM tests
(splat atomics are the only order-dependent sums); the same holds for the un-batched schedule
(one wavefront per iteration).
"""
import numpy as np
import pytest

from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _close(a, b):
    return (np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5).mean()


@pytest.fixture(scope="module")
def scenes(gpu_ctx):
    specs = [W.room_scene(1), W.vcm_scene(seed=5, n_meshes=12, nu=16, nv=8),
             W.vcm_scene(seed=9, n_meshes=6, nu=8, nv=8, textures=False)]
    built = [N.build_render_scene(gpu_ctx, sc) for sc in specs]
    yield list(zip(specs, built))
    for b in built:
        b.release()


def _jobs(scenes):
    rng = np.random.default_rng(4)
    jobs = []
    for k in range(9):
        sc, scene = scenes[k % len(scenes)]
        w = int(rng.choice([16, 24, 40, 64, 100]))
        h = int(rng.choice([16, 30, 48]))
        jobs.append({"scene": scene, "camera": sc["camera"], "width": w, "height": h,
                     "iterations": int(rng.integers(1, 6)), "max_path_length": int(rng.choice([3, 6, 10])),
                     "base_radius": float(sc.get("merge_radius", 0.005)), "radius_alpha": 0.75,
                     "seed": int(rng.integers(0, 1 << 40)), "iteration_base": int(rng.integers(0, 3))})
    return jobs


def test_batch_equals_jobs_alone(scenes):
    jobs = _jobs(scenes)
    batch, st = N.render_vcm_batch(jobs, return_stats=True)
    assert st["samples"] == sum(j["width"] * j["height"] * j["iterations"] for j in jobs)
    assert st["wavefronts"] == 1  # all of these fit one wavefront
    for img, j in zip(batch, jobs):
        alone = N.render_vcm(j["scene"], j["camera"], j["width"], j["height"], j["iterations"],
                             j["max_path_length"], j["base_radius"], j["radius_alpha"], j["seed"],
                             iteration_base=j["iteration_base"])
        assert np.isfinite(img).all()
        assert _close(img, alone) >= 0.999, (j["width"], j["height"], j["iterations"])


def test_unbatched_schedule_gives_the_same_images(scenes):
    jobs = _jobs(scenes)[:4]
    a = N.render_vcm_batch(jobs)
    b, st = N.render_vcm_batch(jobs, flags=N.VCM_ONE_SLICE_PER_WAVEFRONT, return_stats=True)
    assert st["wavefronts"] == sum(j["iterations"] for j in jobs)
    for x, y in zip(a, b):
        assert _close(x, y) >= 0.999


def test_batch_errors(scenes, gpu_ctx):
    jobs = _jobs(scenes)[:2]
    fb = [gpu_ctx.empty(3 * j["width"] * j["height"] * 4) for j in jobs]
    n = len(jobs)
    arr = (N.VcmJob * n)()
    sc = (N.C.c_void_p * n)()
    for k, j in enumerate(jobs):
        sc[k] = j["scene"].handle
        arr[k] = N.VcmJob(N.Camera.make(j["camera"], j["width"], j["height"]), 1, 0, 5, 0.01, 0.75, 0, 1)
    try:
        same = (N.C.c_void_p * n)(fb[0].ptr, fb[0].ptr)
        assert N.lib.sgpu_render_vcm_batch(sc, arr, same, n, 0, None, None) == N.SGPU_E_INVALID
        ok = (N.C.c_void_p * n)(fb[0].ptr, fb[1].ptr)
        assert N.lib.sgpu_render_vcm_batch(sc, arr, ok, n, 8, None, None) == N.SGPU_E_INVALID
        arr[1].flags = 1
        assert N.lib.sgpu_render_vcm_batch(sc, arr, ok, n, 0, None, None) == N.SGPU_E_INVALID
        arr[1].flags = 0
        assert N.lib.sgpu_render_vcm_batch(sc, arr, ok, 0, 0, None, None) == N.SGPU_E_INVALID
        assert N.lib.sgpu_render_vcm_batch(sc, arr, ok, n, 0, None, None) == N.SGPU_OK
    finally:
        for b in fb:
            b.free()
