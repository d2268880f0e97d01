"""The request loop on the GPU (selas_amd/serving.py GpuEngine): requests arriving while others are
in flight share wavefronts, and each result equals the job rendered alone (sgpu_render_vcm) per
pixel, to the bar of the other VCM tests (splat atomics are the only order-dependent sums); the
plain one-job-at-a-time schedule gives the same images."""
import numpy as np
import pytest

from selas_amd import native as N
from selas_amd import serving as S
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _close(a, b):
    return (np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5).mean()


SPECS = {0: lambda: dict(W.room_scene(3), merge_radius=0.005),
         1: lambda: W.vcm_scene(seed=21, n_meshes=8, nu=16, nv=8),
         2: lambda: W.vcm_scene(seed=22, n_meshes=6, nu=12, nv=6, textures=False)}


def _requests(n=12, rate=400.0):
    rng = np.random.default_rng(8)
    jobs = [{"scene": int(rng.integers(0, 3)), "width": int(rng.choice([16, 24, 40, 64])),
             "height": int(rng.choice([16, 32, 48])), "iterations": int(rng.integers(1, 7)),
             "seed": int(rng.integers(0, 1 << 40)), "max_path_length": int(rng.choice([4, 8, 10]))}
            for _ in range(n)]
    return S.poisson_requests(3, rate, jobs)


def _alone(engine, r):
    sc, cam, radius, alpha = engine.scenes[r.scene]
    return N.render_vcm(sc, cam, r.width, r.height, r.iterations, r.max_path_length, radius, alpha, r.seed)


@pytest.mark.parametrize("plain", [False, True])
def test_served_results_equal_jobs_alone(gpu_ctx, plain):
    eng = S.GpuEngine(gpu_ctx, lambda sid: SPECS[sid](), wavefront_paths=6000, max_path_length=10, plain=plain)
    try:
        reqs = _requests()
        loop = S.RequestLoop(eng, hbm_budget=1 << 34, max_slots=1 if plain else 8, keep_images=True)
        loop.run(reqs)
        assert all(r.status == "done" for r in reqs)
        if not plain:
            assert loop.peak_in_flight > 1  # the requests did overlap
        assert sorted(eng.resident_scenes()) == sorted({r.scene for r in reqs})
        for r in reqs:
            assert r.image.shape == (r.height, r.width, 3)
            assert _close(r.image, _alone(eng, r)) >= 0.999, r
        st = S.summarize(reqs, loop.makespan)
        assert st["jobs"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0
    finally:
        eng.close()


def test_admission_bounded_by_footprint_on_gpu(gpu_ctx):
    eng = S.GpuEngine(gpu_ctx, lambda sid: SPECS[sid](), wavefront_paths=6000, max_path_length=10)
    try:
        for sid in SPECS:
            eng.load_scene(sid)
        scene_bytes = sum(eng.scene_bytes(s) for s in SPECS)
        assert all(eng.scene_bytes(s) > 0 for s in SPECS)
        reqs = _requests(10, rate=1000.0)
        biggest = max(12 * r.width * r.height for r in reqs)
        loop = S.RequestLoop(eng, hbm_budget=scene_bytes + 2 * biggest, max_slots=8, keep_images=True)
        loop.run(reqs)
        assert all(r.status == "done" for r in reqs)
        assert loop.peak_used <= loop.budget and loop.evictions == 0
        for r in reqs[:4]:
            assert _close(r.image, _alone(eng, r)) >= 0.999
    finally:
        eng.close()


def test_eviction_and_reload_match_alone(gpu_ctx):
    """An HBM budget that holds one scene at a time: scenes are evicted (released while the server
    lives) and built again for later requests; every result equals the job alone on a fresh build."""
    eng = S.GpuEngine(gpu_ctx, lambda sid: SPECS[sid](), wavefront_paths=6000, max_path_length=10)
    try:
        sizes = {}
        for sid in SPECS:
            eng.load_scene(sid)
            sizes[sid] = eng.scene_bytes(sid)
            eng.evict_scene(sid)
        rng = np.random.default_rng(4)
        order = [0, 1, 2, 1, 0, 2, 0]
        jobs = [{"scene": s, "width": 32, "height": 24, "iterations": int(rng.integers(1, 4)),
                 "seed": int(rng.integers(0, 1 << 40)), "max_path_length": 8} for s in order]
        reqs = S.poisson_requests(5, 1000.0, jobs)
        loop = S.RequestLoop(eng, hbm_budget=max(sizes.values()) + 4 * 12 * 32 * 24, max_slots=4, keep_images=True)
        loop.run(reqs)
        assert all(r.status == "done" for r in reqs)
        assert loop.evictions >= 3
        for sid in list(eng.resident_scenes()):
            eng.evict_scene(sid)
        for r in reqs:
            eng.load_scene(r.scene)
            assert _close(r.image, _alone(eng, r)) >= 0.999, r
    finally:
        eng.close()
