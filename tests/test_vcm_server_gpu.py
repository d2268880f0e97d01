"""The persistent VCM executor (sgpu_vcm_server_*): jobs admitted while others are running share
wavefronts, and each job's image is the one it gets rendered alone (per pixel, to the bar of the
other VCM tests: splat atomics are the only order-dependent sums)."""
import numpy as np
import pytest

from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _close(a, b):
    return (np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5).mean()


@pytest.fixture(scope="module")
def two_scenes(gpu_ctx):
    specs = [W.room_scene(2), W.vcm_scene(seed=11, n_meshes=8, nu=16, nv=8)]
    built = [N.build_render_scene(gpu_ctx, sc) for sc in specs]
    yield list(zip(specs, built))
    for b in built:
        b.release()


def _job(sc, scene, w, h, its, seed, L=8):
    return {"scene": scene, "camera": sc["camera"], "width": w, "height": h, "iterations": its,
            "max_path_length": L, "base_radius": float(sc.get("merge_radius", 0.005)), "radius_alpha": 0.75,
            "seed": seed}


def test_jobs_admitted_mid_run_match_alone(gpu_ctx, two_scenes):
    (sa, a), (sb, b) = two_scenes
    jobs = [_job(sa, a, 48, 32, 7, 1), _job(sb, b, 64, 40, 3, 2), _job(sa, a, 20, 20, 5, 3),
            _job(sb, b, 96, 64, 2, 4), _job(sa, a, 30, 16, 9, 5, L=4)]
    # small wavefronts, so that jobs run over several steps and overlap
    srv = N.VcmServer(gpu_ctx, wavefront_paths=8192, max_path_length=10)
    fbs = [gpu_ctx.empty(12 * j["width"] * j["height"]) for j in jobs]
    try:
        ids = {}
        retired = []
        ids[srv.submit(jobs[0]["scene"], jobs[0], fbs[0])] = 0
        ids[srv.submit(jobs[1]["scene"], jobs[1], fbs[1])] = 1
        steps = 0
        pending = [2, 3, 4]
        while True:
            done, active = srv.step()
            steps += 1
            retired += [ids[d] for d in done]
            if pending and steps % 2 == 0:  # new arrivals while others are in flight
                k = pending.pop(0)
                ids[srv.submit(jobs[k]["scene"], jobs[k], fbs[k])] = k
                active += 1
            if active == 0 and not pending:
                break
            assert steps < 500
        assert sorted(retired) == list(range(len(jobs)))
        st = srv.stats()
        assert st["samples"] == sum(j["width"] * j["height"] * j["iterations"] for j in jobs)
        for k, j in enumerate(jobs):
            mixed = fbs[k].download(np.float32, 3 * j["width"] * j["height"]).reshape(j["height"], j["width"], 3)
            alone = N.render_vcm(j["scene"], j["camera"], j["width"], j["height"], j["iterations"],
                                 j["max_path_length"], j["base_radius"], j["radius_alpha"], j["seed"])
            assert _close(mixed, alone) >= 0.999, k
    finally:
        srv.close()
        for f in fbs:
            f.free()


def test_server_limits_and_footprint(gpu_ctx, two_scenes):
    (sa, a), _ = two_scenes
    f1 = N.VcmServer.footprint(gpu_ctx, 1 << 16, 10)
    f2 = N.VcmServer.footprint(gpu_ctx, 1 << 18, 10)
    assert 0 < f1 < f2
    srv = N.VcmServer(gpu_ctx, wavefront_paths=1024, max_path_length=6)
    fb = gpu_ctx.empty(12 * 64 * 64)
    try:
        with pytest.raises(RuntimeError):
            srv.submit(a, _job(sa, a, 64, 64, 1, 1), fb)  # larger than a wavefront
        with pytest.raises(RuntimeError):
            srv.submit(a, _job(sa, a, 16, 16, 1, 1, L=8), fb)  # longer paths than configured
        assert srv.step() == ([], 0)
    finally:
        srv.close()
        fb.free()


def test_scene_held_while_jobs_admitted(gpu_ctx):
    """A scene with admitted jobs can be neither released nor given new lights or an image-based
    light (the jobs hold its device pointers); once its jobs retire both work again."""
    spec = W.room_scene(2)
    sc = N.build_render_scene(gpu_ctx, spec)
    srv = N.VcmServer(gpu_ctx, wavefront_paths=4096, max_path_length=6)
    fb = gpu_ctx.empty(12 * 32 * 32)
    try:
        srv.submit(sc, _job(spec, sc, 32, 32, 3, 7, L=6), fb)
        for call in (sc.release, lambda: sc.set_ibl(np.ones((4, 8, 3), np.float32))):
            with pytest.raises(N.SgpuError) as e:
                call()
            assert e.value.code == N.SGPU_E_STATE
        assert sc.handle is not None
        active = 1
        while active:
            _, active = srv.step()
        sc.set_ibl(np.ones((4, 8, 3), np.float32))
        sc.set_ibl(None)
    finally:
        srv.close()
        fb.free()
        sc.release()
    assert sc.handle is None


def test_released_scene_address_reuse(gpu_ctx):
    """Scenes released between a live server's jobs and new scenes built in their place (often at
    the same host address) are traversed through fresh table entries: each job's image equals the
    job rendered alone on its own scene."""
    srv = N.VcmServer(gpu_ctx, wavefront_paths=8192, max_path_length=8)
    try:
        for rnd, seed in enumerate((31, 32, 33)):
            spec = W.vcm_scene(seed=seed, n_meshes=4 + 2 * rnd, nu=8 + 4 * rnd, nv=8)
            sc = N.build_render_scene(gpu_ctx, spec)
            j = _job(spec, sc, 40, 24, 3, 100 + rnd, L=8)
            fb = gpu_ctx.empty(12 * 40 * 24)
            try:
                srv.submit(sc, j, fb)
                active = 1
                while active:
                    _, active = srv.step()
                served = fb.download(np.float32, 3 * 40 * 24).reshape(24, 40, 3)
                alone = N.render_vcm(sc, j["camera"], 40, 24, 3, 8, j["base_radius"], 0.75, j["seed"])
                assert _close(served, alone) >= 0.999, rnd
            finally:
                fb.free()
                sc.release()
    finally:
        srv.close()
