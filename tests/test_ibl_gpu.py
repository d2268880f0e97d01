"""GPU parity of the image-based light (sgpu_scene_set_ibl; EmitIblLightSample, DirectIblLightSample
and ConnectToSkyLight in sgpu_render_vcm, EvaluateBackground in sgpu_render_pt) against the
oracle's CPU restatement, through the C ABI.

Bars, written here (those of the other VCM and path-tracer parity tests):
  * per pixel: |gpu - cpu| <= 1e-3 |cpu| + 1e-5 for at least 99 % of the pixel values, and
    RMSE(gpu - cpu) <= 2e-3 x mean radiance (VCM: splat atomics make sums order-dependent);
  * the path tracer's misses read the same texels: 99 % of pixel values within the same bound;
  * closed forms: a Lambert floor under a uniform sky has the radiance rho x L within 0.5 %, and a
    camera that sees only a uniform sky returns exposure x L exactly;
  * batches: a job's image does not depend on what it is batched with (a scene with the light
    next to scenes without it).
"""
import numpy as np
import pytest

import oracle as O
import scenes
from selas_amd import native as N

pytestmark = pytest.mark.gpu


def _close(gpu, cpu):
    close = np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5
    rmse = np.sqrt(np.mean((gpu - cpu) ** 2))
    return close.mean(), rmse / max(cpu.mean(), 1e-12)


@pytest.fixture(scope="module")
def cube(gpu_ctx):
    sc = scenes.cube_scene(lights=True)
    scene = N.build_render_scene(gpu_ctx, sc)
    yield sc, scene, O.RenderScene(sc)
    scene.release()


def test_vcm_with_sky_matches_cpu_per_pixel(cube):
    sc, scene, rs = cube
    w, h, its = 64, 48, 3
    print(f"cube + sky {w}x{h} its={its}", flush=True)
    gpu, st = N.render_vcm(scene, sc["camera"], w, h, its, 12, 0.02, 0.75, 4321, return_stats=True)
    cpu, cst = O.render_vcm(rs, sc["camera"], w, h, its, 12, 0.02, 0.75, 4321)
    assert np.isfinite(gpu).all() and cpu.mean() > 0.01
    frac, rel_rmse = _close(gpu, cpu)
    assert frac >= 0.99, f"only {frac:.4f} of pixel values agree"
    assert rel_rmse <= 2e-3, rel_rmse
    assert abs(st["light_vertices"] / its / cst["light_vertices_per_iteration"] - 1) < 1e-3


def test_vcm_sky_only_matches_cpu_per_pixel(gpu_ctx):
    sc = scenes.cube_scene(lights=False)
    sc["ibl"]["rotation"] = -2.5
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        rs = O.RenderScene(sc)
        print("cube, sky only", flush=True)
        gpu = N.render_vcm(scene, sc["camera"], 48, 32, 2, 12, 0.02, 0.75, 99)
        cpu, _ = O.render_vcm(rs, sc["camera"], 48, 32, 2, 12, 0.02, 0.75, 99)
        frac, rel_rmse = _close(gpu, cpu)
        assert frac >= 0.99 and rel_rmse <= 2e-3, (frac, rel_rmse)
    finally:
        scene.release()


def test_pt_with_sky_matches_cpu_per_pixel(cube):
    sc, scene, rs = cube
    w, h, spp = 48, 32, 4
    print(f"pt cube + sky {w}x{h} spp={spp}", flush=True)
    gpu = N.render_pt(scene, sc["camera"], w, h, spp, 12, 7, N.PT_NEE_COSINE_ONCE)
    cpu = O.render_pt(rs, sc["camera"], w, h, spp, 12, 7, 1)
    assert np.isfinite(gpu).all()
    close = np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5
    assert close.mean() >= 0.99, close.mean()


def test_vcm_floor_under_uniform_sky_closed_form(gpu_ctx):
    sc = scenes.uniform_sky_plane()
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        print("uniform sky floor", flush=True)
        v = N.render_vcm(scene, sc["camera"], 64, 48, 96, 10, 0.03, 0.75, 5)
        assert abs(v.mean() / 0.6 ** 2.2 - 1) < 0.005, v.mean() / 0.6 ** 2.2
        cam = {"position": (0.0, 2.0, 0.0), "look_at": (0.3, 5.0, 0.2), "up": (0.0, 0.0, 1.0), "fov": 50.0}
        scene.set_ibl(np.full((16, 32, 3), 1.5, np.float32), 0.0, 2.5)
        sky = N.render_vcm(scene, cam, 24, 16, 2, 10, 0.03, 0.75, 1)
        assert (sky == np.float32(1.5) * np.float32(2.5)).all()
    finally:
        scene.release()


def test_removing_the_sky_restores_the_background(gpu_ctx):
    sc = scenes.cube_scene(lights=True)
    plain = dict(sc)
    plain.pop("ibl")
    a = N.build_render_scene(gpu_ctx, sc)
    b = N.build_render_scene(gpu_ctx, plain)
    try:
        args = (sc["camera"], 32, 24, 2, 12, 0.02, 0.75, 3)
        with_sky = N.render_vcm(a, *args)
        a.set_ibl(None)
        without = N.render_vcm(a, *args)
        ref = N.render_vcm(b, *args)
        assert with_sky.mean() > 1.5 * without.mean()
        frac, rel_rmse = _close(without, ref)
        assert frac >= 0.99 and rel_rmse <= 2e-3
    finally:
        a.release()
        b.release()


def test_set_ibl_rejects_bad_input(gpu_ctx):
    sc = scenes.uniform_sky_plane()
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        bad = np.ones((4, 8, 3), np.float32)
        bad[1, 2, 0] = -1.0
        for rgb, rot, exp in [(bad, 0.0, 1.0), (np.full((4, 8, 3), np.nan, np.float32), 0.0, 1.0),
                              (np.ones((4, 8, 3), np.float32), float("inf"), 1.0),
                              (np.ones((4, 8, 3), np.float32), 0.0, -1.0)]:
            with pytest.raises(RuntimeError):
                scene.set_ibl(rgb, rot, exp)
        assert N.lib.sgpu_scene_set_ibl(scene.handle, 0, 4, N.fptr(np.ones(12, np.float32)), 0.0, 1.0) != 0
        assert N.lib.sgpu_scene_set_ibl(scene.handle, 16385, 1, N.fptr(np.ones(3 * 16385, np.float32)), 0.0,
                                        1.0) != 0
        fresh = N.Scene(gpu_ctx)
        try:
            fresh.add_soup(sc["geoms"][0])
            with pytest.raises(RuntimeError):
                fresh.set_ibl(np.ones((4, 8, 3), np.float32))  # not committed yet
        finally:
            fresh.release()
        # the light that was set before the failures is still the scene's
        v = N.render_vcm(scene, sc["camera"], 32, 24, 8, 10, 0.03, 0.75, 5)
        assert abs(v.mean() / 0.6 ** 2.2 - 1) < 0.05
    finally:
        scene.release()


def test_batched_jobs_with_and_without_sky(gpu_ctx):
    sky = scenes.cube_scene(lights=True)
    plain = scenes.cube_scene(lights=True, ibl=False)
    sa = N.build_render_scene(gpu_ctx, sky)
    sb = N.build_render_scene(gpu_ctx, plain)
    try:
        jobs = [(sa, 40, 32, 2, 11), (sb, 48, 24, 3, 12), (sa, 24, 24, 1, 13)]
        alone = [N.render_vcm(s, sky["camera"], w, h, its, 12, 0.02, 0.75, seed) for s, w, h, its, seed in jobs]
        batch = N.render_vcm_batch([dict(scene=s, camera=sky["camera"], width=w, height=h, iterations=its,
                                         max_path_length=12, base_radius=0.02, radius_alpha=0.75, seed=seed)
                                    for s, w, h, its, seed in jobs])
        for a, b in zip(alone, batch):
            frac, rel_rmse = _close(b, a)
            assert frac >= 0.99 and rel_rmse <= 2e-3, (frac, rel_rmse)
    finally:
        sa.release()
        sb.release()
