"""GPU parity of sgpu_render_vcm (the replacement of VCM::GenerateImage) against the oracle's CPU
restatement, through the C ABI.

Both sides draw each path's random numbers from the same PCG32 stream (one per iteration, path and
pass), evaluate cos/sin/acos/pow in double rounded to float, and use bit-identical ray casts and
photon grids, so paths take the same branches and images agree pixel by pixel up to the order of
the atomic adds of light-tracing splats (and a path whose double-evaluated function lands within
2^-29 of a float rounding boundary on one side only). Bars, written here:
  * per pixel: |gpu - cpu| <= 1e-3 |cpu| + 1e-5 for at least 99 % of the pixel values;
  * whole image: RMSE(gpu - cpu) <= 2e-3 x mean radiance;
  * closed form: the floor lit by a point and a quad light (tests/scenes.py) within 1 % in mean;
  * convergence: with the same seeds, RMSE(gpu, 4096-iteration CPU reference) is within 2 % of
    RMSE(cpu, same reference) at equal iterations.
"""
import numpy as np
import pytest

import oracle as O
import scenes
from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(gpu_ctx, sc):
    return N.build_render_scene(gpu_ctx, sc), O.RenderScene(sc)


def _close(gpu, cpu):
    close = np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5
    rmse = np.sqrt(np.mean((gpu - cpu) ** 2))
    return close.mean(), rmse / max(cpu.mean(), 1e-12)


@pytest.fixture(scope="module")
def room(gpu_ctx):
    sc = W.room_scene(1)
    scene, rs = _pair(gpu_ctx, sc)
    yield sc, scene, rs
    scene.release()


def test_vcm_matches_cpu_per_pixel_room(room):
    sc, scene, rs = room
    w, h, its = 80, 64, 3
    print(f"room {w}x{h} its={its}", flush=True)
    gpu, st = N.render_vcm(scene, sc["camera"], w, h, its, 10, 0.0052, 0.75, 1234, return_stats=True)
    cpu, cst = O.render_vcm(rs, sc["camera"], w, h, its, 10, 0.0052, 0.75, 1234)
    assert np.isfinite(gpu).all()
    assert cpu.mean() > 0.01
    frac, rel_rmse = _close(gpu, cpu)
    assert frac >= 0.99, f"only {frac:.4f} of pixel values agree"
    assert rel_rmse <= 2e-3, rel_rmse
    assert st["samples"] == w * h * its
    lv = st["light_vertices"] / its
    assert abs(lv / cst["light_vertices_per_iteration"] - 1) < 1e-3


def test_vcm_matches_cpu_per_pixel_displaced_meshes(gpu_ctx):
    """The VCM workload's scene kind (displaced-grid meshes with per-mesh Disney parameters,
    transparent meshes included) at reduced size."""
    sc = W.vcm_scene(seed=5, n_meshes=12, nu=16, nv=8)
    scene, rs = _pair(gpu_ctx, sc)
    try:
        w, h, its = 64, 48, 2
        r = sc["merge_radius"]
        print(f"displaced {w}x{h} its={its}", flush=True)
        gpu = N.render_vcm(scene, sc["camera"], w, h, its, 10, r, 0.75, 77)
        cpu, _ = O.render_vcm(rs, sc["camera"], w, h, its, 10, r, 0.75, 77)
        assert np.isfinite(gpu).all() and cpu.mean() > 0
        frac, rel_rmse = _close(gpu, cpu)
        assert frac >= 0.99, f"only {frac:.4f} of pixel values agree"
        assert rel_rmse <= 2e-3, rel_rmse
    finally:
        scene.release()


def test_vcm_plane_closed_form(gpu_ctx):
    sc = scenes.plane_scene()
    scene, _ = _pair(gpu_ctx, sc)
    try:
        w, h = 128, 96
        ref = scenes.plane_image(sc, w, h)
        print("plane closed form", flush=True)
        gpu = N.render_vcm(scene, sc["camera"], w, h, 32, 10, 0.03, 0.75, 3)
        ratio = gpu.mean((0, 1)) / ref.mean((0, 1))
        assert np.all(np.abs(ratio - 1) < 0.01), ratio
    finally:
        scene.release()


def test_vcm_converges_like_cpu(room):
    sc, scene, rs = room
    w, h = 32, 24
    r = 0.0052
    print("convergence", flush=True)
    ref, _ = O.render_vcm(rs, sc["camera"], w, h, 4096, 10, r, 0.75, 99)
    gpu = N.render_vcm(scene, sc["camera"], w, h, 16, 10, r, 0.75, 7)
    cpu, _ = O.render_vcm(rs, sc["camera"], w, h, 16, 10, r, 0.75, 7)
    e_gpu = np.sqrt(np.mean((gpu - ref) ** 2))
    e_cpu = np.sqrt(np.mean((cpu - ref) ** 2))
    assert abs(e_gpu / e_cpu - 1.0) < 0.02, (e_gpu, e_cpu)
    gpu256 = N.render_vcm(scene, sc["camera"], w, h, 256, 10, r, 0.75, 11)
    assert np.sqrt(np.mean((gpu256 - ref) ** 2)) < 0.5 * e_gpu


def test_vcm_iteration_base_splits_a_render(room):
    sc, scene, _ = room
    args = (sc["camera"], 48, 32)
    full = N.render_vcm(scene, *args, 4, 10, 0.005, 0.75, 4)
    a = N.render_vcm(scene, *args, 2, 10, 0.005, 0.75, 4, iteration_base=0)
    b = N.render_vcm(scene, *args, 2, 10, 0.005, 0.75, 4, iteration_base=2)
    frac, rel_rmse = _close(full, 0.5 * (a + b))
    assert frac >= 0.999 and rel_rmse <= 1e-4


def test_vcm_stats_and_timing_flags(room):
    """Phase timing and traversal counting change nothing in the image."""
    sc, scene, _ = room
    args = (sc["camera"], 48, 32, 2, 10, 0.005, 0.75, 6)
    plain, s0 = N.render_vcm(scene, *args, return_stats=True)
    timed, s1 = N.render_vcm(scene, *args, return_stats=True, flags=N.VCM_KERNEL_TIMES | N.VCM_COUNT_TRAVERSAL)
    assert np.array_equal(plain, timed) or _close(timed, plain)[0] >= 0.999
    assert s0["closest_ms"] == 0 and s0["nodes_visited"] == 0
    assert s1["closest_ms"] > 0 and s1["camera_ms"] > 0 and s1["grid_ms"] > 0
    assert s1["nodes_visited"] >= s1["closest_rays"] > 0 and s1["tris_tested"] > 0
    assert s1["rays"] == s1["closest_rays"] + s1["shadow_rays"] == s0["rays"]


def test_vcm_short_paths_and_no_lights(room, gpu_ctx):
    sc, scene, rs = room
    for L in (1, 2, 3):
        gpu = N.render_vcm(scene, sc["camera"], 24, 16, 1, L, 0.005, 0.75, 3)
        cpu, _ = O.render_vcm(rs, sc["camera"], 24, 16, 1, L, 0.005, 0.75, 3)
        frac, _ = _close(gpu, cpu)
        assert frac >= 0.99, (L, frac)
    dark = dict(sc)
    dark["lights"] = sc["lights"][:0]
    s2 = N.build_render_scene(gpu_ctx, dark)
    try:
        out, st = N.render_vcm(s2, sc["camera"], 16, 16, 2, 10, 0.005, 0.75, 1, return_stats=True)
        assert not out.any() and st["light_vertices"] == 0
    finally:
        s2.release()


def test_vcm_errors(room, gpu_ctx):
    sc, scene, _ = room
    fb = gpu_ctx.empty(3 * 16 * 16 * 4)
    try:
        def job(**kw):
            a = dict(iterations=1, iteration_base=0, max_path_length=10, base_radius=0.01, radius_alpha=0.75,
                     flags=0, seed=1)
            a.update(kw)
            return N.VcmJob(N.Camera.make(sc["camera"], kw.pop("w", 16), 16), a["iterations"], a["iteration_base"],
                            a["max_path_length"], a["base_radius"], a["radius_alpha"], a["flags"], a["seed"])

        def call(j):
            return N.lib.sgpu_render_vcm(scene.handle, N.C.byref(j), fb.ptr, None, None)

        assert call(job(iterations=0)) == N.SGPU_E_INVALID
        assert call(job(max_path_length=0)) == N.SGPU_E_INVALID
        assert call(job(max_path_length=32)) == N.SGPU_E_RANGE
        assert call(job(base_radius=0.0)) == N.SGPU_E_INVALID
        assert call(job(base_radius=float("nan"))) == N.SGPU_E_INVALID
        assert call(job(radius_alpha=1.5)) == N.SGPU_E_INVALID
        assert call(job(flags=8)) == N.SGPU_E_INVALID
        assert call(job(w=0)) == N.SGPU_E_INVALID
        assert N.lib.sgpu_render_vcm(scene.handle, None, fb.ptr, None, None) == N.SGPU_E_INVALID
        bare = N.Scene(gpu_ctx)
        bare.add_soup(sc["geoms"][0])
        bare.commit()
        assert N.lib.sgpu_render_vcm(bare.handle, N.C.byref(job()), fb.ptr, None, None) == N.SGPU_E_STATE
        bare.release()
        assert call(job()) == N.SGPU_OK
    finally:
        fb.free()
