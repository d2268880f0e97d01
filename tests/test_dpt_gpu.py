"""GPU parity of sgpu_render_deferred_pt (the replacement of DeferredPathTracer::GenerateImage)
against the oracle's restatement of its estimator (oracle/orc_pt.c, orc_render_dpt), through the C
ABI.

Both sides draw each path's random numbers from the same PCG32 stream, take the same CMJ primary
rays and bit-identical ray casts, so images agree pixel by pixel up to the device libm. Bars,
written here: |gpu - cpu| <= 1e-3 |cpu| + 1e-5 for at least 99 % of the pixel values and RMSE
<= 1e-3 x the mean; with and without the hit sort the images are bit-identical (a path's random
numbers and the order of its contributions do not depend on the order in which hits are shaded).
"""
import numpy as np
import pytest

import oracle as O
import scenes
from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _agree(gpu, cpu):
    assert np.isfinite(gpu).all()
    assert cpu.mean() > 0.005
    close = np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5
    assert close.mean() >= 0.99, f"only {close.mean():.4f} of pixel values agree"
    rmse = np.sqrt(np.mean((gpu - cpu) ** 2))
    assert rmse <= 1e-3 * cpu.mean(), rmse / cpu.mean()


def _background_plane():
    sc = scenes.plane_scene()
    sc["background"] = np.array([0.4, 0.5, 0.7], np.float32)
    return sc


CASES = {
    "cube_sky_lights": lambda: scenes.cube_scene(),
    "plane_background": _background_plane,
    "room_disney": lambda: W.room_scene(1),
    "textured_meshes": lambda: W.sized_vcm_scene(5, 2e4),
    "instanced_sky": lambda: scenes.instanced_scene(ibl=True),
}


@pytest.mark.parametrize("flags", [0, N.DPT_SORT_HITS | N.DPT_SORT_RAYS])
@pytest.mark.parametrize("case", list(CASES))
def test_deferred_matches_oracle_per_pixel(gpu_ctx, case, flags):
    sc = CASES[case]()
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        rs = O.InstancedRenderScene(sc) if "instances" in sc else O.RenderScene(sc)
        w, h = 64, 48
        print(f"deferred {case} flags={flags}", flush=True)
        gpu = N.render_deferred_pt(scene, sc["camera"], w, h, 2, 2, 8, 77, flags=flags)
        cpu = O.render_dpt(rs, sc["camera"], w, h, 2, 2, 8, 77)
        _agree(gpu, cpu)
    finally:
        scene.release()


@pytest.mark.parametrize("case", ["textured", "instanced"])
def test_hit_sort_does_not_change_the_image(gpu_ctx, case):
    sc = W.sized_vcm_scene(6, 5e4) if case == "textured" else scenes.instanced_scene(ibl=True)
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        print(f"deferred sort on/off {case}", flush=True)
        b = N.render_deferred_pt(scene, sc["camera"], 96, 64, 3, 3, 10, 5, flags=0)
        assert b.mean() > 0.005
        for flags in (N.DPT_SORT_HITS, N.DPT_SORT_RAYS, N.DPT_SORT_HITS | N.DPT_SORT_RAYS):
            a = N.render_deferred_pt(scene, sc["camera"], 96, 64, 3, 3, 10, 5, flags=flags)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), flags
    finally:
        scene.release()


def test_uniform_sky_closed_form_on_device(gpu_ctx):
    sc = scenes.uniform_sky_plane()
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        v = N.render_deferred_pt(scene, sc["camera"], 64, 48, 4, 4, 10, 3)
        assert abs(v.mean() / 0.6 ** 2.2 - 1) < 0.003
    finally:
        scene.release()


def test_deferred_errors(gpu_ctx):
    sc = scenes.cube_scene()
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        for bad in (dict(spp_x=0), dict(spp_y=0), dict(max_path_length=0), dict(flags=8),
                    dict(spp_x=2048, spp_y=1024)):
            args = dict(spp_x=1, spp_y=1, max_path_length=4, flags=0)
            args.update(bad)
            with pytest.raises(N.SgpuError):
                N.render_deferred_pt(scene, sc["camera"], 8, 8, args["spp_x"], args["spp_y"],
                                     args["max_path_length"], 1, flags=args["flags"])
    finally:
        scene.release()
