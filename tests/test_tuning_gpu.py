"""Launch-shape tuning (sgpu_set_tuning): values round-trip, out-of-range values are refused, and
images do not depend on the launch shape (slices are independent of how they are packed)."""
import numpy as np
import pytest

from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_tuning_roundtrip_and_limits(gpu_ctx):
    old = gpu_ctx.tuning()
    try:
        gpu_ctx.set_tuning(vcm_wavefront_paths=1 << 14, trace_blocks_per_cu=3, trace_tree=0, trace_lds_nodes=9,
                           vcm_graphs=0, vcm_order=3, vcm_merge_lanes=32, bvh_builder=0)
        want = {"vcm_wavefront_paths": 1 << 14, "trace_blocks_per_cu": 3, "trace_tree": 0, "trace_lds_nodes": 9,
                "vcm_graphs": 0, "vcm_order": 3, "vcm_merge_lanes": 32, "bvh_builder": 0}
        assert gpu_ctx.tuning() == want
        for bad in ({"vcm_wavefront_paths": 1 << 11}, {"vcm_wavefront_paths": (1 << 24) + 1},
                    {"trace_blocks_per_cu": 0}, {"trace_blocks_per_cu": 33}, {"trace_tree": 2},
                    {"trace_lds_nodes": 10}, {"vcm_graphs": 2}, {"vcm_order": 4}, {"vcm_merge_lanes": 4},
                    {"vcm_merge_lanes": 64}, {"bvh_builder": 2}):
            with pytest.raises(N.SgpuError):
                gpu_ctx.set_tuning(**bad)
        assert gpu_ctx.tuning() == want
        info = gpu_ctx.device_info()
        assert info["compute_units"] > 0 and info["total_bytes"] >= info["free_bytes"] > 0 and info["name"]
    finally:
        gpu_ctx.set_tuning(**old)


def test_images_do_not_depend_on_launch_shape(gpu_ctx):
    sc = W.vcm_scene(seed=31, n_meshes=8, nu=16, nv=8)
    scene = N.build_render_scene(gpu_ctx, sc)
    old = gpu_ctx.tuning()
    try:
        imgs = []
        for paths, blocks in ((1 << 12, 1), (1 << 14, 8), (1 << 16, 32)):
            gpu_ctx.set_tuning(vcm_wavefront_paths=paths, trace_blocks_per_cu=blocks)
            img, st = N.render_vcm(scene, sc["camera"], 48, 40, 9, 10, sc["merge_radius"], 0.75, 5, return_stats=True)
            imgs.append((img, st["wavefronts"]))
        # 1920 paths per iteration: 2, 8 and all 9 iterations per wavefront
        assert [w for _, w in imgs] == [5, 2, 1]
        for img, _ in imgs[1:]:
            close = np.abs(img - imgs[0][0]) <= 1e-3 * np.abs(imgs[0][0]) + 1e-5
            assert close.mean() >= 0.999
    finally:
        gpu_ctx.set_tuning(**old)
        scene.release()


def test_graph_replays_match_eager_launches(gpu_ctx):
    """The light and camera passes replayed from captured graphs (queue counts read on the device,
    launches at capacity) give the images of eager launches sized from counts read back: a job of
    5 wavefronts (5 replays of each graph, eager grid builds between them) and a two-scene batch."""
    scs = [W.vcm_scene(seed=33, n_meshes=8, nu=16, nv=8), W.vcm_scene(seed=34, n_meshes=6, nu=12, nv=6, textures=False)]
    scenes = [N.build_render_scene(gpu_ctx, sc) for sc in scs]
    old = gpu_ctx.tuning()
    try:
        out = {}
        for graphs in (0, 1):
            gpu_ctx.set_tuning(vcm_wavefront_paths=1 << 12, vcm_graphs=graphs)
            sc = scs[0]
            img, st = N.render_vcm(scenes[0], sc["camera"], 48, 40, 9, 10, sc["merge_radius"], 0.75, 5,
                                   return_stats=True)
            assert st["wavefronts"] == 5
            jobs = [{"scene": scenes[k], "camera": scs[k]["camera"], "width": 32, "height": 24, "iterations": 3,
                     "max_path_length": 8, "base_radius": float(scs[k]["merge_radius"]), "radius_alpha": 0.75,
                     "seed": 11 + k} for k in range(2)]
            out[graphs] = (img, st, N.render_vcm_batch(jobs))
        (a, sa, ba), (b, sb, bb) = out[0], out[1]
        assert sa["samples"] == sb["samples"] and sa["closest_rays"] == sb["closest_rays"]
        assert sa["shadow_rays"] == sb["shadow_rays"] and sa["light_vertices"] == sb["light_vertices"]
        for x, y in [(a, b)] + list(zip(ba, bb)):
            close = np.abs(x - y) <= 1e-4 * np.abs(y) + 1e-6
            assert close.mean() >= 0.999
    finally:
        gpu_ctx.set_tuning(**old)
        for s in scenes:
            s.release()


@pytest.mark.parametrize("graphs,lanes", [(0, 16), (1, 16), (1, 32)])
def test_images_do_not_depend_on_work_order(gpu_ctx, graphs, lanes):
    """vcm_order (merge range searches in spatial bins, hits shaded grouped by material) and the lanes
    per range search reorder the work of a bounce and nothing else: a textured job of 3 wavefronts and a two-scene batch give the
    images of the queue order (99.9 % of the values within 1e-4: the light-tracing splats are summed
    with atomics, whose order follows the shading order) and the same ray and merge counts."""
    scs = [W.vcm_scene(seed=35, n_meshes=8, nu=16, nv=8), W.vcm_scene(seed=36, n_meshes=6, nu=12, nv=6, textures=False)]
    scenes = [N.build_render_scene(gpu_ctx, sc) for sc in scs]
    old = gpu_ctx.tuning()
    try:
        out = {}
        for order in (0, N.ORDER_MERGE_CELLS, N.ORDER_SHADE_MATERIAL, N.ORDER_MERGE_CELLS | N.ORDER_SHADE_MATERIAL):
            print(f"vcm_order {order}, graphs {graphs}, merge lanes {lanes}", flush=True)
            gpu_ctx.set_tuning(vcm_wavefront_paths=1 << 12, vcm_graphs=graphs, vcm_order=order, vcm_merge_lanes=lanes)
            sc = scs[0]
            img, st = N.render_vcm(scenes[0], sc["camera"], 48, 40, 6, 10, sc["merge_radius"], 0.75, 5,
                                   return_stats=True)
            assert st["wavefronts"] == 3
            # merge counters (eager launches): range searches and photons are the same in every order
            _, cnt = N.render_vcm(scenes[0], sc["camera"], 48, 40, 6, 10, sc["merge_radius"], 0.75, 5,
                                  return_stats=True, flags=N.VCM_COUNT_MERGE)
            jobs = [{"scene": scenes[k], "camera": scs[k]["camera"], "width": 32, "height": 24, "iterations": 3,
                     "max_path_length": 8, "base_radius": float(scs[k]["merge_radius"]), "radius_alpha": 0.75,
                     "seed": 21 + k} for k in range(2)]
            out[order] = (img, st, cnt, N.render_vcm_batch(jobs))
        a, sa, ca, ba = out[0]
        assert ca["merge_queries"] > 0 and ca["merge_near"] > 0
        for order, (b, sb, cb, bb) in out.items():
            for k in ("samples", "closest_rays", "shadow_rays", "light_vertices"):
                assert sa[k] == sb[k], (order, k)
            for k in ("merge_queries", "merge_visited", "merge_near"):
                assert ca[k] == cb[k], (order, k)
            for x, y in [(a, b)] + list(zip(ba, bb)):
                assert np.isfinite(y).all()
                close = np.abs(x - y) <= 1e-4 * np.abs(x) + 1e-6
                assert close.mean() >= 0.999, (order, close.mean())
    finally:
        gpu_ctx.set_tuning(**old)
        for s in scenes:
            s.release()
