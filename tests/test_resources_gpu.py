"""GPU: a scene read from Selas resource files (tests/golden/resources, written by the reference's
own serializer) committed to the device and rendered: the path tracer and VCM per pixel against the
oracle on the same loaded description (99 % of pixels within 1e-3; VCM RMSE <= 2e-3 of the mean),
and the loaded scene's closest hits on its camera rays against the oracle's traversal."""
import numpy as np
import pytest

import oracle as O
import resource_fixtures as F
from selas_amd import native as N
from selas_amd import resources as R

pytestmark = pytest.mark.gpu


def _loaded():
    return R.load_scene(F.FIXTURE_ROOT, "rooms", camera_index=1, ptex_loader=lambda path: F.floor_ptex())


def _close(gpu, cpu):
    close = np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5
    rmse = np.sqrt(np.mean((gpu - cpu) ** 2))
    return close.mean(), rmse / max(cpu.mean(), 1e-12)


def test_loaded_scene_renders_match_oracle(gpu_ctx):
    ls = _loaded()
    sc = ls.description
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        rs = O.InstancedRenderScene(sc)
        print("resources: rooms, path tracer", flush=True)
        g = N.render_pt(scene, ls.camera, 64, 48, 4, 8, 3, N.PT_NEE_COSINE_ONCE)
        c = O.render_pt(rs, ls.camera, 64, 48, 4, 8, 3, 1)
        frac, _ = _close(g, c)
        assert c.mean() > 0.01 and frac >= 0.99, frac
        print("resources: rooms, VCM", flush=True)
        g = N.render_vcm(scene, ls.camera, 64, 48, 3, 8, 0.02, 0.75, 29)
        c, _ = O.render_vcm(rs, ls.camera, 64, 48, 3, 8, 0.02, 0.75, 29)
        frac, rel_rmse = _close(g, c)
        assert c.mean() > 0.01 and frac >= 0.99 and rel_rmse <= 2e-3, (frac, rel_rmse)
    finally:
        scene.release()


def test_loaded_scene_hits_match_oracle(gpu_ctx):
    """Camera rays of the scene's camera: (t, u, v, prim, geom, inst0, inst1) and occlusion of every
    ray equal the oracle's traversal of the composed instances (ids as Selas numbers them)."""
    import instancing as I
    from selas_amd.workloads import geom_tri9

    ls = _loaded()
    sc = ls.description
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        models = [I.Model([("quads", g["positions"], g["quads"]) if "quads" in g else ("soup", geom_tri9(g))
                           for g in m["geoms"]]) for m in sc["models"]]
        cm = lambda m: np.ascontiguousarray(np.asarray(m).T.reshape(-1), np.float32)  # noqa: E731
        top = [([(models[mi], cm(m2)) for mi, m2 in sc["subscenes"][idx]], cm(m)) for _, idx, m in sc["instances"]]
        cam = ls.camera
        rays = N.gen_camera_rays(cam["position"], cam["look_at"], cam["up"], cam["fov"], 96, 64, 1)
        n = rays["org_x"].shape[0]
        print(f"resources: rooms, {n} camera rays", flush=True)
        dr = N.DeviceRays(gpu_ctx, rays)
        dh = N.DeviceHits(gpu_ctx, n, with_ng=False, with_inst=True)
        scene.intersect(dr, dh)
        occ = gpu_ctx.empty(4 * n)
        scene.occluded(dr, occ)
        N.synchronize()
        scene.check()
        got = dh.to_host()
        ref = I.oracle_hits(I.compose(top), rays)
        assert (ref["prim_id"] != N.INVALID_ID).mean() > 0.5
        assert len(set(ref["inst_id0"][ref["prim_id"] != N.INVALID_ID].tolist())) == 3
        for k in ("prim_id", "geom_id", "inst_id0", "inst_id1"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
        for k in ("t", "u", "v"):
            np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32), err_msg=k)
        np.testing.assert_array_equal(occ.download(np.uint32, n), ref["occluded"])
    finally:
        scene.release()


def test_loaded_scene_with_curves_renders_match_oracle(gpu_ctx, tmp_path):
    """The fixture scene with hair on its panel model (ModelResource curves: InitializeCurves, the
    curveModelName material): loaded, committed and rendered by the path tracer and VCM (with the
    scene's image-based light, whose emission sphere takes in the curves) per pixel against the
    oracle's instanced scene."""
    from test_curves_gpu import _close as _close_curves
    from test_curves_gpu import _hair

    files, _ = F.fixture_resources()
    tp = np.array([(-1, 0, 0), (1, 0, 0), (1, 1.5, 0), (-1, 1.5, 0), (0, 2.2, 0)], np.float32)
    v, idx = _hair(21, 60, 8, (0.0, 1.5, 0.0), 0.8, 0.02, step=0.08)
    brass = R.murmur3_32("brass")
    old = files[("model", "panel")]
    d, g = F.model_dict([("panel_main", brass, [[0, 1, 2], [0, 2, 3]]), ("panel_top", R.murmur3_32("nothing"),
                                                                         [[3, 2, 4]])], tp,
                        list(old["materials"]), list(old["material_hashes"]), curves=[(0, len(idx))],
                        curve_indices=idx, curve_vertices=v)
    files[("model", "panel")], files[("geometry", "panel")] = d, g
    F.write_tree(str(tmp_path), files, F.py_write)
    ls = R.load_scene(str(tmp_path), "rooms", camera_index=1, ptex_loader=lambda path: F.floor_ptex())
    sc = ls.description
    assert any("curves" in gg for m in sc["models"] for gg in m["geoms"])
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        rs = O.InstancedRenderScene(sc)
        print("resources with curves: path tracer", flush=True)
        gi = N.render_pt(scene, ls.camera, 64, 48, 4, 8, 3, N.PT_NEE_COSINE_ONCE)
        c = O.render_pt(rs, ls.camera, 64, 48, 4, 8, 3, 1)
        frac, _ = _close(gi, c)
        assert c.mean() > 0.01 and frac >= 0.99, frac
        print("resources with curves: VCM", flush=True)
        gi = N.render_vcm(scene, ls.camera, 64, 48, 3, 8, 0.02, 0.75, 29)
        c, _ = O.render_vcm(rs, ls.camera, 64, 48, 3, 8, 0.02, 0.75, 29)
        frac, rel_rmse = _close_curves(gi, c)
        assert c.mean() > 0.01 and frac >= 0.99 and rel_rmse <= 2e-3, (frac, rel_rmse)
    finally:
        scene.release()
