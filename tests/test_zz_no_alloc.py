"""sgpu_intersect_noalloc (no device allocation: the traversal stacks spill into a fixed per-GPU
array) against sgpu_intersect through the C ABI: bit-exact t, u, v, primID, geomID, Ng and
instance ids on two calls in a row, the second with a smaller input, on a flat scene deep enough
for the stacks to spill past their LDS part and on a two-level instanced scene; plus its argument
errors. Collected last (its file name), so that it also runs after every other GPU test.
"""
import numpy as np
import pytest

import instancing as I
from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _both(ctx, scene, rays, with_inst):
    dr = N.DeviceRays(ctx, rays)
    ref = N.DeviceHits(ctx, dr.n, with_ng=True, with_inst=with_inst)
    got = N.DeviceHits(ctx, dr.n, with_ng=True, with_inst=with_inst)
    scene.intersect(dr, ref)
    scene.intersect_noalloc(dr, got)
    N.synchronize()
    scene.check()
    return ref.to_host(), got.to_host()


def _assert_bits(ref, got):
    assert ref.keys() == got.keys()
    for k in ref:
        a, b = ref[k].view(np.uint32), got[k].view(np.uint32)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{k}: {bad.size} mismatches, first {bad[:5]}"


def _subset(rays, start, n):
    return {k: np.ascontiguousarray(v[start:start + n]) for k, v in rays.items()}


def test_flat_scene_two_calls_match_intersect(gpu_ctx):
    t9 = N.gen_cluster_soup(3, 128, 1024, 0.03, 1e-3, 2e-2)
    scene = N.Scene(gpu_ctx)
    scene.add_soup(t9)
    scene.commit()
    try:
        cam = N.gen_camera_rays(W.CAMERA_POS, W.CAMERA_LOOK_AT, W.CAMERA_UP, 50.0, 256, 256, 1)
        bnc = N.gen_bounce_rays(4, t9, 65536)
        rays = {k: np.concatenate([cam[k], bnc[k]]) for k in N.RAY_FIELDS}
        info = scene.info()
        print(f"flat: {info}, {rays['org_x'].size} rays then 7000", flush=True)
        ref, got = _both(gpu_ctx, scene, rays, False)
        assert (ref["prim_id"] != N.INVALID_ID).mean() > 0.3
        _assert_bits(ref, got)
        # second call, smaller input: a different slice of the rays (the array keeps the last call's
        # spilled stacks; every entry is written before it is read)
        small = _subset(rays, 60000, 7000)
        ref2, got2 = _both(gpu_ctx, scene, small, False)
        _assert_bits(ref2, got2)
        assert (ref2["prim_id"] != N.INVALID_ID).any()
    finally:
        scene.release()


def _soup(seed, n, spread=1.0, size=0.25):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    return (c + rng.normal(scale=size, size=(n, 3, 3))).reshape(n, 9).astype(np.float32)


def _rays(seed, n, spread=9.0, target=4.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    d = rng.uniform(-target, target, (n, 3)) - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tfar = np.full(n, np.inf, np.float32)
    tfar[::3] = rng.uniform(1.0, 15.0, tfar[::3].shape[0]).astype(np.float32)
    return {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": d[:, 0], "dir_y": d[:, 1],
            "dir_z": d[:, 2], "tnear": np.zeros(n, np.float32), "tfar": tfar}


def test_instanced_scene_two_calls_match_intersect(gpu_ctx):
    rng = np.random.default_rng(9)
    m = I.Model([("soup", _soup(1, 3000, size=0.1)), ("soup", _soup(2, 500, spread=0.5))])
    dm = m.device(gpu_ctx)
    sub = N.Scene(gpu_ctx)
    for _ in range(3):
        sub.add_instance(dm, np.asarray(I.random_transform(rng, spread=1.5), np.float32).reshape(4, 4).T)
    sub.commit()
    top = N.Scene(gpu_ctx)
    for k in range(6):
        child = sub if k % 2 == 0 else dm
        top.add_instance(child, np.asarray(I.random_transform(rng), np.float32).reshape(4, 4).T)
    top.commit()
    try:
        rays = _rays(5, 40000)
        print("instanced: 40000 rays then 5000", flush=True)
        ref, got = _both(gpu_ctx, top, rays, True)
        assert (ref["prim_id"] != N.INVALID_ID).mean() > 0.1
        assert (ref["inst_id1"] != N.INVALID_ID).any()
        _assert_bits(ref, got)
        small = _subset(rays, 1234, 5000)
        ref2, got2 = _both(gpu_ctx, top, small, True)
        _assert_bits(ref2, got2)
    finally:
        top.release()
        sub.release()
        dm.release()


def test_noalloc_errors(gpu_ctx):
    scene = N.Scene(gpu_ctx)
    try:
        rays = N.DeviceRays(gpu_ctx, _rays(6, 64))
        hits = N.DeviceHits(gpu_ctx, 64)
        with pytest.raises(N.SgpuError):
            scene.intersect_noalloc(rays, hits)  # not committed
        scene.add_soup(_soup(7, 100))
        scene.commit()
        with pytest.raises(N.SgpuError):
            N.check(N.lib.sgpu_intersect_noalloc(scene.handle, N.C.byref(rays.soa), None, 64, None))
        scene.intersect_noalloc(rays, hits)
        N.synchronize()
    finally:
        scene.release()
