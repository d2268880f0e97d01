"""GPU parity of sgpu_intersect / sgpu_occluded (the replacements of rtcIntersect1/8 and
rtcOccluded1/8) against the oracle, through the C ABI.

Bar: bit-exact primID, geomID, t, u, v and occlusion flags against the oracle's FP32 restatement
(brute force on small inputs, the oracle's own SAH tree on larger ones), plus size-independent
invariants on the full 2M-triangle / 16M-ray workload.
"""
import numpy as np
import pytest

import oracle as O
from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_trace(ctx, tri9, rays, geoms=None):
    """Closest hit + occlusion on the GPU for a triangle soup split into `geoms` geometries."""
    scene = N.Scene(ctx)
    bounds = [0, tri9.shape[0]] if geoms is None else geoms
    gids = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        gids.append(scene.add_soup(tri9[a:b]))
    scene.commit()
    dr = N.DeviceRays(ctx, rays)
    dh = N.DeviceHits(ctx, dr.n)
    scene.intersect(dr, dh)
    occ = ctx.empty(4 * dr.n)
    scene.occluded(dr, occ)
    N.synchronize()
    scene.check()
    out = dh.to_host()
    out["occluded"] = occ.download(np.uint32, dr.n)
    out["info"] = scene.info()
    scene.release()
    return out


def global_prim(out, bounds):
    """Map (geomID, primID) back to the soup's global index."""
    g = out["geom_id"]
    p = out["prim_id"]
    res = np.full(g.shape, N.INVALID_ID, np.uint32)
    hit = g != N.INVALID_ID
    res[hit] = np.asarray(bounds, np.uint32)[g[hit]] + p[hit]
    return res


def assert_same(gpu, ref, bounds):
    gp = global_prim(gpu, bounds)
    mism = np.nonzero(gp != ref["prim"])[0]
    assert mism.size == 0, f"{mism.size} primID mismatches, first {mism[:5]}"
    for k in ("t", "u", "v"):
        a = gpu[k].view(np.uint32)
        b = ref[k].view(np.uint32)
        assert np.array_equal(a, b), f"{k}: {np.count_nonzero(a != b)} bit mismatches"


def dense_soup(seed, n_clusters=64, per=64):
    return N.gen_cluster_soup(seed, n_clusters, per, 0.08, 2e-2, 2e-1)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_soup_matches_brute_force(gpu_ctx, seed):
    t9 = dense_soup(seed)
    cam = N.gen_camera_rays(W.CAMERA_POS, W.CAMERA_LOOK_AT, W.CAMERA_UP, 50.0, 64, 64, 2)
    bnc = N.gen_bounce_rays(seed, t9, 8192)
    rays = {k: np.concatenate([cam[k], bnc[k]]) for k in N.RAY_FIELDS}
    bounds = [0, 1000, 1001, 2500, t9.shape[0]]  # several geometries, one with a single triangle
    gpu = gpu_trace(gpu_ctx, t9, rays, bounds)
    ref = O.closest_brute_f32(t9, rays)
    assert (ref["prim"] != N.INVALID_ID).mean() > 0.3
    assert_same(gpu, ref, bounds)
    assert np.array_equal(gpu["occluded"], O.occluded_brute_f32(t9, rays))
    # geometric normal = (v1 - v0) x (v2 - v0) of the hit triangle
    hit = ref["prim"] != N.INVALID_ID
    tri = t9[ref["prim"][hit]]
    e1 = tri[:, 3:6] - tri[:, 0:3]
    e2 = tri[:, 6:9] - tri[:, 0:3]
    ng = np.cross(e1.astype(np.float64), e2.astype(np.float64))
    got = np.stack([gpu["ng_x"][hit], gpu["ng_y"][hit], gpu["ng_z"][hit]], 1)
    assert np.allclose(got, ng, rtol=1e-5, atol=1e-9)
    assert (gpu["ng_x"][~hit] == 0).all()


def test_edge_cases(gpu_ctx):
    # duplicates (exact ties -> smaller index), degenerate triangles, axis-aligned rays,
    # tfar exactly at a hit, tnear beyond a hit, rays behind the scene
    base = np.array([[0, 0, 2, 4, 0, 2, 0, 4, 2],
                     [0, 0, 2, 4, 0, 2, 0, 4, 2],
                     [0, 0, 3, 1, 1, 3, 2, 2, 3],      # zero area
                     [-1, -1, 5, 6, -1, 5, -1, 6, 5]], np.float32)
    o = np.array([[1, 1, 0], [1, 1, 0], [1, 1, 0], [1, 1, 0], [1, 1, 10], [0.5, 0.5, 0], [1, 1, 2.5],
                  [10, 10, 0], [1, 1, -1]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, -1], [0, 0, 1], [0, 0, 1],
                  [0, 0, 1], [0, 0, -1]], np.float32)
    tnear = np.array([0, 0, 2.0, 0, 0, 0, 0, 0, 0], np.float32)
    tfar = np.array([np.inf, 2.0, np.inf, 1.5, np.inf, np.inf, np.inf, np.inf, np.inf], np.float32)
    rays = {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": d[:, 0], "dir_y": d[:, 1],
            "dir_z": d[:, 2], "tnear": tnear, "tfar": tfar}
    gpu = gpu_trace(gpu_ctx, base, rays)
    ref = O.closest_brute_f32(base, rays)
    assert_same(gpu, ref, [0, 4])
    assert np.array_equal(gpu["occluded"], O.occluded_brute_f32(base, rays))
    assert list(ref["prim"][:4]) == [0, 0, 3, N.INVALID_ID]


def test_empty_and_single_triangle_scenes(gpu_ctx):
    rays = N.gen_camera_rays(W.CAMERA_POS, W.CAMERA_LOOK_AT, W.CAMERA_UP, 50.0, 16, 16, 1)
    empty = np.zeros((0, 9), np.float32)
    gpu = gpu_trace(gpu_ctx, empty, rays)
    assert (gpu["prim_id"] == N.INVALID_ID).all() and (gpu["geom_id"] == N.INVALID_ID).all()
    assert np.isinf(gpu["t"]).all() and (gpu["occluded"] == 0).all()
    one = np.array([[0, 0, 0.5, 1, 0, 0.5, 0, 1, 0.5]], np.float32)
    gpu = gpu_trace(gpu_ctx, one, rays)
    ref = O.closest_brute_f32(one, rays)
    assert (ref["prim"] == 0).sum() > 10
    assert_same(gpu, ref, [0, 1])


def test_transform_places_instance(gpu_ctx):
    # a unit triangle moved by a column-major transform (scale 2, translate z by 3) equals the
    # pre-transformed soup
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    m = np.zeros((4, 4), np.float32)  # m[col][row]
    m[0, 0] = m[1, 1] = m[2, 2] = 2.0
    m[3, 3] = 1.0
    m[3, 2] = 3.0
    scene = N.Scene(gpu_ctx)
    scene.add_triangles(pos, np.array([0, 1, 2], np.uint32), m.reshape(-1))
    scene.commit()
    rays = {"org_x": np.array([0.5], np.float32), "org_y": np.array([0.5], np.float32),
            "org_z": np.array([0], np.float32), "dir_x": np.zeros(1, np.float32), "dir_y": np.zeros(1, np.float32),
            "dir_z": np.ones(1, np.float32), "tnear": np.zeros(1, np.float32),
            "tfar": np.full(1, np.inf, np.float32)}
    dr = N.DeviceRays(gpu_ctx, rays)
    dh = N.DeviceHits(gpu_ctx, 1)
    scene.intersect(dr, dh)
    h = dh.to_host()
    assert h["prim_id"][0] == 0 and h["t"][0] == 3.0
    assert h["u"][0] == 0.25 and h["v"][0] == 0.25
    scene.release()


def test_medium_scene_matches_oracle_tree(gpu_ctx):
    t9 = W.traversal_scene(scale=1 / 32, seed=9)  # 65,536 triangles
    rays = W.traversal_rays(t9, 128, 128, 2, n_bounce=32768, seed=9)
    gpu = gpu_trace(gpu_ctx, t9, rays)
    bvh = O.CpuBvh(t9)
    ref = bvh.closest(rays)
    assert (ref["prim"] != N.INVALID_ID).sum() > 2000
    assert_same(gpu, ref, [0, t9.shape[0]])
    assert np.array_equal(gpu["occluded"], bvh.occluded(rays))


def test_full_traversal_workload(gpu_ctx):
    """2,097,152 triangles, 16,777,216 rays: exact match with the oracle tree on a strided
    subsample; any-hit/closest-hit consistency on every ray."""
    t9 = W.traversal_scene()
    rays = W.traversal_rays(t9)
    n = rays["org_x"].shape[0]
    assert t9.shape[0] == 2097152 and n == 16777216
    scene = N.Scene(gpu_ctx)
    scene.add_soup(t9)
    scene.commit()
    info = scene.info()
    assert info["n_tris"] == 2097152 and info["n_nodes"] == 2097151
    dr = N.DeviceRays(gpu_ctx, rays)
    dh = N.DeviceHits(gpu_ctx, n, with_ng=False)
    scene.intersect(dr, dh)
    h = dh.to_host()
    hit = h["prim_id"] != N.INVALID_ID
    assert 0.05 < hit.mean() < 0.95
    # any hit with tfar = +inf: occluded exactly where a closest hit exists
    occ = gpu_ctx.empty(4 * n)
    scene.occluded(dr, occ)
    o1 = occ.download(np.uint32, n)
    assert np.array_equal(o1, hit.astype(np.uint32))
    # any hit with tfar = 0.5 t on a seeded half: never occluded there
    tfar, expected = W.anyhit_tfar(h["t"], h["prim_id"])
    dr.set_field("tfar", tfar)
    scene.occluded(dr, occ)
    o2 = occ.download(np.uint32, n)
    assert np.array_equal(o2, expected)
    N.synchronize()
    scene.check()
    # bit-exact against the oracle's own tree on every 97th ray (173k rays)
    sel = np.arange(0, n, 97)
    sub = {k: np.ascontiguousarray(v[sel]) for k, v in rays.items()}
    ref = O.CpuBvh(t9).closest(sub)
    got = {"prim_id": h["prim_id"][sel], "geom_id": h["geom_id"][sel], "t": h["t"][sel], "u": h["u"][sel],
           "v": h["v"][sel]}
    assert_same(got, ref, [0, t9.shape[0]])
    scene.release()


@pytest.mark.parametrize("seed", [4, 5])
def test_wide_and_binary_trees_agree(gpu_ctx, seed):
    """The 8-wide compressed tree (default) and the binary tree it is collapsed from give the same
    hits, bit for bit (both only cull; hits are the triangle test's), with far fewer node visits."""
    t9 = W.traversal_scene(scale=1 / 16, seed=seed)  # 131,072 triangles in Gaussian clusters
    rays = W.traversal_rays(t9, 128, 128, 2, n_bounce=32768, seed=seed)
    old = gpu_ctx.tuning()
    try:
        out, stats = {}, {}
        for tree in (0, 1, 2):  # 2: the 8-wide tree with its top nodes read from LDS
            gpu_ctx.set_tuning(trace_tree=min(tree, 1), trace_lds_nodes=9 if tree == 2 else 0)
            out[tree] = gpu_trace(gpu_ctx, t9, rays)
            if tree == 2:
                continue
            scene = N.Scene(gpu_ctx)
            scene.add_soup(t9)
            scene.commit()
            stats[tree] = scene.intersect_stats(N.DeviceRays(gpu_ctx, rays))
            scene.release()
        for other in (1, 2):
            for k in ("prim_id", "geom_id", "occluded"):
                assert np.array_equal(out[0][k], out[other][k]), (other, k)
            for k in ("t", "u", "v"):
                assert np.array_equal(out[0][k].view(np.uint32), out[other][k].view(np.uint32)), (other, k)
        assert stats[0]["bytes_node"] == 64 and stats[1]["bytes_node"] == 128
        assert stats[1]["nodes_visited"] < 0.6 * stats[0]["nodes_visited"]
    finally:
        gpu_ctx.set_tuning(**old)


@pytest.mark.parametrize("builder", [0, 1])
def test_both_tree_builders_match_brute_force(gpu_ctx, builder):
    """The binary tree under the 8-wide one built as a linear BVH (0) or by locally-ordered
    clustering (1): the same bit-exact hits and occlusion flags against FP32 brute force, on a soup
    with coincident and degenerate triangles (equal centroids, zero-area slivers) and 2 geometries.
    Prints the nodes visited per closest-hit ray of each tree."""
    old = gpu_ctx.tuning()
    try:
        gpu_ctx.set_tuning(bvh_builder=builder)
        rng = np.random.default_rng(41)
        t9 = dense_soup(7, 32, 64)
        dup = t9[:200].copy()  # coincident copies: equal Morton codes, equal boxes
        sliver = t9[200:260].copy()
        sliver[:, 6:9] = sliver[:, 0:3] + 1e-7 * rng.normal(size=(60, 3)).astype(np.float32)  # near-degenerate
        tri9 = np.concatenate([t9, dup, sliver]).astype(np.float32)
        bounds = [0, 1000, tri9.shape[0]]
        cam = N.gen_camera_rays(W.CAMERA_POS, W.CAMERA_LOOK_AT, W.CAMERA_UP, 50.0, 48, 48, 1)
        bnc = N.gen_bounce_rays(8, tri9, 3000)
        rays = {k: np.concatenate([cam[k], bnc[k]]) for k in N.RAY_FIELDS}
        print(f"builder {builder}: {tri9.shape[0]} triangles, {rays['org_x'].size} rays", flush=True)
        gpu = gpu_trace(gpu_ctx, tri9, rays, bounds)
        assert_same(gpu, O.closest_brute_f32(tri9, rays), bounds)
        assert np.array_equal(gpu["occluded"], O.occluded_brute_f32(tri9, rays))
        scene = N.Scene(gpu_ctx)
        scene.add_soup(tri9)
        scene.commit()
        st = scene.intersect_stats(N.DeviceRays(gpu_ctx, rays))
        scene.release()
        print(f"builder {builder}: {st['nodes_visited'] / st['rays']:.2f} nodes/ray, depth {gpu['info']['max_depth']}",
              flush=True)
        assert st["nodes_visited"] > 0
    finally:
        gpu_ctx.set_tuning(**old)
