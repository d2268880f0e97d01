"""The oracle's instanced traversal (oracle/orc_trace.c, orc_inst_closest / orc_inst_occluded)
pinned on the CPU.

Embree cannot run here, so instancing is pinned by what it means: an instance placed by an
identity transform hits exactly what its scene hits; a transformed instance hits what the
transformed triangles hit (same triangle, same instance, t within float rounding of the double
brute force, for rays whose hit is not within rounding of another); identical instances tie to
the earlier one; an occlusion query agrees with the closest hit; and the composition helpers the
GPU tests use (the library's double-precision product and inverse, restated in tests/instancing.py)
invert each other.
"""
import numpy as np

import oracle as O
import instancing as I
from selas_amd import native as N


def _soup(seed, n=600, spread=1.5, size=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    v = c + rng.normal(scale=size, size=(n, 3, 3))
    return v.reshape(n, 9).astype(np.float32)


def _rays(seed, n, spread=6.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    target = rng.uniform(-2.0, 2.0, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    return {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": d[:, 0], "dir_y": d[:, 1],
            "dir_z": d[:, 2], "tnear": np.zeros(n, np.float32), "tfar": np.full(n, np.inf, np.float32)}


def _world(model_tri9, l2w_rows):
    m = np.array(l2w_rows, np.float64).reshape(3, 4)
    p = model_tri9.reshape(-1, 3).astype(np.float64)
    w = p @ m[:, :3].T + m[:, 3]
    return w.reshape(-1, 9).astype(np.float32)


def test_identity_instance_hits_what_its_scene_hits():
    t9 = _soup(1)
    bvh = O.CpuBvh(t9)
    rays = _rays(2, 3000)
    eye = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    got = O.InstancedScene([(eye, bvh)]).closest(rays)
    ref = bvh.closest(rays)
    hit = ref["prim"] != N.INVALID_ID
    assert hit.mean() > 0.3
    for k in ("t", "u", "v", "prim"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert (got["inst"][hit] == 0).all()


def test_transformed_instances_hit_the_transformed_triangles():
    rng = np.random.default_rng(3)
    t9 = _soup(4, n=300, spread=1.0, size=0.25)
    model = I.Model([("soup", t9)])
    top = [(model, I.random_transform(rng, spread=3.0)) for _ in range(5)]
    composed = I.compose(top)
    rays = _rays(5, 4000)
    got = O.InstancedScene([(w, model.bvh) for w, _, _, _ in composed]).closest(rays)
    world = np.concatenate([_world(t9, I.rows_from_colmajor(m)) for _, m in top])
    ref = O.closest_brute_f64(world, rays)
    hit = ref["prim"] != N.INVALID_ID
    assert hit.mean() > 0.2
    # rays whose closest hit is clear of every other candidate by more than float rounding
    clear = hit & (ref["margin"] > 1e-4)
    assert clear.sum() > 0.8 * hit.sum()
    n_tri = t9.shape[0]
    assert np.array_equal(got["inst"][clear], ref["prim"][clear] // n_tri)
    assert np.array_equal(got["prim"][clear], ref["prim"][clear] % n_tri)
    np.testing.assert_allclose(got["t"][clear], ref["t"][clear], rtol=2e-5, atol=1e-6)
    # misses agree away from triangle edges
    assert ((got["prim"] == N.INVALID_ID) == ~hit).mean() > 0.995


def test_identical_instances_tie_to_the_earlier():
    t9 = _soup(6)
    bvh = O.CpuBvh(t9)
    eye = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    got = O.InstancedScene([(eye, bvh), (eye, bvh), (eye, bvh)]).closest(_rays(7, 2000))
    hit = got["prim"] != N.INVALID_ID
    assert hit.mean() > 0.3 and (got["inst"][hit] == 0).all()


def test_occlusion_agrees_with_closest_hit():
    rng = np.random.default_rng(8)
    model = I.Model([("soup", _soup(9, n=200))])
    composed = I.compose([(model, I.random_transform(rng, spread=2.0)) for _ in range(4)])
    sc = O.InstancedScene([(w, model.bvh) for w, _, _, _ in composed])
    rays = _rays(10, 3000)
    rays["tfar"] = np.random.default_rng(11).uniform(0.5, 12.0, 3000).astype(np.float32)
    c = sc.closest(rays)
    occ = sc.occluded(rays)
    assert np.array_equal(occ == 1, c["prim"] != N.INVALID_ID)
    assert 0.1 < occ.mean() < 0.9


def test_composition_helpers_invert():
    rng = np.random.default_rng(12)
    for _ in range(20):
        a = I.rows_from_colmajor(I.random_transform(rng))
        b = I.rows_from_colmajor(I.random_transform(rng))
        ab = I.affine_mul(a, b)
        inv = I.affine_inverse(ab)
        m = np.vstack([np.array(ab).reshape(3, 4), [0, 0, 0, 1]])
        mi = np.vstack([np.array(inv).reshape(3, 4), [0, 0, 0, 1]])
        np.testing.assert_allclose(m @ mi, np.eye(4), atol=1e-12)
        ma = np.vstack([np.array(a).reshape(3, 4), [0, 0, 0, 1]])
        mb = np.vstack([np.array(b).reshape(3, 4), [0, 0, 0, 1]])
        np.testing.assert_allclose(m, ma @ mb, atol=1e-12)


def test_instanced_render_equals_the_flattened_scene():
    """Rendering through two instance levels (rays moved into each placed scene, normals carried
    back by the instance's linear part, geometries numbered as one list) gives the image of the same
    geometry moved to world space, up to the float rounding of the two placements (which moves a few
    hits): 99 % of pixel values within 1e-3, means within 0.2 %."""
    import scenes
    sc = scenes.instanced_scene(ibl=True)
    rs = O.InstancedRenderScene(sc)
    rf = O.RenderScene(scenes.flatten_instanced(sc))
    for a, b in [(O.render_pt(rs, sc["camera"], 64, 48, 16, 8, 3, 1)[1], O.render_pt(rf, sc["camera"], 64, 48, 16, 8, 3, 1)[1]),
                 (O.render_vcm(rs, sc["camera"], 64, 48, 8, 8, 0.02, 0.75, 3)[0],
                  O.render_vcm(rf, sc["camera"], 64, 48, 8, 8, 0.02, 0.75, 3)[0])]:
        assert b.mean() > 0.01
        assert np.mean(np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5) >= 0.99
        assert abs(a.mean() / b.mean() - 1) < 2e-3
