"""GPU parity of instanced scenes (sgpu_scene_add_instance: Embree's RTC_GEOMETRY_TYPE_INSTANCE and
Selas's subscene instances) against the oracle, through the C ABI.

Bar: bit-exact t, u, v, primID, geomID, instID[0], instID[1] and occlusion flags against the
oracle's brute force over the composed instances (whose world-to-local transforms are the
library's, restated in tests/instancing.py), geometric normals in the placed scene's space within
float rounding; errors for the forms the library refuses.
"""
import numpy as np
import pytest

import instancing as I
from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _soup(seed, n, spread=1.0, size=0.25):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    return (c + rng.normal(scale=size, size=(n, 3, 3))).reshape(n, 9).astype(np.float32)


def _quads(seed, nu=12, nv=9):
    rng = np.random.default_rng(seed)
    xs, zs = np.linspace(-1, 1, nu + 1), np.linspace(-1, 1, nv + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    P = np.stack([X, 0.2 * rng.normal(size=X.shape), Z], -1).reshape(-1, 3).astype(np.float32)
    q = []
    for i in range(nu):
        for j in range(nv):
            a = i * (nv + 1) + j
            q.append([a, a + nv + 1, a + nv + 2, a + 1])
    return P, np.array(q, np.uint32)


def _rays(seed, n, spread=9.0, target=4.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    d = rng.uniform(-target, target, (n, 3)) - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tfar = np.full(n, np.inf, np.float32)
    tfar[::3] = rng.uniform(1.0, 15.0, tfar[::3].shape[0]).astype(np.float32)
    return {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": d[:, 0], "dir_y": d[:, 1],
            "dir_z": d[:, 2], "tnear": np.zeros(n, np.float32), "tfar": tfar}


def _gpu(ctx, scene, rays):
    dr = N.DeviceRays(ctx, rays)
    dh = N.DeviceHits(ctx, dr.n, with_inst=True)
    scene.intersect(dr, dh)
    occ = ctx.empty(4 * dr.n)
    scene.occluded(dr, occ)
    N.synchronize()
    scene.check()
    out = dh.to_host()
    out["occluded"] = occ.download(np.uint32, dr.n)
    return out


def _assert_same(gpu, ref):
    for k in ("prim_id", "geom_id", "inst_id0", "inst_id1"):
        bad = np.nonzero(gpu[k] != ref[k])[0]
        assert bad.size == 0, f"{k}: {bad.size} mismatches, first {bad[:5]}"
    for k in ("t", "u", "v"):
        a, b = gpu[k].view(np.uint32), ref[k].view(np.uint32)
        assert np.array_equal(a, b), f"{k}: {np.count_nonzero(a != b)} bit mismatches"
    assert np.array_equal(gpu["occluded"], ref["occluded"])
    hit = ref["prim_id"] != N.INVALID_ID
    got = np.stack([gpu["ng_x"][hit], gpu["ng_y"][hit], gpu["ng_z"][hit]], 1)
    np.testing.assert_allclose(got, ref["ng"][hit], rtol=1e-5, atol=1e-7)


@pytest.fixture(scope="module")
def models():
    a = I.Model([("soup", _soup(1, 400)), ("soup", _soup(2, 1)), ("soup", _soup(3, 150, spread=0.6))])
    p, q = _quads(4)
    b = I.Model([("quads", p, q), ("soup", _soup(5, 50))])
    return a, b


@pytest.mark.parametrize("seed", [1, 2])
def test_two_level_instances_match_oracle(gpu_ctx, models, seed):
    rng = np.random.default_rng(100 + seed)
    ma, mb = models
    da, db = ma.device(gpu_ctx), mb.device(gpu_ctx)
    # a subscene placing both models, instanced twice in the scene next to a direct model instance
    sub_desc = [(ma, I.random_transform(rng, spread=1.5)), (mb, I.random_transform(rng, spread=1.5)),
                (ma, I.random_transform(rng, spread=1.5))]
    top_desc = [(sub_desc, I.random_transform(rng)), (mb, I.random_transform(rng)), (sub_desc, I.random_transform(rng)),
                (ma, I.random_transform(rng))]
    sub = N.Scene(gpu_ctx)
    for m, x in sub_desc:
        sub.add_instance(da if m is ma else db, np.asarray(x, np.float32).reshape(4, 4).T)
    sub.commit()
    top = N.Scene(gpu_ctx)
    ids = []
    for c, x in top_desc:
        child = sub if isinstance(c, list) else (da if c is ma else db)
        ids.append(top.add_instance(child, np.asarray(x, np.float32).reshape(4, 4).T))
    assert ids == list(range(len(top_desc)))
    top.commit()
    try:
        rays = _rays(seed, 20000)
        print(f"instances seed={seed}", flush=True)
        gpu = _gpu(gpu_ctx, top, rays)
        ref = I.oracle_hits(I.compose(top_desc), rays)
        assert (ref["prim_id"] != N.INVALID_ID).mean() > 0.2
        assert (ref["inst_id1"] != N.INVALID_ID).any() and (ref["inst_id1"] == N.INVALID_ID).any()
        _assert_same(gpu, ref)
        # one level: the subscene queried on its own
        gsub = _gpu(gpu_ctx, sub, rays)
        rsub = I.oracle_hits(I.compose(sub_desc), rays)
        _assert_same(gsub, rsub)
        assert (gsub["inst_id1"] == N.INVALID_ID).all()
        st = top.intersect_stats(N.DeviceRays(gpu_ctx, rays))
        assert st["rays"] == 20000 and st["nodes_visited"] > 0
    finally:
        top.release()
        sub.release()
        da.release()
        db.release()


def test_many_instances_of_one_model(gpu_ctx):
    rng = np.random.default_rng(7)
    m = I.Model([("soup", _soup(8, 2000, spread=1.0, size=0.1))])
    dm = m.device(gpu_ctx)
    desc = [(m, I.random_transform(rng, scale=(0.05, 0.2), spread=8.0)) for _ in range(300)]
    top = N.Scene(gpu_ctx)
    for _, x in desc:
        top.add_instance(dm, np.asarray(x, np.float32).reshape(4, 4).T)
    top.commit()
    try:
        rays = _rays(9, 20000, spread=12.0, target=8.0)
        print("300 instances", flush=True)
        gpu = _gpu(gpu_ctx, top, rays)
        ref = I.oracle_hits(I.compose(desc), rays)
        assert (ref["prim_id"] != N.INVALID_ID).mean() > 0.05
        _assert_same(gpu, ref)
        info = top.info()
        assert info["n_tris"] == 0 and info["n_nodes"] == 299
    finally:
        top.release()
        dm.release()


def test_flat_scenes_report_no_instance(gpu_ctx):
    t9 = _soup(10, 500)
    s = N.Scene(gpu_ctx)
    s.add_soup(t9)
    s.commit()
    try:
        out = _gpu(gpu_ctx, s, _rays(11, 4000, spread=3.0, target=1.0))
        assert (out["prim_id"] != N.INVALID_ID).any()
        assert (out["inst_id0"] == N.INVALID_ID).all() and (out["inst_id1"] == N.INVALID_ID).all()
    finally:
        s.release()


def test_instancing_errors(gpu_ctx, models):
    ma, _ = models
    da = ma.device(gpu_ctx)
    eye = np.eye(4)
    g1, g2, g3 = N.Scene(gpu_ctx), N.Scene(gpu_ctx), N.Scene(gpu_ctx)
    try:
        with pytest.raises(N.SgpuError):
            da.add_instance(da, eye)  # committed, and a scene cannot place itself
        fresh = N.Scene(gpu_ctx)
        fresh.add_soup(_soup(12, 3))
        with pytest.raises(N.SgpuError):
            fresh.add_instance(da, eye)  # holds geometries
        with pytest.raises(N.SgpuError):
            g1.add_instance(fresh, eye)  # child not committed
        fresh.release()
        g1.add_instance(da, eye)
        with pytest.raises(N.SgpuError):
            g1.add_soup(_soup(13, 3))  # holds instances
        g1.commit()
        g2.add_instance(g1, eye)
        g2.commit()
        with pytest.raises(N.SgpuError):
            g3.add_instance(g2, eye)  # a third level
        with pytest.raises(N.SgpuError):
            da.release()  # placed by g1
        singular = np.eye(4)
        singular[1, 1] = 0.0
        g3.add_instance(da, singular)
        with pytest.raises(N.SgpuError):
            g3.commit()
    finally:
        g3.release()
        g2.release()
        g1.release()
        da.release()


def _close(gpu, cpu):
    close = np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5
    rmse = np.sqrt(np.mean((gpu - cpu) ** 2))
    return close.mean(), rmse / max(cpu.mean(), 1e-12)


@pytest.mark.parametrize("ibl", [False, True])
def test_instanced_renders_match_oracle(gpu_ctx, ibl):
    """The path tracer and VCM on a scene with two instance levels (placed scenes with Lambert,
    Disney and textured materials): per pixel against the oracle (99 % within 1e-3, VCM RMSE
    <= 2e-3 of the mean)."""
    import oracle as O
    import scenes
    sc = scenes.instanced_scene(ibl=ibl)
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        rs = O.InstancedRenderScene(sc)
        print(f"instanced render ibl={ibl}", flush=True)
        g = N.render_pt(scene, sc["camera"], 64, 48, 4, 8, 5, N.PT_NEE_COSINE_ONCE)
        c = O.render_pt(rs, sc["camera"], 64, 48, 4, 8, 5, 1)
        frac, _ = _close(g, c)
        assert c[1].mean() > 0.01 and frac >= 0.99, frac
        g = N.render_vcm(scene, sc["camera"], 64, 48, 3, 8, 0.02, 0.75, 17)
        c, _ = O.render_vcm(rs, sc["camera"], 64, 48, 3, 8, 0.02, 0.75, 17)
        frac, rel_rmse = _close(g, c)
        assert c.mean() > 0.01 and frac >= 0.99 and rel_rmse <= 2e-3, (frac, rel_rmse)
    finally:
        scene.release()


def test_batch_mixes_instanced_and_flat_scenes(gpu_ctx):
    import scenes
    inst = scenes.instanced_scene(ibl=True)
    flat = scenes.cube_scene(lights=True)
    si = N.build_render_scene(gpu_ctx, inst)
    sf = N.build_render_scene(gpu_ctx, flat)
    try:
        jobs = [(si, inst["camera"], 40, 32, 2, 21), (sf, flat["camera"], 48, 24, 3, 22), (si, inst["camera"], 24, 24, 1, 23)]
        alone = [N.render_vcm(s, cam, w, h, its, 10, 0.02, 0.75, seed) for s, cam, w, h, its, seed in jobs]
        batch = N.render_vcm_batch([dict(scene=s, camera=cam, width=w, height=h, iterations=its, max_path_length=10,
                                         base_radius=0.02, radius_alpha=0.75, seed=seed)
                                    for s, cam, w, h, its, seed in jobs])
        for a, b in zip(alone, batch):
            frac, rel_rmse = _close(b, a)
            assert a.mean() > 0.005 and frac >= 0.99 and rel_rmse <= 2e-3, (frac, rel_rmse)
    finally:
        si.release()
        sf.release()
