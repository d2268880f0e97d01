"""Pins the oracle's vertex-attribute interpolation and shading frames (rtcInterpolate0 in
CalculateSurfaceParams, SurfaceParameters.cpp:155-190; Embree 3's interpolation rule) with what
the mathematics fixes:

* a field linear in position, given at the vertices, interpolates to its value at the hit point
  (any correct barycentric weighting does; swapped or misassigned weights do not);
* on a twisted (non-planar) quad the hit lies on one triangle of Embree's split, and the
  interpolated linear field equals the field at the hit point, which a bilinear patch rule would
  not give;
* vertex normals equal to the face normals, and tangents equal to the frame MakeOrthogonalCoordinate-
  System builds (w = 1), render the image of the plain geometry; tilted normals do not.

Embree itself cannot run here (headers only), so agreement with Embree's compiled interpolation is
parity unpinned beyond these properties.
"""
import numpy as np

import oracle as O
from selas_amd import workloads as W


def test_linear_field_triangle():
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = rng.normal(size=(3, 3))
        A = rng.normal(size=(4, 3))
        c = rng.normal(size=4)
        corner = (p @ A.T + c).astype(np.float32)  # a 4-float attribute per corner
        u, v = rng.random(2) * 0.5
        got = O.interp_attr(corner.reshape(-1), 4, 3, 0, u, v)
        x = (1 - u - v) * p[0] + u * p[1] + v * p[2]
        want = A @ x + c
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(corner).max())
        # sensitivity: swapped weights give another value
        swapped = A @ ((1 - u - v) * p[0] + v * p[1] + u * p[2]) + c
        if abs(u - v) > 0.05:
            assert np.abs(swapped - want).max() > 1e-4 * (1 + np.abs(want).max())


def test_primitive_offset():
    """prim selects corners prim * corners .. + corners - 1."""
    vals = np.arange(2 * 3 * 2, dtype=np.float32)  # 2 triangles, 3 corners, 2 floats
    got = O.interp_attr(vals, 2, 3, 1, 0.0, 0.0)
    assert np.array_equal(got, vals[6:8])
    got = O.interp_attr(vals, 2, 3, 1, 1.0, 0.0)
    assert np.array_equal(got, vals[8:10])
    got = O.interp_attr(vals, 2, 3, 1, 0.0, 1.0)
    assert np.array_equal(got, vals[10:12])


def test_twisted_quad_follows_the_hit_triangle():
    rng = np.random.default_rng(2)
    # a strongly twisted quad (v0, v1, v2, v3) and Embree's split (v0, v1, v3), (v2, v3, v1)
    q = np.array([[0, 0, 0], [1, 0, 0.6], [1, 1, -0.5], [0, 1, 0.4]], np.float64)
    tri9 = np.stack([np.concatenate([q[0], q[1], q[3]]), np.concatenate([q[2], q[3], q[1]])]).astype(np.float32)
    A = rng.normal(size=(3, 3))
    c = rng.normal(size=3)
    corner = (q @ A.T + c).astype(np.float32).reshape(-1)
    bvh = O.CpuBvh(tri9)
    n = 400
    tx = rng.uniform(0.05, 0.95, (n, 2))
    o = np.stack([tx[:, 0], tx[:, 1], np.full(n, 3.0)], 1).astype(np.float32)
    d = np.tile(np.array([0, 0, -1], np.float32), (n, 1))
    rays = {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": d[:, 0], "dir_y": d[:, 1],
            "dir_z": d[:, 2], "tnear": np.zeros(n, np.float32), "tfar": np.full(n, np.inf, np.float32)}
    h = bvh.closest(rays)
    assert (h["prim"] != 0xFFFFFFFF).all()
    bilinear_gap = 0.0
    for i in range(n):
        tri, bu, bv = int(h["prim"][i]), h["u"][i], h["v"][i]
        # the quad coordinates a hit reports: (u, v) on the first triangle, (1 - u, 1 - v) on the second
        qu, qv = (bu, bv) if tri == 0 else (np.float32(1) - bu, np.float32(1) - bv)
        got = O.interp_attr(corner, 3, 4, 0, qu, qv)
        x = o[i].astype(np.float64) + h["t"][i] * d[i]
        np.testing.assert_allclose(got, A @ x + c, rtol=2e-5, atol=2e-5 * np.abs(corner).max())
        # the bilinear patch through the four corners disagrees away from the diagonal
        qu64, qv64 = float(qu), float(qv)
        bl = ((1 - qu64) * (1 - qv64) * q[0] + qu64 * (1 - qv64) * q[1] + qu64 * qv64 * q[2] + (1 - qu64) * qv64 * q[3])
        bilinear_gap = max(bilinear_gap, float(np.abs(A @ bl + c - (A @ x + c)).max()))
    assert bilinear_gap > 0.05


def _panel_scene(normals=None, tangents=None):
    """A Lambert/Disney floor panel (a flat 4 x 3 quad grid) in a small box under a quad light."""
    sc = W.attributed_room_scene(1)
    panel = W.attributed_panel(4, 3, (0.1, 0.02, 0.2), (0.0, 0.0, 0.6), (0.8, 0.0, 0.0), 5)
    # flatten the panel (no bump) so that the face normals are the plane's
    pos = panel["positions"].copy()
    pos[:, 1] = 0.02
    panel["positions"] = pos
    if normals is None:
        for k in ("normals", "tangents", "uvs"):
            panel.pop(k)
    else:
        panel["normals"] = normals(pos.shape[0])
        panel["tangents"] = tangents(pos.shape[0]) if tangents else None
        panel["uvs"] = None
    geoms = sc["geoms"][:12] + [panel]
    mats = list(sc["materials"][:12]) + [W.material(1, (0.7, 0.7, 0.7), 0.0, 0.25, anisotropic=0.6)]
    return dict(sc, geoms=geoms, materials=np.stack(mats), geom_material=np.arange(len(geoms), dtype=np.uint32))


def _make_frame_t(n):
    n = np.asarray(n, np.float32)
    if abs(n[0]) > abs(n[1]):
        return np.array([-n[2], 0, n[0]], np.float32) / np.sqrt(n[0] * n[0] + n[2] * n[2])
    return np.array([0, n[2], -n[1]], np.float32) / np.sqrt(n[1] * n[1] + n[2] * n[2])


def test_face_normals_and_frame_tangents_render_the_plain_image():
    up = np.array([0, 1, 0], np.float32)
    t = _make_frame_t(up)
    plain = _panel_scene()
    same = _panel_scene(lambda k: np.tile(up, (k, 1)), lambda k: np.tile(np.append(t, 1.0), (k, 1)).astype(np.float32))
    tilted = _panel_scene(lambda k: np.tile(np.array([0.3, 1, 0.2], np.float32) / np.float32(np.sqrt(1.13)), (k, 1)),
                          lambda k: np.tile(np.array([1, 0, 0, 1], np.float32), (k, 1)))
    cam = plain["camera"]
    imgs = [O.render_pt(O.RenderScene(s), cam, 48, 40, 8, 6, 9)[1] for s in (plain, same, tilted)]
    a, b, c = imgs
    assert a.mean() > 0.01
    assert np.mean(np.abs(a - b) <= 1e-3 * np.abs(a) + 1e-5) >= 0.99
    assert abs(b.mean() / a.mean() - 1) < 1e-3
    # the shading frame is used: tilted normals change the image
    assert np.mean(np.abs(a - c) <= 1e-3 * np.abs(a) + 1e-5) < 0.9
    # VCM as well
    v = [O.render_vcm(O.RenderScene(s), cam, 32, 24, 2, 6, 0.01, 0.75, 4)[0] for s in (plain, same)]
    assert np.mean(np.abs(v[0] - v[1]) <= 1e-3 * np.abs(v[0]) + 1e-5) >= 0.99


def test_attributed_scene_renders():
    """The attributed room through both integrators: finite, lit, and different from the same
    geometry shaded with face normals."""
    sc = W.attributed_room_scene(1)
    flat = dict(sc, geoms=[{k: v for k, v in g.items() if k not in ("normals", "tangents", "uvs")}
                           if isinstance(g, dict) else g for g in sc["geoms"]])
    rs, rf = O.RenderScene(sc), O.RenderScene(flat)
    a = O.render_pt(rs, sc["camera"], 40, 32, 4, 6, 2)
    b = O.render_pt(rf, sc["camera"], 40, 32, 4, 6, 2)
    assert np.isfinite(a).all() and a[1].mean() > 0.01
    assert np.mean(np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-5) < 0.99
    v, _ = O.render_vcm(rs, sc["camera"], 40, 32, 2, 8, 0.005, 0.75, 3)
    assert np.isfinite(v).all() and v.mean() > 0.01
