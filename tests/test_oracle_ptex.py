"""The per-face texture filter of the oracle (oracle/orc_ptex.c) pinned on the CPU.

* Bit-exact against PtexFilter{f_bspline}.eval of the reference's own Ptex reader and filters
  (Middleware/ptex/Include, compiled by oracle/ref.mk) on Ptex files written by
  tests/golden/ptx.py: a 5 x 4 quad grid at 16 x 16 texels (borders, valence-4 corners), a cube at
  8 x 8 (neighbours rotated by every amount, valence-3 corners) and five quads around a valence-5
  vertex at 4 x 4 (the wide-footprint branch of the kernel builder). Vectors recorded by
  tests/golden/make_golden.py.
* What the filter's mathematics fixes: the cubic B-spline weights sum to one and reproduce linear
  functions, so a constant texture filters to its value and a texture linear in the grid's texel
  coordinate filters to that linear function wherever the 4 x 4 footprint stays inside the mesh.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["grid", "cube", "star"])
def test_ptex_matches_reference(name):
    g = np.load(os.path.join(GOLD, f"ptex_{name}.npz"))
    tex = O.PtexTexture(g["texels"], int(g["res_log2"]), g["adjfaces"], g["adjedges"])
    out = tex.eval(g["faces"], g["u"], g["v"])
    assert np.array_equal(out.view(np.uint32), g["out"].view(np.uint32))
    # the recorded points reach every face edge and corner
    assert (g["u"] == 0).any() and (g["u"] == 1).any() and (g["v"] == 0).any() and (g["v"] == 1).any()


def _grid(nu, nv):
    import sys

    sys.path.insert(0, GOLD)
    import ptx

    return ptx.grid_adjacency(nu, nv)


def test_constant_texture_is_reproduced():
    af, ae = _grid(3, 2)
    tex = O.PtexTexture(np.full((6, 16, 16, 3), 77, np.uint8), 4, af, ae)
    rng = np.random.default_rng(1)
    out = tex.eval(rng.integers(0, 6, 500), rng.random(500), rng.random(500))
    np.testing.assert_allclose(out, 77 / 255, rtol=2e-6)


def test_linear_texture_is_reproduced():
    """Texel (i, j) of face (fi, fj) holds a + b * (global u texel index) in channel 0 and
    a + c * (global v texel index) in channel 1; away from the mesh border the filter returns the
    linear function at the query point."""
    nu, nv, res = 4, 3, 16
    af, ae = _grid(nu, nv)
    tex = np.zeros((nu * nv, res, res, 3), np.uint8)
    for fj in range(nv):
        for fi in range(nu):
            f = fj * nu + fi
            gi = fi * res + np.arange(res)
            gj = fj * res + np.arange(res)
            tex[f, :, :, 0] = (10 + 3 * gi)[None, :]
            tex[f, :, :, 1] = (20 + 4 * gj)[:, None]
            tex[f, :, :, 2] = 99
    t = O.PtexTexture(tex, 4, af, ae)
    rng = np.random.default_rng(2)
    n = 2000
    faces = rng.integers(0, nu * nv, n)
    u, v = rng.random(n), rng.random(n)
    gu = (faces % nu) * res + u * res - 0.5  # continuous global texel coordinate
    gv = (faces // nu) * res + v * res - 0.5
    inside = (gu > 1.5) & (gu < nu * res - 2.5) & (gv > 1.5) & (gv < nv * res - 2.5)
    out = t.eval(faces, u, v) * 255
    np.testing.assert_allclose(out[inside, 0], 10 + 3 * gu[inside], atol=2e-3)
    np.testing.assert_allclose(out[inside, 1], 20 + 4 * gv[inside], atol=2e-3)
    np.testing.assert_allclose(out[:, 2], 99, rtol=1e-5)
    assert inside.mean() > 0.5


# ---- triangle meshes: PtexTriangleFilter -------------------------------------------------------------

TRI_MESHES = ["tgrid", "tgrid_rot", "tetra", "icosa", "fan", "tgrid_r1", "tgrid_r0"]


def _tri(name):
    g = np.load(os.path.join(GOLD, f"ptex_tri_{name}.npz"))
    return g, O.PtexTexture(g["texels"], int(g["res_log2"]), g["adjfaces"], g["adjedges"], tri=True)


@pytest.mark.parametrize("name", TRI_MESHES)
def test_tri_ptex_matches_reference(name):
    """Bit-exact against PtexTriangleFilter (served by getFilter for triangle meshes) on a split
    grid, a tetrahedron (valence 3), an icosahedron (valence 5), an open fan (valence 7) and the
    grid at 2 x 2 and 1 x 1 texels, where the kernel is wider than a face: with the
    single-precision expf the reference's exp call resolves to."""
    g, tex = _tri(name)
    O.set_libm_single(True)
    try:
        out = tex.eval(g["faces"], g["u"], g["v"])
    finally:
        O.set_libm_single(False)
    assert np.array_equal(out.view(np.uint32), g["out"].view(np.uint32))


@pytest.mark.parametrize("name", TRI_MESHES)
def test_tri_ptex_device_exp_within_two_ulp(name):
    """With exp evaluated in double and rounded once (the device's way), results differ from the
    reference's where glibc's expf is not correctly rounded: at most 2 ulp, and rarely."""
    g, tex = _tri(name)
    out = tex.eval(g["faces"], g["u"], g["v"])
    ulps = np.abs(out.view(np.int32).astype(np.int64) - g["out"].view(np.int32).astype(np.int64))
    assert ulps.max() <= 2 and (ulps == 0).mean() >= 0.995


def test_tri_constant_texture_is_reproduced():
    """The Gaussian's weights are normalised by their sum: a constant texture filters to its value."""
    import sys

    sys.path.insert(0, GOLD)
    import ptx

    af, ae = ptx.tri_mesh_adjacency(ptx.icosa_tris())
    tex = O.PtexTexture(np.full((20, 8, 8, 3), 131, np.uint8), 3, af, ae, tri=True)
    rng = np.random.default_rng(5)
    u, v = rng.random(800), rng.random(800)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    out = tex.eval(rng.integers(0, 20, 800), u, v)
    np.testing.assert_allclose(out, 131 / 255, rtol=2e-6)


@pytest.mark.parametrize("name", ["tgrid", "tgrid_rot", "tetra", "icosa", "fan"])
def test_tri_ptex_edge_points_agree_from_both_faces(name):
    """A point on an edge shared by two faces, away from the edge's ends, filters to the same value
    from either face (the kernel pieces are the same, only summed in another order): pins the edge
    convention (edge e runs from corner e to corner e + 1) and every case of the reorientation
    across an edge. (Near a vertex of valence other than 6 the filter's pieces depend on the face
    it starts from, so the ends are left out.)"""
    g, tex = _tri(name)
    corners = np.array([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)], np.float32)
    rng = np.random.default_rng(9)
    fa, ua, va, fb, ub, vb = [], [], [], [], [], []
    cases = set()
    for f in range(g["adjfaces"].shape[0]):
        for e in range(3):
            nf, ne = int(g["adjfaces"][f, e]), int(g["adjedges"][f, e])
            if nf < 0:
                continue
            cases.add((e, ne))
            for t in rng.uniform(0.35, 0.65, 6).astype(np.float32):
                p = corners[e] + t * (corners[(e + 1) % 3] - corners[e])
                q = corners[ne] + (1 - t) * (corners[(ne + 1) % 3] - corners[ne])
                fa.append(f), ua.append(p[0]), va.append(p[1])
                fb.append(nf), ub.append(q[0]), vb.append(q[1])
    a = tex.eval(np.array(fa), np.array(ua, np.float32), np.array(va, np.float32))
    b = tex.eval(np.array(fb), np.array(ub, np.float32), np.array(vb, np.float32))
    np.testing.assert_allclose(a, b, rtol=3e-5, atol=1e-6)
    if name == "tgrid_rot":
        assert len(cases) == 9  # every (edge, neighbour's edge) pair, every case of the reorientation
