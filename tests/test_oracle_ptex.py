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
