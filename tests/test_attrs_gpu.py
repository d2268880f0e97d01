"""GPU parity of per-vertex shading attributes (sgpu_scene_set_vertex_attributes: Selas's vertex
attribute slots, interpolated by CalculateSurfaceParams) against the oracle, through the C ABI.

The attributed room: the icosphere as a smooth triangle mesh (vertex normals, tangents with a w sign
that flips across x = 0, UVs; anisotropic Disney) and a curved quad panel with tilted vertex normals
and tangents. Rendered by the path tracer, the deferred path tracer and VCM, flat and through two
instance levels. Bar, as for the other renderers: per pixel |gpu - cpu| <= 1e-3 |cpu| + 1e-5 for at
least 99 % of the values, relative RMSE <= 2e-3 of the mean.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from selas_amd import native as N
from selas_amd import workloads as W

pytestmark = pytest.mark.gpu


def _close(gpu, cpu):
    frac = np.mean(np.abs(gpu - cpu) <= 1e-3 * np.abs(cpu) + 1e-5)
    rel = np.sqrt(np.mean((gpu - cpu) ** 2)) / max(float(cpu.mean()), 1e-12)
    return frac, rel


@pytest.fixture(scope="module")
def room(gpu_ctx):
    sc = W.attributed_room_scene(1)
    scene = N.build_render_scene(gpu_ctx, sc)
    yield sc, scene, O.RenderScene(sc)
    scene.release()


def test_pt_with_attributes_matches_oracle(room):
    sc, scene, rs = room
    print("pt attributed room 80x64x8", flush=True)
    gpu = N.render_pt(scene, sc["camera"], 80, 64, 8, 8, 21)
    cpu = O.render_pt(rs, sc["camera"], 80, 64, 8, 8, 21)
    assert np.isfinite(gpu).all() and cpu[1].mean() > 0.01
    frac, rel = _close(gpu, cpu)
    assert frac >= 0.99 and rel <= 2e-3, (frac, rel)


def test_dpt_with_attributes_matches_oracle(room):
    sc, scene, rs = room
    print("dpt attributed room 64x48x(2x2)", flush=True)
    gpu = N.render_deferred_pt(scene, sc["camera"], 64, 48, 2, 2, 10, 5)
    cpu = O.render_dpt(rs, sc["camera"], 64, 48, 2, 2, 10, 5)
    assert np.isfinite(gpu).all() and cpu.mean() > 0.01
    frac, rel = _close(gpu, cpu)
    assert frac >= 0.99 and rel <= 2e-3, (frac, rel)


def test_vcm_with_attributes_matches_oracle(room):
    sc, scene, rs = room
    print("vcm attributed room 64x48x3", flush=True)
    gpu = N.render_vcm(scene, sc["camera"], 64, 48, 3, 10, 0.005, 0.75, 17)
    cpu, _ = O.render_vcm(rs, sc["camera"], 64, 48, 3, 10, 0.005, 0.75, 17)
    assert np.isfinite(gpu).all() and cpu.mean() > 0.01
    frac, rel = _close(gpu, cpu)
    assert frac >= 0.99 and rel <= 2e-3, (frac, rel)


def _instanced_attributed():
    """The attributed sphere and panel as placed scenes: the sphere twice through a subscene and
    once directly, the panel once; room walls as a third model."""
    base = W.attributed_room_scene(1)
    walls = {"geoms": base["geoms"][:12], "materials": base["materials"][:12],
             "geom_material": np.arange(12, dtype=np.uint32)}
    sphere = {"geoms": [base["geoms"][12]], "materials": base["materials"][12:13],
              "geom_material": np.zeros(1, np.uint32)}
    panel = {"geoms": [base["geoms"][13]], "materials": base["materials"][13:14],
             "geom_material": np.zeros(1, np.uint32)}

    def m4(s=1.0, t=(0, 0, 0), ry=0.0):
        c, si = np.cos(ry), np.sin(ry)
        m = np.array([[c * s, 0, si * s, t[0]], [0, s, 0, t[1]], [-si * s, 0, c * s, t[2]], [0, 0, 0, 1]])
        return np.ascontiguousarray(m.T.reshape(-1), np.float32)

    return {
        "models": [walls, sphere, panel],
        "subscenes": [[(1, m4(0.5, (-0.22, -0.1, 0.0), 0.4)), (1, m4(0.4, (0.25, -0.12, 0.1), -0.7))]],
        "instances": [("model", 0, m4()), ("subscene", 0, m4(1.0, (0.0, 0.0, 0.1), 0.2)),
                      ("model", 1, m4(0.6, (0.1, 0.25, 0.3), 1.1)), ("model", 2, m4(1.0, (0.0, 0.05, 0.0), 0.05))],
        "lights": base["lights"],
        "background": base["background"],
        "camera": base["camera"],
    }


def test_instanced_attributes_match_oracle(gpu_ctx):
    sc = _instanced_attributed()
    scene = N.build_render_scene(gpu_ctx, sc)
    try:
        rs = O.InstancedRenderScene(sc)
        print("instanced attributed pt 64x48x8", flush=True)
        gpu = N.render_pt(scene, sc["camera"], 64, 48, 8, 8, 3)
        cpu = O.render_pt(rs, sc["camera"], 64, 48, 8, 8, 3)
        assert cpu[1].mean() > 0.01
        frac, rel = _close(gpu, cpu)
        assert frac >= 0.99 and rel <= 2e-3, (frac, rel)
        print("instanced attributed vcm 48x40x2", flush=True)
        gv = N.render_vcm(scene, sc["camera"], 48, 40, 2, 8, 0.005, 0.75, 9)
        cv, _ = O.render_vcm(rs, sc["camera"], 48, 40, 2, 8, 0.005, 0.75, 9)
        frac, rel = _close(gv, cv)
        assert frac >= 0.99 and rel <= 2e-3, (frac, rel)
    finally:
        scene.release()


def test_attribute_errors(gpu_ctx):
    mesh = W.icosphere_mesh(1, (0, 0, 0), 1.0)
    nv = mesh["positions"].shape[0]
    scene = N.Scene(gpu_ctx)
    try:
        g = scene.add_triangles(mesh["positions"], mesh["triangles"])
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32).ctypes.data
        call = lambda n, t, uv, k=nv, geom=g: N.lib.sgpu_scene_set_vertex_attributes(scene.handle, geom, f(n), f(t),
                                                                                     f(uv), k)
        assert call(None, mesh["tangents"], None) == N.SGPU_E_INVALID  # tangents need normals
        assert call(mesh["normals"], None, mesh["uvs"]) == N.SGPU_E_INVALID  # UVs need tangents
        assert call(mesh["normals"], None, None, k=nv - 1) == N.SGPU_E_INVALID  # wrong vertex count
        assert call(mesh["normals"], None, None, geom=5) == N.SGPU_E_INVALID  # no such geometry
        bad = mesh["normals"].copy()
        bad[3, 1] = np.nan
        assert call(bad, None, None) == N.SGPU_E_INVALID
        assert call(mesh["normals"], mesh["tangents"], mesh["uvs"]) == N.SGPU_OK
        assert call(mesh["normals"], None, None) == N.SGPU_OK  # replaced
        scene.commit()
        assert call(mesh["normals"], None, None) == N.SGPU_E_STATE
        assert scene.device_bytes() > 0
    finally:
        scene.release()
