"""Selas's binary resource files (selas_amd/resources.py) against the reference's own serializer.

Pinned three ways: (1) the fixture scene's files in tests/golden/resources were written by the
reference's SerializeToBinary (tests/golden/make_golden.py --only resources): they read back to the
values they were written from, this module's writer reproduces them byte for byte, and every array
is found where the reference's AttachToBinary found it; (2) where oracle/_ref is built, random
resources (empty arrays, spare capacity, bytes after a string's terminator) are written by both
writers and compared byte for byte, with the record sizes against the reference's sizeof; (3) the
name hash against MurmurHash3's published test vectors, the material search against the
reference's BinarySearch, the scene camera against SetupSceneCamera."""
import os
import struct

import numpy as np
import pytest

import oracle as O
import resource_fixtures as F
from selas_amd import assets as A
from selas_amd import resources as R

GOLD = np.load(os.path.join(F.FIXTURE_ROOT, "reference.npz"))
LIB = F.ref_lib()
needs_ref = pytest.mark.skipif(LIB is None, reason="oracle/_ref (the reference build) is not built here")
FILES, PTEX_REL = F.fixture_resources()
READ = {"model": R.read_model_resource, "geometry": R.read_model_geometry, "subscene": R.read_subscene_resource,
        "scene": R.read_scene_resource, "ibl": R.read_ibl_resource}


def _fixture_bytes(kind, name):
    with open(R.asset_path(F.FIXTURE_ROOT, kind, name), "rb") as f:
        return f.read()


def _same(got, want, where=""):
    if isinstance(want, dict):
        for k, v in want.items():
            if k == "carrays":
                for an, meta in v.items():
                    for mk, mv in meta.items():
                        assert got["carrays"][an][mk] == mv, (where, an, mk)
                continue
            _same(got[k], v, f"{where}.{k}")
    elif isinstance(want, (np.ndarray, list, tuple)) and not isinstance(want, (str, bytes)):
        g = np.asarray(got)
        w = np.asarray(want)
        if g.dtype.kind == "f" and not w.dtype.names:
            w = w.astype(np.float32)  # values given as Python floats were stored as float32
        if w.dtype.names:
            assert g.dtype == w.dtype and g.tobytes() == w.tobytes(), where
        else:
            np.testing.assert_array_equal(g.reshape(w.shape) if w.size == g.size else g, w, err_msg=where)
    else:
        assert got == want or (isinstance(want, float) and np.float32(got) == np.float32(want)), (where, got, want)


@pytest.mark.parametrize("key", sorted(FILES))
def test_fixture_files_read_back_and_rewrite_byte_identical(key):
    kind, name = key
    blob = _fixture_bytes(kind, name)
    got = READ[kind](blob)
    want = dict(FILES[key])
    if kind == "geometry":
        for n, dt, cols in R.GEOMETRY_BUFFERS:
            w = np.asarray(want.get(n, np.zeros(0, dt)), dt)
            np.testing.assert_array_equal(got[n].reshape(-1), w.reshape(-1), err_msg=n)
    else:
        _same(got, want, f"{kind}:{name}")
    assert F.py_write(kind, got) == blob
    assert F.py_write(kind, FILES[key]) == blob


@pytest.mark.parametrize("key", sorted(FILES))
def test_arrays_are_where_the_reference_attach_finds_them(key):
    kind, name = key
    got = READ[kind](_fixture_bytes(kind, name))
    ref = GOLD[f"attach_{kind}_{name}"]
    if kind in ("geometry", "ibl"):
        assert list(ref) == got["offsets"]
    else:
        arrays = {"model": R.MODEL_ARRAYS, "subscene": R.SUBSCENE_ARRAYS, "scene": R.SCENE_ARRAYS}[kind]
        mine = []
        for an, _ in arrays:
            mine += [got[an].shape[0], got["carrays"][an]["offset"]]
        assert list(ref) == mine


def test_root_records_end_where_the_first_array_starts():
    """The reference's AttachToBinary finds the first array's data right after the root record:
    the records' sizes (ModelResourceData 296, ModelGeometryData 128, SubsceneResourceData 648,
    SceneResourceData 1200, ImageBasedLightResourceData 72 bytes)."""
    assert GOLD["attach_model_crate"][1] == 296
    assert GOLD["attach_geometry_crate"][0] == 128
    assert GOLD["attach_subscene_sub_a"][1] == 648
    assert GOLD["attach_scene_rooms"][1] == 1200
    assert GOLD["attach_ibl_sky"][0] == 72
    empty = R.write_model_resource({"aabox": ((0, 0, 0), (0, 0, 0)), "total_vertex_count": 0,
                                    "total_curve_vertex_count": 0, **{k: 0 for k in R.SIZE_FIELDS},
                                    **{n: [] for n, _ in R.MODEL_ARRAYS}})
    assert len(empty) == 296


@needs_ref
def test_record_sizes_match_reference_sizeof():
    s = np.zeros(14, np.uint64)
    LIB.ref_resource_sizes(s.ctypes.data)
    mine = [296, 128, 648, 1200, 72, R.CAMERA_DT.itemsize, R.PATH_DT.itemsize, R.MATERIAL_DT.itemsize,
            R.MESH_DT.itemsize, R.CURVE_DT.itemsize, R.INSTANCE_DT.itemsize, R.LIGHT_DT.itemsize,
            R.RANGE_DT.itemsize, 32]
    assert list(s) == mine


def _junk_strings(rng, n, size):
    """Fixed strings with text, a terminator and arbitrary bytes after it (all serialized)."""
    out = np.zeros(n, f"S{size}")
    raw = out.view(np.uint8).reshape(n, size)
    for i in range(n):
        k = int(rng.integers(0, size - 1))
        raw[i, :k] = rng.integers(33, 127, k)
        raw[i, k] = 0
        raw[i, k + 1:] = rng.integers(0, 256, size - k - 1) * (rng.random() < 0.5)
    return out


def _random(kind, rng):
    def arr(dt, n):
        a = np.zeros(n, dt)
        b = a.view(np.uint8)
        b[:] = rng.integers(0, 256, b.size)
        return a

    def mats(n):
        m = arr(R.MATERIAL_DT, n)
        m["base_color_texture"] = _junk_strings(rng, n, 512)
        m["scalars"] = rng.normal(size=(n, 14))
        m["base_color"] = rng.random((n, 3))
        return m

    n = lambda: int(rng.integers(0, 5))  # noqa: E731
    if kind == "model":
        d = {"aabox": (tuple(rng.normal(size=3)), tuple(rng.normal(size=3))),
             "total_vertex_count": int(rng.integers(0, 1 << 31)), "total_curve_vertex_count": int(rng.integers(0, 99)),
             "curve_model_name": int(rng.integers(0, 1 << 32)), "pad": int(rng.integers(0, 1 << 32)),
             **{k: int(rng.integers(0, 1 << 40)) for k in R.SIZE_FIELDS}}
        cams = arr(R.CAMERA_DT, n())
        cams["name"] = _junk_strings(rng, cams.shape[0], 128)
        cams["fov_degrees"] = rng.random(cams.shape[0])
        meshes = arr(R.MESH_DT, n())
        meshes["name"] = _junk_strings(rng, meshes.shape[0], 64)
        k = n()
        d.update(cameras=cams, texture_resource_names=_junk_strings(rng, n(), 512), materials=mats(k),
                 material_hashes=np.sort(rng.integers(0, 1 << 32, k).astype(np.uint32)), meshes=meshes,
                 curves=arr(R.CURVE_DT, n()))
        d["carrays"] = {a: {"capacity": np.asarray(d[a]).shape[0] + int(rng.integers(0, 3))} for a, _ in R.MODEL_ARRAYS}
        return d
    if kind == "geometry":
        return {name: rng.random(int(rng.integers(0, 7)) * cols).astype(dt).reshape(-1, cols) if cols > 1 else
                rng.integers(0, 99, int(rng.integers(0, 7))).astype(dt) for name, dt, cols in R.GEOMETRY_BUFFERS}
    if kind == "subscene":
        inst = arr(R.INSTANCE_DT, n())
        inst["local_to_world"] = rng.normal(size=(inst.shape[0], 4, 4))
        k = n()
        return {"name": "sub_" + str(int(rng.integers(0, 999))), "light_set_index": int(rng.integers(0, 1 << 40)),
                "model_names": _junk_strings(rng, n(), 512), "model_instances": inst,
                "scene_material_names": rng.integers(0, 1 << 32, k).astype(np.uint32), "scene_materials": mats(k)}
    if kind == "scene":
        inst = arr(R.INSTANCE_DT, n())
        inst["local_to_world"] = rng.normal(size=(inst.shape[0], 4, 4))
        lights = arr(R.LIGHT_DT, n())
        for f in ("position", "direction", "x", "z", "radiance"):
            lights[f] = rng.normal(size=(lights.shape[0], 3))
        cams = arr(R.CAMERA_DT, n())
        cams["name"] = _junk_strings(rng, cams.shape[0], 128)
        return {"name": "scene", "ibl_name": "ibl_" + "x" * int(rng.integers(0, 30)),
                "background_intensity": tuple(rng.random(4)), "subscene_names": _junk_strings(rng, n(), 512),
                "subscene_instances": inst, "lights": lights, "light_set_ranges": arr(R.RANGE_DT, n()), "cameras": cams,
                "carrays": {"lights": {"capacity": lights.shape[0] + 5, "flags": 0}}}
    w, h = int(rng.integers(1, 6)), int(rng.integers(1, 5))
    mw, mh = (int(rng.integers(1, 4)), int(rng.integers(1, 4))) if rng.random() < 0.5 else (0, 0)
    return {"width": w, "height": h, "marginal": rng.random(h).astype(np.float32),
            "conditional": rng.random((h, w)).astype(np.float32), "miss_width": mw, "miss_height": mh,
            "rotation": float(np.float32(rng.normal())), "exposure_scale": float(np.float32(rng.random())),
            "light": rng.random((h, w, 3)).astype(np.float32), "miss": rng.random((mh, mw, 3)).astype(np.float32)}


@needs_ref
@pytest.mark.parametrize("kind", ["model", "geometry", "subscene", "scene", "ibl"])
def test_random_resources_byte_identical_to_reference_writer(kind):
    rng = np.random.default_rng(["model", "geometry", "subscene", "scene", "ibl"].index(kind) + 100)
    for _ in range(25):
        d = _random(kind, rng)
        ref = F.ref_write(LIB, kind, d)
        assert F.py_write(kind, d) == ref
        got = READ[kind](ref)
        assert F.py_write(kind, got) == ref
        offs = F.ref_attach(LIB, kind, ref)
        if kind in ("geometry", "ibl"):
            assert list(offs) == got["offsets"]


def test_murmur3_published_vectors():
    """MurmurHash3_x86_32's published test vectors."""
    cases = [(b"", 0, 0), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39), (b"\0\0\0\0", 0, 0x2362F9DE),
             (b"aaaa", 0x9747B28C, 0x5A97808A), (b"aaa", 0x9747B28C, 0x283E0130), (b"aa", 0x9747B28C, 0x5D211726),
             (b"a", 0x9747B28C, 0x7FA09EA6), (b"abcd", 0x9747B28C, 0xF0478627), (b"abc", 0x9747B28C, 0xC84A62DD),
             (b"ab", 0x9747B28C, 0x74875592), (b"Hello, world!", 0x9747B28C, 0x24884CBA),
             (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD), (b"hello", 0, 0x248BFA47)]
    for data, seed, want in cases:
        assert R.murmur3_32(data, seed) == want, data


@needs_ref
def test_murmur3_matches_reference():
    rng = np.random.default_rng(9)
    for n in list(range(0, 20)) + [63, 64, 65, 200]:
        for _ in range(5):
            data = bytes(rng.integers(1, 256, n).astype(np.uint8))
            seed = int(rng.integers(0, 1 << 32))
            assert R.murmur3_32(data, seed) == LIB.ref_murmur3_32(data, n, seed)


@needs_ref
def test_binary_search_matches_reference():
    """Sorted keys with duplicates and unsorted keys (the bisection's answer either way), keys at
    most the largest value (beyond it the reference reads before its array)."""
    rng = np.random.default_rng(11)
    for trial in range(300):
        n = int(rng.integers(1, 40))
        data = rng.integers(0, 50, n).astype(np.uint32)
        if trial % 3:
            data.sort()
        keys = rng.integers(0, int(data.max()) + 1, 10).astype(np.uint32)
        for key in list(keys) + list(data[:3]):
            r = LIB.ref_binary_search(data.ctypes.data, n, int(key))
            want = -1 if r == (1 << 64) - 1 else int(r)
            assert R.binary_search(data, int(key)) == want


def test_binary_search_is_a_lower_bound_on_sorted_keys():
    rng = np.random.default_rng(12)
    for _ in range(200):
        data = np.sort(rng.integers(0, 30, int(rng.integers(0, 25))).astype(np.uint32))
        for key in range(0, 32):
            i = int(np.searchsorted(data, key, "left"))
            want = i if i < data.size and data[i] == key else -1
            assert R.binary_search(data, key) == want


def test_find_mesh_material_follows_the_reference_order():
    """Model material by hash first, then the first subscene material of that name, else the
    default (base colour 0.6, DisneySolid, ior 1.5)."""
    crate = FILES[("model", "crate")]
    sub_a, sub_b = FILES[("subscene", "sub_a")], FILES[("subscene", "sub_b")]
    body, lid = crate["meshes"]
    m, where = R.find_mesh_material(crate, sub_a["scene_material_names"], sub_a["scene_materials"],
                                    int(body["material_hash"]))
    assert where[0] == "model" and crate["material_hashes"][where[1]] == body["material_hash"]
    _, where = R.find_mesh_material(crate, sub_a["scene_material_names"], sub_a["scene_materials"],
                                    int(lid["material_hash"]))
    assert where == ("subscene", 0)
    _, where = R.find_mesh_material(crate, sub_b["scene_material_names"], sub_b["scene_materials"],
                                    int(lid["material_hash"]))
    assert where == ("subscene", 1)
    panel = FILES[("model", "panel")]
    m, where = R.find_mesh_material(panel, [], [], int(panel["meshes"][1]["material_hash"]))
    assert where == ("default", -1)
    rec = R.material_record(m)
    assert np.all(rec[0:3] == np.float32(0.6)) and rec[19] == np.float32(1.5) and rec[20] == 1 and rec[21] == 0
    assert np.count_nonzero(rec) == 5
    # a subscene that names the hash twice: the first entry
    names = np.array([5, int(lid["material_hash"]), int(lid["material_hash"])], np.uint32)
    mats = np.array([F.material((0, 0, 0)), F.material((1, 0, 0)), F.material((0, 1, 0))], R.MATERIAL_DT)
    _, where = R.find_mesh_material(crate, names, mats, int(lid["material_hash"]))
    assert where == ("subscene", 1)


def test_material_record_slots():
    """The 22-float record holds MaterialResourceData's values in the slots build_material uses."""
    rng = np.random.default_rng(3)
    m = F.material((0.1, 0.2, 0.3), shader=R.SHADER_THIN, flags=A.USES_PTEX | A.TRANSPARENT, rng=rng)
    rec = R.material_record(m)
    np.testing.assert_array_equal(rec[0:3], m["base_color"])
    np.testing.assert_array_equal(rec[3:6], m["transmittance_color"])
    for attr, slot in A.SLOTS.items():
        assert rec[slot] == m["scalars"][A.ATTRIBUTES.index(attr)], attr
    assert rec[15] == 0 and rec[20] == 0 and rec[21] == A.USES_PTEX | A.TRANSPARENT
    imp = A.ImportedMaterial(shader_name="DisneyThin", base_color=tuple(m["base_color"]),
                             transmittance_color=tuple(m["transmittance_color"]),
                             scalars={a: float(m["scalars"][i]) for i, a in enumerate(A.ATTRIBUTES)}, uses_ptex=True)
    built = A.build_material(imp)["record"]
    np.testing.assert_array_equal(rec[:21], built[:21])
    m["shader"] = R.SHADER_DIRAC
    with pytest.raises(ValueError, match="DiracTransparent"):
        R.material_record(m)


@pytest.mark.parametrize("index", [0, 1])
def test_scene_camera_matches_setup_scene_camera(index):
    """Camera 0 of the fixture is invalid (fov 0): the default camera; camera 1 is used as it is.
    The basis the oracle's InitializeRayCastCamera makes of the chosen settings equals the
    reference's SetupSceneCamera output recorded in the fixture."""
    scene = R.read_scene_resource(_fixture_bytes("scene", "rooms"))
    cam = R.scene_camera(scene, index)
    cam10 = np.array(list(cam["position"]) + list(cam["look_at"]) + list(cam["up"]) + [cam["fov"]], np.float32)
    _, basis = O.camera_rays(cam10, 64, 48, np.zeros((0, 6), np.int32))
    ref = GOLD[f"camera_{index}"]
    np.testing.assert_array_equal(basis, ref[:10])
    np.testing.assert_array_equal(np.float32(cam["position"]), ref[10:13])
    if LIB is not None:
        np.testing.assert_array_equal(F.ref_scene_camera(LIB, _fixture_bytes("scene", "rooms"), index, 64, 48), ref)
    assert R.scene_camera(scene, 7)["fov"] == 35.0


def _fixture_ptex(path):
    assert path.endswith(os.path.join("Content", PTEX_REL)), path
    return F.floor_ptex()


def test_load_scene_assembles_the_fixture():
    ls = R.load_scene(F.FIXTURE_ROOT, "rooms", camera_index=1, ptex_loader=_fixture_ptex)
    d = ls.description
    assert ls.dropped == []
    # crate (sub_a's paint), panel, floor, crate (sub_b's paint)
    assert len(d["models"]) == 4
    crate_a, panel, floor, crate_b = d["models"]
    assert [len(s) for s in d["subscenes"]] == [3, 2]
    assert [i[0] for i in d["subscenes"][0]] == [0, 1, 0] and [i[0] for i in d["subscenes"][1]] == [2, 3]
    src = FILES[("subscene", "sub_a")]["model_instances"]
    for (mi, m), inst in zip(d["subscenes"][0], src):
        m4 = np.asarray(inst["local_to_world"], np.float64).T
        np.testing.assert_array_equal(m, m4)
    assert [(k, i) for k, i, _ in d["instances"]] == [("subscene", 1), ("subscene", 0), ("subscene", 0)]
    # a point through an instance matrix goes where Selas's MatrixMultiplyPoint puts it
    inst = FILES[("scene", "rooms")]["subscene_instances"][2]
    p = np.array([0.3, -1.2, 2.0])
    rows = np.asarray(inst["local_to_world"], np.float64)
    np.testing.assert_allclose((d["instances"][2][2] @ np.r_[p, 1.0])[:3], p @ rows[:3, :3] + rows[3, :3])
    # geometry: the crate's meshes are quads over the vertices they use, with their attributes
    cgeo = FILES[("geometry", "crate")]
    body, lid = crate_a["geoms"]
    assert body["quads"].shape == (5, 4) and lid["quads"].shape == (1, 4)
    np.testing.assert_array_equal(body["positions"][body["quads"]], cgeo["positions"][cgeo["indices"][:20].reshape(5, 4)])
    np.testing.assert_array_equal(lid["tangents"][lid["quads"]], cgeo["tangents"][cgeo["indices"][20:].reshape(1, 4)])
    assert "triangles" in panel["geoms"][0] and panel["geoms"][0]["triangles"].shape == (2, 3)
    assert "normals" not in panel["geoms"][0]
    # materials: wood (model), paint blue in sub_a and red in sub_b (subscene), default on the panel top
    recs = lambda mdl: mdl["materials"][mdl["geom_material"]]  # noqa: E731
    np.testing.assert_array_equal(recs(crate_a)[1][:3], np.float32([0.1, 0.4, 0.8]))
    np.testing.assert_array_equal(recs(crate_b)[1][:3], np.float32([0.8, 0.1, 0.1]))
    np.testing.assert_array_equal(recs(crate_a)[0], recs(crate_b)[0])
    assert recs(panel)[1][0] == np.float32(0.6) and recs(panel)[0][10] == 1.0
    assert recs(floor)[0][21] == A.USES_PTEX and floor["geoms"][0]["ptex"]["res_log2"] == F.FLOOR_PTEX_RES
    # light set 0 (the first two lights), background, image-based light, camera 1
    np.testing.assert_array_equal(d["lights"], FILES[("scene", "rooms")]["lights"][:2])
    np.testing.assert_array_equal(d["background"], np.float32([0.1, 0.2, 0.3]))
    np.testing.assert_array_equal(d["ibl"]["rgb"], FILES[("ibl", "sky")]["light"])
    assert d["ibl"]["rotation"] == np.float32(0.3) and d["ibl"]["exposure"] == np.float32(0.8)
    assert ls.camera["fov"] == np.float32(55.0)


def _variant(tmp_path, edit):
    files = {k: dict(v) for k, v in FILES.items()}
    edit(files)
    F.write_tree(str(tmp_path), files, F.py_write)
    return str(tmp_path)


def test_load_scene_refuses_what_the_device_does_not_render(tmp_path):
    def curves(files):
        d, g = F.model_dict([("m", 1, [[0, 1, 2]])], [(0, 0, 0), (1, 0, 0), (0, 1, 0)], curves=[(0, 1)],
                            curve_indices=[0], curve_vertices=np.ones((4, 4)))
        files[("model", "panel")], files[("geometry", "panel")] = d, g

    # curves load (InitializeCurves: after the meshes, the material of curveModelName, no texture)
    root = _variant(tmp_path / "curves", curves)
    ls = R.load_scene(root, "rooms", ptex_loader=_fixture_ptex)
    assert not ls.dropped
    panel = [m for m in ls.description["models"] if any("curves" in g for g in m["geoms"])]
    assert len(panel) == 1 and "triangles" in panel[0]["geoms"][0]
    cg = panel[0]["geoms"][1]
    np.testing.assert_array_equal(cg["curves"], [0])
    assert cg["curve_vertices"].shape == (4, 4)

    def bad_curve(files):
        # two segments over four control points: the second (from index 1) needs a fifth
        d, g = F.model_dict([("m", 1, [[0, 1, 2]])], [(0, 0, 0), (1, 0, 0), (0, 1, 0)], curves=[(0, 2)],
                            curve_indices=[0, 1], curve_vertices=np.ones((4, 4)))
        files[("model", "panel")], files[("geometry", "panel")] = d, g

    root = _variant(tmp_path / "badcurve", bad_curve)
    with pytest.raises(R.ResourceError, match="curve vertices"):
        R.load_scene(root, "rooms", ptex_loader=_fixture_ptex)
    # no texture loader: the floor's texture is refused, or dropped with its flag cleared
    with pytest.raises(ValueError, match="per-face texture"):
        R.load_scene(F.FIXTURE_ROOT, "rooms")
    ls = R.load_scene(F.FIXTURE_ROOT, "rooms", strict=False)
    assert ls.description["models"][2]["materials"][0][21] == 0 and len(ls.dropped) == 1

    def two_sets(files):
        s = dict(files[("subscene", "sub_b")])
        s["light_set_index"] = 1
        files[("subscene", "sub_b")] = s

    root = _variant(tmp_path / "sets", two_sets)
    with pytest.raises(ValueError, match="light sets"):
        R.load_scene(root, "rooms", ptex_loader=_fixture_ptex)

    def no_set(files):
        for k in ("sub_a", "sub_b"):
            s = dict(files[("subscene", k)])
            s["light_set_index"] = 9
            files[("subscene", k)] = s

    root = _variant(tmp_path / "noset", no_set)
    assert R.load_scene(root, "rooms", ptex_loader=_fixture_ptex).description["lights"].shape[0] == 0

    def uv_texture(files):
        m = dict(files[("model", "crate")])
        mats = m["materials"].copy()
        used = list(m["material_hashes"]).index(m["meshes"][0]["material_hash"])
        mats["base_color_texture"][used] = b"wood_albedo"
        m["materials"] = mats
        files[("model", "crate")] = m

    root = _variant(tmp_path / "uv", uv_texture)
    with pytest.raises(ValueError, match="UV base-colour texture"):
        R.load_scene(root, "rooms", ptex_loader=_fixture_ptex)

    def dual_ibl(files):
        i = dict(files[("ibl", "sky")])
        i["miss_width"], i["miss_height"], i["miss"] = 2, 1, np.ones((1, 2, 3), np.float32)
        files[("ibl", "sky")] = i

    root = _variant(tmp_path / "ibl", dual_ibl)
    with pytest.raises(ValueError, match="miss map"):
        R.load_scene(root, "rooms", ptex_loader=_fixture_ptex)


def test_malformed_files_are_refused():
    blob = _fixture_bytes("model", "crate")
    with pytest.raises(R.ResourceError, match="ends inside"):
        R.read_model_resource(blob[:100])
    with pytest.raises(R.ResourceError, match="past the file"):
        R.read_model_resource(blob[:-1])
    bad = bytearray(blob)
    bad[296 - 8:296] = struct.pack("<Q", R.POINTER_KEY)  # the curves' pointer never written
    with pytest.raises(R.ResourceError, match="never written"):
        R.read_model_resource(bytes(bad))
    geo = bytearray(_fixture_bytes("geometry", "crate"))
    geo[64:72] = struct.pack("<Q", 136)  # indices at an offset that is not 16-byte aligned
    with pytest.raises(R.ResourceError, match="aligned"):
        R.read_model_geometry(bytes(geo))


def test_mesh_faces_follow_indices_per_face():
    """indicesPerFace 3: triangles; otherwise quads from the first 4 indices of each face at a
    stride of indicesPerFace (UINT4 at that stride, ModelResource.cpp:273-277); fewer than 3, or an
    index past the vertices, is refused."""
    idx = np.arange(20, dtype=np.uint32)
    mesh = np.zeros((), R.MESH_DT)
    mesh["indices_per_face"], mesh["index_count"], mesh["index_offset"] = 5, 12, 3
    f = R._mesh_faces(idx, mesh, 40, "m")
    np.testing.assert_array_equal(f, [[3, 4, 5, 6], [8, 9, 10, 11]])
    mesh["indices_per_face"] = 2
    with pytest.raises(R.ResourceError):
        R._mesh_faces(idx, mesh, 40, "m")
    mesh["indices_per_face"] = 3
    with pytest.raises(R.ResourceError, match="vertex index"):
        R._mesh_faces(idx, mesh, 10, "m")
