"""Selas resource files for the tests: the reference's own writer (oracle/_ref, when built) driven
from the dicts selas_amd.resources reads and writes, random resources, and the fixture scene whose
files tests/golden/make_golden.py writes with the reference writer into tests/golden/resources/.

The fixture scene ("rooms"): three models -- a quad crate with normals, tangents and UVs whose two
meshes take their materials from the model (binary search) and from the subscene, a triangle panel
whose meshes take a model material and the default, a quad floor with a per-face texture -- in two
subscenes with different subscene materials (so the crate resolves to two device models), placed
three times at two levels, three quad lights in two light sets, a background, an image-based light
and two cameras (the first invalid)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from selas_amd import resources as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libselas_ref.so")
FIXTURE_ROOT = os.path.join(ROOT, "tests", "golden", "resources")
U64 = C.c_uint64
P = C.c_void_p


def ref_lib():
    if not os.path.exists(REF):
        return None
    lib = C.CDLL(REF)
    if not hasattr(lib, "ref_model_resource_write"):
        return None
    lib.ref_model_resource_write.restype = U64
    lib.ref_model_resource_write.argtypes = [P] * 15 + [P, U64]
    lib.ref_model_geometry_write.restype = U64
    lib.ref_model_geometry_write.argtypes = [P, P, P, U64]
    lib.ref_subscene_write.restype = U64
    lib.ref_subscene_write.argtypes = [P, U64] + [P] * 9 + [P, U64]
    lib.ref_scene_write.restype = U64
    lib.ref_scene_write.argtypes = [P] * 13 + [P, U64]
    lib.ref_ibl_write.restype = U64
    lib.ref_ibl_write.argtypes = [U64, U64, P, P, U64, U64, C.c_float, C.c_float, P, P, P, U64]
    lib.ref_attach_offsets.restype = C.c_int
    lib.ref_attach_offsets.argtypes = [C.c_int, P, U64, P]
    lib.ref_scene_camera.argtypes = [P, U64, U64, C.c_uint32, C.c_uint32, P]
    lib.ref_murmur3_32.restype = C.c_uint32
    lib.ref_murmur3_32.argtypes = [C.c_char_p, C.c_int32, C.c_uint32]
    lib.ref_binary_search.restype = U64
    lib.ref_binary_search.argtypes = [P, U64, C.c_uint32]
    lib.ref_resource_sizes.argtypes = [P]
    return lib


class _Args:
    """Keeps the numpy buffers handed to a reference call alive."""

    def __init__(self):
        self.keep = []

    def __call__(self, a, dtype=None):
        a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
        self.keep.append(a)
        return a.ctypes.data if a.size else None


def _counts(d, arrays):
    counts = np.array([np.asarray(d[n]).shape[0] for n, _ in arrays], np.uint64)
    meta = d.get("carrays", {})
    caps = np.array([meta.get(n, {}).get("capacity", c) for (n, _), c in zip(arrays, counts)], np.uint64)
    return counts, caps


def _strs(a, size):
    a = np.asarray(a, f"S{size}")
    return np.frombuffer(a.tobytes(), np.uint8) if a.size else np.zeros(0, np.uint8)


def _cams(cams):
    cams = np.asarray(cams, R.CAMERA_DT)
    f = np.zeros((cams.shape[0], 12), np.float32)
    if cams.shape[0]:
        f[:, 0:3] = cams["position"]
        f[:, 3] = cams["fov_degrees"]
        f[:, 4:7] = cams["look_at"]
        f[:, 7] = cams["znear"]
        f[:, 8:11] = cams["up"]
        f[:, 11] = cams["zfar"]
    return _strs(cams["name"], 128), f


def _mats(mats):
    mats = np.asarray(mats, R.MATERIAL_DT)
    u = np.zeros((mats.shape[0], 2), np.uint32)
    f = np.zeros((mats.shape[0], 20), np.float32)
    if mats.shape[0]:
        u[:, 0] = mats["shader"]
        u[:, 1] = mats["flags"]
        f[:, 0:3] = mats["base_color"]
        f[:, 3:6] = mats["transmittance_color"]
        f[:, 6:20] = mats["scalars"]
    return _strs(mats["base_color_texture"], 512), u, f


def _call(lib, fn, args, a):
    size = fn(*args, None, 0)
    out = np.zeros(size, np.uint8)
    size2 = fn(*args, out.ctypes.data, size)
    assert size2 == size
    return out.tobytes()


def ref_write(lib, kind: str, d: dict) -> bytes:
    """The reference writer's file for a resource dict."""
    a = _Args()
    if kind == "model":
        lo, hi = d["aabox"]
        counts, caps = _counts(d, R.MODEL_ARRAYS)
        cn, cf = _cams(d["cameras"])
        mt, mu, mf = _mats(d["materials"])
        meshes = np.asarray(d["meshes"], R.MESH_DT)
        mesh_u = np.zeros((meshes.shape[0], 7), np.uint32)
        for k, name in enumerate(("index_count", "index_offset", "vertex_count", "vertex_offset", "material_hash",
                                  "indices_per_face", "name_hash")):
            mesh_u[:, k] = meshes[name]
        curves = np.asarray(d["curves"], R.CURVE_DT)
        curve_u = np.stack([curves["index_offset"], curves["index_count"]], 1).astype(np.uint32)
        args = [a(np.r_[lo, hi], np.float32),
                a([d["total_vertex_count"], d["total_curve_vertex_count"], d.get("curve_model_name", 0),
                   d.get("pad", 0)], np.uint32),
                a([d[k] for k in R.SIZE_FIELDS], np.uint64), a(counts), a(caps), a(cn), a(cf),
                a(_strs(d["texture_resource_names"], 512)), a(mt), a(mu), a(mf),
                a(d["material_hashes"], np.uint32), a(mesh_u), a(_strs(meshes["name"], 64)), a(curve_u)]
        return _call(lib, lib.ref_model_resource_write, args, a)
    if kind == "geometry":
        bufs = [np.ascontiguousarray(np.asarray(d.get(n, np.zeros(0, dt)), dt)) for n, dt, _ in R.GEOMETRY_BUFFERS]
        sizes = a([b.nbytes for b in bufs], np.uint64)
        ptrs = a([b.ctypes.data if b.size else 0 for b in bufs], np.uint64)
        a.keep += bufs
        return _call(lib, lib.ref_model_geometry_write, [sizes, ptrs], a)
    if kind == "subscene":
        counts, caps = _counts(d, R.SUBSCENE_ARRAYS)
        inst = np.asarray(d["model_instances"], R.INSTANCE_DT)
        mt, mu, mf = _mats(d["scene_materials"])
        args = [a(_strs([R.to_fixed(d["name"], 512)], 512)), d["light_set_index"], a(counts), a(caps),
                a(_strs(d["model_names"], 512)), a(inst["index"], np.uint64),
                a(inst["local_to_world"].reshape(-1, 16), np.float32), a(d["scene_material_names"], np.uint32),
                a(mt), a(mu), a(mf)]
        return _call(lib, lib.ref_subscene_write, args, a)
    if kind == "scene":
        counts, caps = _counts(d, R.SCENE_ARRAYS)
        inst = np.asarray(d["subscene_instances"], R.INSTANCE_DT)
        lights = np.asarray(d["lights"], R.LIGHT_DT)
        lf = np.concatenate([lights[k].reshape(-1, 3) for k in ("position", "direction", "x", "z", "radiance")], 1) \
            if lights.shape[0] else np.zeros((0, 15), np.float32)
        ranges = np.asarray(d["light_set_ranges"], R.RANGE_DT)
        cn, cf = _cams(d["cameras"])
        args = [a(_strs([R.to_fixed(d["name"], 512)], 512)), a(_strs([R.to_fixed(d["ibl_name"], 512)], 512)),
                a(d["background_intensity"], np.float32), a(counts), a(caps), a(_strs(d["subscene_names"], 512)),
                a(inst["index"], np.uint64), a(inst["local_to_world"].reshape(-1, 16), np.float32),
                a(lights["type"], np.uint32), a(lf, np.float32),
                a(np.stack([ranges["start"], ranges["count"]], 1) if ranges.shape[0] else np.zeros(0), np.uint32),
                a(cn), a(cf)]
        return _call(lib, lib.ref_scene_write, args, a)
    if kind == "ibl":
        args = [d["width"], d["height"], a(d["marginal"], np.float32), a(d["conditional"], np.float32),
                d["miss_width"], d["miss_height"], d["rotation"], d["exposure_scale"], a(d["light"], np.float32),
                a(d["miss"], np.float32)]
        return _call(lib, lib.ref_ibl_write, args, a)
    raise KeyError(kind)


def ref_attach(lib, kind: str, blob: bytes) -> np.ndarray:
    buf = np.frombuffer(blob, np.uint8).copy()
    out = np.zeros(16, np.uint64)
    n = lib.ref_attach_offsets(("model", "geometry", "subscene", "scene", "ibl").index(kind), buf.ctypes.data,
                               buf.size, out.ctypes.data)
    return out[:n]


def ref_scene_camera(lib, blob: bytes, index: int, width: int, height: int) -> np.ndarray:
    buf = np.frombuffer(blob, np.uint8).copy()
    out = np.zeros(13, np.float32)
    lib.ref_scene_camera(buf.ctypes.data, buf.size, index, width, height, out.ctypes.data)
    return out


# ---- resource dicts ---------------------------------------------------------------------------------

def material(base, shader=R.SHADER_SOLID, flags=0, texture="", rng=None, **scalars):
    m = R.default_material()
    m["shader"] = shader
    m["flags"] = flags
    m["base_color"] = base
    m["base_color_texture"] = texture.encode()
    if rng is not None:
        m["transmittance_color"] = rng.uniform(0, 1, 3)
        m["scalars"] = rng.uniform(0, 1, 14)
        m["scalars"][R.assets.ATTRIBUTES.index("ior")] = rng.uniform(1.1, 2.0)
    for k, v in scalars.items():
        m["scalars"][R.assets.ATTRIBUTES.index(k)] = v
    return m


def camera(name, position, look_at, up=(0, 1, 0), fov=45.0, znear=0.01, zfar=1000.0):
    c = np.zeros((), R.CAMERA_DT)
    c["name"] = name.encode()
    c["position"], c["look_at"], c["up"] = position, look_at, up
    c["fov_degrees"], c["znear"], c["zfar"] = fov, znear, zfar
    return c


def instance(index, m4):
    """An Instance record from a 4 x 4 matrix acting on column vectors (stored transposed)."""
    r = np.zeros((), R.INSTANCE_DT)
    r["index"] = index
    r["local_to_world"] = np.asarray(m4, np.float32).T
    return r


def affine(scale=1.0, ry=0.0, t=(0, 0, 0)):
    c, s = np.cos(ry), np.sin(ry)
    m = np.eye(4)
    m[:3, :3] = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]) * scale
    m[:3, 3] = t
    return m


def model_dict(meshes, positions, materials=(), hashes=(), normals=None, tangents=None, uvs=None, curves=(),
               curve_indices=None, curve_vertices=None, cameras=(), textures=(), capacity_extra=0):
    """A ModelResource dict and its ModelGeometryResource dict. meshes: (name, material hash, faces
    (k, 3 or 4)) in order; indices are laid out mesh after mesh into one buffer."""
    idx, mrec = [], np.zeros(len(meshes), R.MESH_DT)
    off = 0
    for i, (name, mh, faces) in enumerate(meshes):
        faces = np.asarray(faces, np.uint32)
        idx.append(faces.reshape(-1))
        mrec[i] = (faces.size, off, int(faces.max()) - int(faces.min()) + 1, int(faces.min()), mh, faces.shape[1],
                   R.murmur3_32(name), name.encode())
        off += faces.size
    positions = np.asarray(positions, np.float32)
    geo = {"indices": np.concatenate(idx) if idx else np.zeros(0, np.uint32), "positions": positions}
    if normals is not None:
        geo["normals"] = np.asarray(normals, np.float32)
    if tangents is not None:
        geo["tangents"] = np.asarray(tangents, np.float32)
    if uvs is not None:
        geo["uvs"] = np.asarray(uvs, np.float32)
    if curve_indices is not None:
        geo["curve_indices"] = np.asarray(curve_indices, np.uint32)
        geo["curve_vertices"] = np.asarray(curve_vertices, np.float32)
    sizes = {}
    for (n, dt, cols), k in zip(R.GEOMETRY_BUFFERS, R.SIZE_FIELDS):
        sizes[k] = np.asarray(geo.get(n, np.zeros(0, dt)), dt).nbytes
    lo, hi = positions.min(0), positions.max(0)
    d = {"aabox": (tuple(float(x) for x in lo), tuple(float(x) for x in hi)),
         "total_vertex_count": positions.shape[0],
         "total_curve_vertex_count": 0 if curve_vertices is None else len(curve_vertices),
         "curve_model_name": R.murmur3_32("curves") if curves else 0, "pad": 0, **sizes,
         "cameras": np.array(list(cameras), R.CAMERA_DT), "texture_resource_names": np.array(list(textures), R.PATH_DT),
         "materials": np.array(list(materials), R.MATERIAL_DT), "material_hashes": np.array(list(hashes), R.HASH_DT),
         "meshes": mrec, "curves": np.array(list(curves), R.CURVE_DT)}
    if capacity_extra:
        d["carrays"] = {n: {"capacity": np.asarray(d[n]).shape[0] + capacity_extra} for n, _ in R.MODEL_ARRAYS}
    return d, geo


def box_quads(lo, hi):
    """A box as 6 outward quads (8 corners) with per-corner normals, tangents, UVs (24 vertices)."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    pos, nrm, tan, uv, quads = [], [], [], [], []
    for axis in range(3):
        for side in (0, 1):
            n = np.zeros(3, np.float32)
            n[axis] = 1.0 if side else -1.0
            u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
            if not side:
                u_ax, v_ax = v_ax, u_ax
            base = len(pos)
            for cu, cv in ((0, 0), (1, 0), (1, 1), (0, 1)):
                p = np.where(np.arange(3) == axis, hi if side else lo, 0).astype(np.float32)
                p[u_ax] = hi[u_ax] if cu else lo[u_ax]
                p[v_ax] = hi[v_ax] if cv else lo[v_ax]
                pos.append(p)
                nrm.append(n)
                t = np.zeros(4, np.float32)
                t[u_ax], t[3] = 1.0, 1.0
                tan.append(t)
                uv.append((cu, cv))
            quads.append([base, base + 1, base + 2, base + 3])
    return (np.array(pos), np.array(quads, np.uint32), np.array(nrm), np.array(tan), np.array(uv, np.float32))


def grid_quads(nu, nv, size, y=0.0):
    xs, zs = np.linspace(-size, size, nu + 1), np.linspace(-size, size, nv + 1)
    pos = np.array([(x, y, z) for z in zs for x in xs], np.float32)
    quads = []
    for j in range(nv):
        for i in range(nu):
            a = j * (nu + 1) + i
            quads.append([a, a + 1, a + nu + 2, a + nu + 1])
    return pos, np.array(quads, np.uint32)


FLOOR_PTEX_RES = 3


def floor_ptex():
    """The floor's per-face texture (4 x 2 faces at 8 x 8 texels) and its adjacency."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import ptx
    rng = np.random.default_rng(17)
    texels = rng.integers(0, 256, (8, 1 << FLOOR_PTEX_RES, 1 << FLOOR_PTEX_RES, 3), dtype=np.uint8)
    af, ae = ptx.grid_adjacency(4, 2)
    return {"res_log2": FLOOR_PTEX_RES, "texels": texels, "adjfaces": af, "adjedges": ae}


def fixture_resources():
    """{(kind, name): dict} of the fixture scene "rooms" and the relative path of its Ptex file."""
    rng = np.random.default_rng(2018)
    wood, paint, brass = (R.murmur3_32(n) for n in ("wood", "paint", "brass"))
    out = {}
    # crate: quads with attributes; body -> model material "wood", lid -> subscene material "paint"
    pos, quads, nrm, tan, uv = box_quads((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5))
    hashes = sorted([wood, brass])
    mats = [material((0.5, 0.3, 0.1), rng=rng) if h == wood else material((0.8, 0.6, 0.2), metallic=1.0,
                                                                           roughness=0.3) for h in hashes]
    out[("model", "crate")], out[("geometry", "crate")] = model_dict(
        [("crate_body", wood, quads[:5]), ("crate_lid", paint, quads[5:])], pos, mats, hashes, nrm, tan, uv,
        cameras=[camera("crate_cam", (2, 2, 2), (0, 0.5, 0))], capacity_extra=3)
    # panel: triangles, no attributes; first mesh "brass" (model), second an unknown name (default)
    tp = np.array([(-1, 0, 0), (1, 0, 0), (1, 1.5, 0), (-1, 1.5, 0), (0, 2.2, 0)], np.float32)
    thin = R.murmur3_32("thin_white")
    by_hash = {brass: material((0.8, 0.6, 0.2), metallic=1.0),
               thin: material((0.9, 0.9, 0.9), shader=R.SHADER_THIN, rng=rng)}
    ph = sorted(by_hash)
    out[("model", "panel")], out[("geometry", "panel")] = model_dict(
        [("panel_main", brass, [[0, 1, 2], [0, 2, 3]]), ("panel_top", R.murmur3_32("nothing"), [[3, 2, 4]])], tp,
        [by_hash[h] for h in ph], ph)
    # floor: a 4 x 2 quad grid with a per-face texture
    fpos, fquads = grid_quads(4, 2, 3.0)
    out[("model", "floor")], out[("geometry", "floor")] = model_dict(
        [("floor_mesh", R.murmur3_32("floor"), fquads)], fpos,
        [material((1.0, 1.0, 1.0), flags=R.assets.USES_PTEX, texture="textures~floor")], [R.murmur3_32("floor")],
        textures=[b"textures~floor"])
    ident = np.eye(4)
    out[("subscene", "sub_a")] = {
        "name": "sub_a", "light_set_index": 0,
        "model_names": np.array([b"crate", b"panel"], R.PATH_DT),
        "model_instances": np.array([instance(0, affine(0.8, 0.3, (-1.0, 0.0, 0.5))),
                                     instance(1, affine(1.0, -0.2, (0.8, 0.0, 1.2))),
                                     instance(0, affine(0.5, 1.1, (1.5, 0.0, -0.8)))], R.INSTANCE_DT),
        "scene_material_names": np.array([paint], R.HASH_DT),
        "scene_materials": np.array([material((0.1, 0.4, 0.8), roughness=0.2)], R.MATERIAL_DT)}
    out[("subscene", "sub_b")] = {
        "name": "sub_b", "light_set_index": 0,
        "model_names": np.array([b"floor", b"crate"], R.PATH_DT),
        "model_instances": np.array([instance(0, ident), instance(1, affine(0.6, 0.7, (-2.0, 0.0, -1.5)))],
                                    R.INSTANCE_DT),
        "scene_material_names": np.array([R.murmur3_32("other"), paint], R.HASH_DT),
        "scene_materials": np.array([material((0.2, 0.2, 0.2)), material((0.8, 0.1, 0.1), roughness=0.6)],
                                    R.MATERIAL_DT),
        "carrays": {"model_instances": {"capacity": 16}}}
    lights = np.zeros(3, R.LIGHT_DT)
    lights[0] = (0, (0.0, 4.0, 0.0), (0.0, -1.0, 0.0), (1.5, 0.0, 0.0), (0.0, 0.0, 1.5), (8.0, 8.0, 7.0))
    lights[1] = (0, (-2.5, 3.0, 2.0), (0.5, -1.0, -0.3), (0.8, 0.0, 0.0), (0.0, 0.0, 0.8), (5.0, 4.0, 3.0))
    lights[2] = (0, (2.0, 3.5, -2.0), (0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 2.0, 3.0))
    out[("scene", "rooms")] = {
        "name": "rooms", "ibl_name": "sky", "background_intensity": (0.1, 0.2, 0.3, 1.0),
        "subscene_names": np.array([b"sub_a", b"sub_b"], R.PATH_DT),
        "subscene_instances": np.array([instance(1, ident), instance(0, affine(1.0, 0.0, (0.0, 0.0, 0.0))),
                                        instance(0, affine(0.7, 2.5, (0.5, 0.0, -2.0)))], R.INSTANCE_DT),
        "lights": lights, "light_set_ranges": np.array([(0, 2), (2, 1)], R.RANGE_DT),
        "cameras": np.array([camera("broken", (0, 0, 0), (0, 0, 1), fov=0.0),
                             camera("main", (0.0, 3.5, -6.0), (0.0, 0.5, 0.5), fov=55.0)], R.CAMERA_DT)}
    h, w = 8, 16
    light = np.zeros((h, w, 3), np.float32)
    light[..., 0] = np.linspace(0.2, 1.0, w)[None, :]
    light[..., 1] = np.linspace(1.0, 0.3, h)[:, None]
    light[..., 2] = 0.5
    light[1, 3] = (20.0, 18.0, 15.0)
    out[("ibl", "sky")] = {"width": w, "height": h, "marginal": rng.uniform(0, 1, h).astype(np.float32),
                           "conditional": rng.uniform(0, 1, (h, w)).astype(np.float32), "miss_width": 0,
                           "miss_height": 0, "rotation": 0.3, "exposure_scale": 0.8, "light": light,
                           "miss": np.zeros((0, 0, 3), np.float32)}
    return out, os.path.join("textures", "floor", "floor_mesh.ptx")


def write_tree(root: str, files: dict, writer) -> None:
    """Writes {(kind, name): dict} under root's _Assets tree with writer(kind, dict) -> bytes."""
    for (kind, name), d in files.items():
        path = R.asset_path(root, kind, name)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "wb") as f:
            f.write(writer(kind, d))


def py_write(kind: str, d: dict) -> bytes:
    return {"model": R.write_model_resource, "geometry": R.write_model_geometry,
            "subscene": R.write_subscene_resource, "scene": R.write_scene_resource,
            "ibl": R.write_ibl_resource}[kind](d)
