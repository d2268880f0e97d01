"""Selas's binary resource files read (and written), and a scene assembled from them for the device.

Selas's offline build writes each asset to <root>/_Assets/<type>/<version>/<name>.bin
(AssetFileUtils::AssetFilePath, Source/Core/Assets/AssetFileUtils.cpp:128-135) and the renderer
reads them back with ReadModelResource / LoadModelGeometry (Source/Core/SceneLib/ModelResource.cpp:
432-472), ReadSubsceneResource (SubsceneResource.cpp:102-113), ReadSceneResource
(SceneResource.cpp:271-282) and ReadImageBasedLightResource (ImageBasedLightResource.cpp:59-72).

The file format is that of SerializeToBinary / AttachToBinary (Source/Core/IoLib/
BinaryStreamSerializer.h:14-44, BinarySerializers.cpp:39-79, SizeSerializer.cpp:11-27):

* the root record's fields, little-endian and back to back in the order of the type's Serialize
  function, each pointer a 64-bit slot; a CArray is (uint32 flags, uint32 padding, uint64 count,
  uint64 capacity, pointer) (ContainersLib/CArray.h:57-62, 239-249) and its elements are stored as
  raw records; a fixed string is its whole buffer;
* then every pointer's data, in field order: the slot holds the data's offset from the start of
  the file; the data is followed by zero bytes up to the pointer's alignment (16 for the geometry
  buffers, ModelResource.cpp:27 and 93-100; none for arrays). AttachToBinary reads the root record
  in place, so the serialized fields coincide with the C++ layout (the record sizes are checked
  against the reference's sizeof in the tests).

The readers return dicts of plain values and numpy arrays (structured dtypes below hold the raw
records); the writers take the same dicts and produce byte-identical files to the reference's
writer. load_scene follows InitializeSceneResource / InitializeSubsceneResource /
InitializeModelResource / LoadModelGeometry to build the instanced scene description that
selas_amd.native.build_render_scene commits to the device; what the device path does not render
is refused (or, with strict=False, dropped and listed).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import assets

# (type directory, kDataVersion) of each resource (ModelResource.cpp:23-26,
# SubsceneResource.cpp:18-19, SceneResource.cpp:22-23, ImageBasedLightResource.cpp:19-20)
KINDS = {
    "model": ("ModelResource", 1539636447),
    "geometry": ("ModelGeometryResource", 1539636447),
    "subscene": ("SubsceneResource", 1539551841),
    "scene": ("SceneResource", 1539147471),
    "ibl": ("IBL", 1538087123),
}
GEOMETRY_ALIGNMENT = 16  # ModelResource::kGeometryDataAlignment
POINTER_KEY = 0x1234567812345678  # the writer's placeholder before a pointer is written (BinarySerializers.cpp:12)

CAMERA_DT = np.dtype([("name", "S128"), ("position", "<f4", 3), ("fov_degrees", "<f4"), ("look_at", "<f4", 3),
                      ("znear", "<f4"), ("up", "<f4", 3), ("zfar", "<f4")])  # CameraSettings, Camera.h:17-26
PATH_DT = np.dtype("S512")  # FilePathString = FixedString<MaxPath_ = 512>
MATERIAL_DT = np.dtype([("base_color_texture", "S512"), ("shader", "<u4"), ("flags", "<u4"),
                        ("base_color", "<f4", 3), ("transmittance_color", "<f4", 3),
                        ("scalars", "<f4", 14)])  # MaterialResourceData, ModelResource.h:68-88
MESH_DT = np.dtype([("index_count", "<u4"), ("index_offset", "<u4"), ("vertex_count", "<u4"),
                    ("vertex_offset", "<u4"), ("material_hash", "<u4"), ("indices_per_face", "<u4"),
                    ("name_hash", "<u4"), ("name", "S64")])  # MeshMetaData, ModelResource.h:106-116
CURVE_DT = np.dtype([("index_offset", "<u4"), ("index_count", "<u4")])  # CurveMetaData, :100-104
INSTANCE_DT = np.dtype([("index", "<u8"), ("local_to_world", "<f4", (4, 4))])  # Instance, SubsceneResource.h
# SceneLight (SceneResource.h:32-40): the layout of sgpu_light, so the record goes to the device as it is
LIGHT_DT = np.dtype([("type", "<u4"), ("position", "<f4", 3), ("direction", "<f4", 3), ("x", "<f4", 3),
                     ("z", "<f4", 3), ("radiance", "<f4", 3)])
RANGE_DT = np.dtype([("start", "<u4"), ("count", "<u4")])  # SceneLightSetRange, :42-46
HASH_DT = np.dtype("<u4")  # Hash32

# the eight geometry buffers of ModelGeometryData (ModelResource.h:148-167): name, element dtype, columns
GEOMETRY_BUFFERS = (("indices", "<u4", 1), ("face_index_counts", "<u4", 1), ("positions", "<f4", 3),
                    ("normals", "<f4", 3), ("tangents", "<f4", 4), ("uvs", "<f4", 2),
                    ("curve_indices", "<u4", 1), ("curve_vertices", "<f4", 4))
# the model's eight size fields, in the same order (ModelResourceData, ModelResource.h:127-134)
SIZE_FIELDS = ("index_size", "face_index_size", "position_size", "normals_size", "tangents_size", "uvs_size",
               "curve_index_size", "curve_vertex_size")
MODEL_ARRAYS = (("cameras", CAMERA_DT), ("texture_resource_names", PATH_DT), ("materials", MATERIAL_DT),
                ("material_hashes", HASH_DT), ("meshes", MESH_DT), ("curves", CURVE_DT))
SUBSCENE_ARRAYS = (("model_names", PATH_DT), ("model_instances", INSTANCE_DT), ("scene_material_names", HASH_DT),
                   ("scene_materials", MATERIAL_DT))
SCENE_ARRAYS = (("subscene_names", PATH_DT), ("subscene_instances", INSTANCE_DT), ("lights", LIGHT_DT),
                ("light_set_ranges", RANGE_DT), ("cameras", CAMERA_DT))

# ShaderType (ModelResource.h:27-34) and the surface-block record of include/selas_gpu.h
SHADER_THIN, SHADER_SOLID, SHADER_DIRAC = 0, 1, 2


class ResourceError(ValueError):
    """A resource file that cannot be read (truncated, an unwritten pointer, a misaligned buffer)."""


def asset_path(root: str, kind: str, name: str) -> str:
    """<root>/_Assets/<type>/<version>/<name>.bin (AssetFileUtils.cpp:128-135)."""
    t, version = KINDS[kind]
    return os.path.join(root, "_Assets", t, str(version), name + ".bin")


def fixed_str(b) -> str:
    """A fixed string's text: its bytes up to the first NUL."""
    return bytes(b).split(b"\0", 1)[0].decode("latin-1")


def to_fixed(s: str, size: int) -> bytes:
    b = s.encode("latin-1")
    if len(b) >= size:
        raise ValueError(f"string of {len(b)} bytes does not fit a {size}-byte fixed string")
    return b + b"\0" * (size - len(b))


def murmur3_32(data, seed: int = 0) -> int:
    """MurmurHash3_x86_32 (Austin Appleby's published MurmurHash3, the x86 32-bit variant), the
    hash Selas keys names and materials with (UtilityLib/MurmurHash.h:20)."""
    if isinstance(data, str):
        data = data.encode("latin-1")
    c1, c2, m32 = 0xCC9E2D51, 0x1B873593, 0xFFFFFFFF
    h = seed & m32
    n = len(data)
    nblocks = n // 4

    def rotl(x, r):
        return ((x << r) | (x >> (32 - r))) & m32

    for i in range(nblocks):
        k = struct.unpack_from("<I", data, 4 * i)[0]
        k = (k * c1) & m32
        k = rotl(k, 15)
        k = (k * c2) & m32
        h ^= k
        h = rotl(h, 13)
        h = (h * 5 + 0xE6546B64) & m32
    tail = data[4 * nblocks:]
    k = 0
    if len(tail) >= 3:
        k ^= tail[2] << 16
    if len(tail) >= 2:
        k ^= tail[1] << 8
    if len(tail) >= 1:
        k ^= tail[0]
        k = (k * c1) & m32
        k = rotl(k, 15)
        k = (k * c2) & m32
        h ^= k
    h ^= n & m32
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & m32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & m32
    h ^= h >> 16
    return h


# ---- reading ------------------------------------------------------------------------------------

class _Reader:
    def __init__(self, src, what: str):
        if isinstance(src, (bytes, bytearray, memoryview)):
            self.buf = bytes(src)
        else:
            with open(src, "rb") as f:
                self.buf = f.read()
        self.what = what
        self.at = 0
        self.offsets = []  # each pointer's data offset, in field order

    def take(self, fmt: str):
        size = struct.calcsize(fmt)
        if self.at + size > len(self.buf):
            raise ResourceError(f"{self.what}: file ends inside its root record ({len(self.buf)} bytes)")
        v = struct.unpack_from(fmt, self.buf, self.at)
        self.at += size
        return v

    def raw(self, size: int) -> bytes:
        if self.at + size > len(self.buf):
            raise ResourceError(f"{self.what}: file ends inside its root record ({len(self.buf)} bytes)")
        b = self.buf[self.at:self.at + size]
        self.at += size
        return b

    def pointer(self, size: int, alignment: int = 0) -> bytes:
        (off,) = self.take("<Q")
        self.offsets.append(off)
        if off == POINTER_KEY:
            raise ResourceError(f"{self.what}: a pointer was never written (the file is unfinished)")
        if off + size > len(self.buf):
            raise ResourceError(f"{self.what}: {size} bytes at offset {off} run past the file's end ({len(self.buf)})")
        if alignment and off % alignment:
            raise ResourceError(f"{self.what}: data at offset {off} is not {alignment}-byte aligned")
        return self.buf[off:off + size]

    def carray(self, dtype: np.dtype):
        flags, padding, count, capacity = self.take("<IIQQ")
        data = self.pointer(count * dtype.itemsize)
        return np.frombuffer(data, dtype).copy(), {"flags": flags, "padding": padding, "capacity": capacity,
                                                   "offset": self.offsets[-1]}


def read_model_resource(src) -> dict:
    """ModelResourceData (ModelResource.h:118-142, Serialize :56-79): a path or the file's bytes."""
    r = _Reader(src, "ModelResource")
    box = r.take("<6f")
    out = {"aabox": (box[:3], box[3:])}
    (out["total_vertex_count"], out["total_curve_vertex_count"], out["curve_model_name"],
     out["pad"]) = r.take("<4I")
    for name, v in zip(SIZE_FIELDS, r.take("<8Q")):
        out[name] = v
    out["carrays"] = {}
    for name, dt in MODEL_ARRAYS:
        out[name], out["carrays"][name] = r.carray(dt)
    return out


def read_model_geometry(src) -> dict:
    """ModelGeometryData (ModelResource.h:148-167, Serialize :82-101): the eight buffers as arrays
    (positions n x 3, normals n x 3, tangents n x 4, uvs n x 2, curve vertices n x 4)."""
    r = _Reader(src, "ModelGeometryResource")
    sizes = r.take("<8Q")
    out = {}
    for (name, dt, cols), size in zip(GEOMETRY_BUFFERS, sizes):
        item = np.dtype(dt).itemsize * cols
        if size % item:
            raise ResourceError(f"ModelGeometryResource: {name} holds {size} bytes, not whole {item}-byte elements")
        a = np.frombuffer(r.pointer(size, GEOMETRY_ALIGNMENT), dt).copy()
        out[name] = a.reshape(-1, cols) if cols > 1 else a
    out["offsets"] = list(r.offsets)
    return out


def read_subscene_resource(src) -> dict:
    """SubsceneResourceData (SubsceneResource.h:33-41, Serialize SubsceneResource.cpp:37-46)."""
    r = _Reader(src, "SubsceneResource")
    out = {"name": fixed_str(r.raw(512))}
    (out["light_set_index"],) = r.take("<Q")
    out["carrays"] = {}
    for name, dt in SUBSCENE_ARRAYS:
        out[name], out["carrays"][name] = r.carray(dt)
    return out


def read_scene_resource(src) -> dict:
    """SceneResourceData (SceneResource.h:48-57, Serialize SceneResource.cpp:56-66)."""
    r = _Reader(src, "SceneResource")
    out = {"name": fixed_str(r.raw(512)), "ibl_name": fixed_str(r.raw(512)), "background_intensity": r.take("<4f")}
    out["carrays"] = {}
    for name, dt in SCENE_ARRAYS:
        out[name], out["carrays"][name] = r.carray(dt)
    return out


def read_ibl_resource(src) -> dict:
    """ImageBasedLightResourceData (ImageBasedLightResource.h:13-32, Serialize
    ImageBasedLightResource.cpp:23-43): density functions (height marginal, height x width
    conditional floats), the light map (height x width x 3) and the miss map (missHeight x
    missWidth x 3, empty for a single-map light)."""
    r = _Reader(src, "IBL")
    w, h = r.take("<2Q")
    marginal = np.frombuffer(r.pointer(4 * h), "<f4").copy()
    conditional = np.frombuffer(r.pointer(4 * w * h), "<f4").copy().reshape(h, w)
    mw, mh = r.take("<2Q")
    rotation, exposure = r.take("<2f")
    light = np.frombuffer(r.pointer(12 * w * h), "<f4").copy().reshape(h, w, 3)
    miss = np.frombuffer(r.pointer(12 * mw * mh), "<f4").copy().reshape(mh, mw, 3)
    return {"width": w, "height": h, "marginal": marginal, "conditional": conditional, "miss_width": mw,
            "miss_height": mh, "rotation": rotation, "exposure_scale": exposure, "light": light, "miss": miss,
            "offsets": list(r.offsets)}


# ---- writing ------------------------------------------------------------------------------------

class _Writer:
    """CBinaryWriteSerializer's two passes in one: root fields with pointer slots, then each
    pointer's data in slot order with its tail padded to the pointer's alignment."""

    def __init__(self):
        self.raw = bytearray()
        self.payloads = []

    def put(self, fmt: str, *vals):
        self.raw += struct.pack(fmt, *vals)

    def bytes_(self, b: bytes):
        self.raw += b

    def pointer(self, data: bytes, alignment: int = 0):
        self.payloads.append((len(self.raw), bytes(data), alignment))
        self.raw += struct.pack("<Q", POINTER_KEY)

    def carray(self, a, dtype: np.dtype, meta=None):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype))
        meta = meta or {}
        count = a.shape[0] if a.ndim else 0
        cap = meta.get("capacity", count)
        if cap < count:
            raise ValueError("a CArray's capacity is below its count")
        self.put("<IIQQ", meta.get("flags", 0), meta.get("padding", 0), count, cap)
        self.pointer(a.tobytes())

    def finish(self) -> bytes:
        out = bytearray(self.raw)
        for slot, data, alignment in self.payloads:
            out[slot:slot + 8] = struct.pack("<Q", len(out))
            out += data
            if alignment:
                out += b"\0" * (-len(data) % alignment)
        return bytes(out)


def _fixed_bytes(v, size):
    return v if isinstance(v, (bytes, bytearray)) and len(v) == size else to_fixed(v, size)


def write_model_resource(d: dict) -> bytes:
    """The file read_model_resource reads (fields as it returns them; "carrays" optional)."""
    w = _Writer()
    lo, hi = d["aabox"]
    w.put("<6f", *lo, *hi)
    w.put("<4I", d["total_vertex_count"], d["total_curve_vertex_count"], d.get("curve_model_name", 0), d.get("pad", 0))
    w.put("<8Q", *[d[k] for k in SIZE_FIELDS])
    meta = d.get("carrays", {})
    for name, dt in MODEL_ARRAYS:
        w.carray(d[name], dt, meta.get(name))
    return w.finish()


def write_model_geometry(d: dict) -> bytes:
    """The file read_model_geometry reads (absent buffers are empty)."""
    w = _Writer()
    arrs = []
    for name, dt, cols in GEOMETRY_BUFFERS:
        a = np.ascontiguousarray(np.asarray(d.get(name, np.zeros(0, dt)), dt))
        arrs.append(a.tobytes())
    w.put("<8Q", *[len(a) for a in arrs])
    for a in arrs:
        w.pointer(a, GEOMETRY_ALIGNMENT)
    return w.finish()


def write_subscene_resource(d: dict) -> bytes:
    w = _Writer()
    w.bytes_(_fixed_bytes(d["name"], 512))
    w.put("<Q", d["light_set_index"])
    meta = d.get("carrays", {})
    for name, dt in SUBSCENE_ARRAYS:
        w.carray(d[name], dt, meta.get(name))
    return w.finish()


def write_scene_resource(d: dict) -> bytes:
    w = _Writer()
    w.bytes_(_fixed_bytes(d["name"], 512))
    w.bytes_(_fixed_bytes(d["ibl_name"], 512))
    w.put("<4f", *d["background_intensity"])
    meta = d.get("carrays", {})
    for name, dt in SCENE_ARRAYS:
        w.carray(d[name], dt, meta.get(name))
    return w.finish()


def write_ibl_resource(d: dict) -> bytes:
    w = _Writer()
    wd, ht = d["width"], d["height"]
    w.put("<2Q", wd, ht)
    w.pointer(np.ascontiguousarray(d["marginal"], "<f4").tobytes())
    w.pointer(np.ascontiguousarray(d["conditional"], "<f4").tobytes())
    w.put("<2Q", d["miss_width"], d["miss_height"])
    w.put("<2f", d["rotation"], d["exposure_scale"])
    w.pointer(np.ascontiguousarray(d["light"], "<f4").tobytes())
    w.pointer(np.ascontiguousarray(d["miss"], "<f4").tobytes())
    return w.finish()


# ---- materials, cameras -----------------------------------------------------------------------------

def default_material() -> np.ndarray:
    """CreateDefaultMaterial (ModelResource.cpp:203-211) as a MATERIAL_DT record: base colour 0.6,
    DisneySolid, ior 1.5, every other value 0."""
    m = np.zeros((), MATERIAL_DT)
    m["base_color"] = 0.6
    m["shader"] = SHADER_SOLID
    m["scalars"][assets.ATTRIBUTES.index("ior")] = 1.5
    return m


def binary_search(data, key: int) -> int:
    """BinarySearch (UtilityLib/BinarySearch.h:11-35): the first index whose value is >= key by
    bisection, returned when that value equals key, else -1. (When every value is below key the
    reference then reads data[-1]; here that is not found.)"""
    index = -1
    low, high = 0, len(data) - 1
    while low <= high:
        mid = (low + high) // 2
        if int(data[mid]) >= key:
            index = mid
            high = mid - 1
        else:
            low = mid + 1
    return index if index >= 0 and int(data[index]) == key else -1


def find_mesh_material(model: dict, scene_material_names, scene_materials, material_hash: int):
    """FindMeshMaterial (ModelResource.cpp:213-232): the model's own material by binary search over
    its materialHashes, else the first subscene material of that name, else the default.
    Returns (record, where) with where in ("model", i), ("subscene", i), ("default", -1)."""
    n = model["materials"].shape[0]
    if n != 0:
        i = binary_search(model["material_hashes"][:n], material_hash)
        if i >= 0:
            return model["materials"][i], ("model", i)
    for i, h in enumerate(scene_material_names):
        if int(h) == material_hash:
            return scene_materials[i], ("subscene", i)
    return default_material(), ("default", -1)


def material_record(m) -> np.ndarray:
    """A MaterialResourceData record as the 22-float record sgpu_scene_set_materials takes (the
    layout assets.build_material fills: colours, the scalar slots, shader, flags)."""
    if int(m["shader"]) == SHADER_DIRAC:
        raise ValueError("DiracTransparent materials are not supported by the device path")
    if int(m["shader"]) not in (SHADER_THIN, SHADER_SOLID):
        raise ValueError(f"unknown shader {int(m['shader'])}")
    rec = np.zeros(22, np.float32)
    rec[0:3] = m["base_color"]
    rec[3:6] = m["transmittance_color"]
    for attr, slot in assets.SLOTS.items():
        rec[slot] = m["scalars"][assets.ATTRIBUTES.index(attr)]
    rec[20] = float(m["shader"])
    rec[21] = float(m["flags"])
    return rec


def valid_camera(c) -> bool:
    """ValidCamera (Camera.cpp:50-53)."""
    return float(c["fov_degrees"]) > 0 and float(c["zfar"]) > float(c["znear"])


def scene_camera(scene: dict, index: int = 0) -> dict:
    """SetupSceneCamera (SceneResource.cpp:336-358): the scene's camera `index` when it is valid,
    else DefaultCameraSettings (Camera.cpp:39-47), as the camera dict the renderers take. An index
    past the scene's cameras takes the default (the reference would read past its array)."""
    cams = scene["cameras"]
    if index < cams.shape[0] and valid_camera(cams[index]):
        c = cams[index]
        return {"position": tuple(float(x) for x in c["position"]), "look_at": tuple(float(x) for x in c["look_at"]),
                "up": tuple(float(x) for x in c["up"]), "fov": float(c["fov_degrees"])}
    return {"position": (0.0, 0.0, 5.0), "look_at": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 35.0}


def instance_matrix(inst) -> np.ndarray:
    """An Instance's localToWorld as a 4 x 4 matrix acting on column vectors. Selas stores rows
    r0..r3 for row vectors (p' = p M, translation in r3: MatrixMultiplyPoint, FloatFuncs.cpp:383-396)
    and hands the same 16 floats to Embree as RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR
    (SubsceneResource.cpp:72-73): both read them as the transpose of the stored rows."""
    return np.asarray(inst["local_to_world"], np.float64).T.copy()


# ---- scene assembly -----------------------------------------------------------------------------------

@dataclass
class LoadedScene:
    """A scene read from resource files: `description` is the instanced scene dict
    selas_amd.native.build_render_scene commits (models, subscenes, instances, lights, background,
    ibl); `camera` is SetupSceneCamera's; `dropped` lists what strict=False left out."""
    description: dict
    camera: dict
    resources: dict = field(default_factory=dict)
    dropped: list = field(default_factory=list)


def _mesh_faces(indices: np.ndarray, mesh, n_verts: int, where: str) -> np.ndarray:
    """InitializeMeshes (ModelResource.cpp:236-298): the mesh's faces from the shared index buffer,
    indexCount / indicesPerFace of them from indexOffset, triangles for 3 indices per face and quads
    (the first 4 indices of each face, at a stride of indicesPerFace) otherwise."""
    ipf, count, off = int(mesh["indices_per_face"]), int(mesh["index_count"]), int(mesh["index_offset"])
    if ipf < 3:
        raise ResourceError(f"{where}: {ipf} indices per face")
    nf = count // ipf
    if off + nf * ipf > indices.shape[0]:
        raise ResourceError(f"{where}: indices {off}..{off + nf * ipf} run past the index buffer ({indices.shape[0]})")
    faces = indices[off:off + nf * ipf].reshape(nf, ipf)
    faces = faces if ipf == 3 else faces[:, :4]
    if nf and int(faces.max()) >= n_verts:
        raise ResourceError(f"{where}: vertex index {int(faces.max())} beyond the {n_verts} vertices")
    return np.ascontiguousarray(faces, np.uint32)


def build_model(model: dict, geometry: dict, name: str, scene_material_names, scene_materials, content_root=None,
                ptex_loader=None, strict=True, dropped=None) -> dict:
    """One model as a scene dict of geometries (meshes in order, then its round B-spline curves, so
    geometry ids are Selas's: InitializeMeshes then InitializeCurves), materials and geom_material. Each mesh keeps the vertices its faces use (indices
    rebased), with the model's normals, tangents and UVs over them (SetMeshVertexAttributes,
    ModelResource.cpp:164-200). Materials follow InitializeGeometryUserDatas
    (ModelResource.cpp:335-371): FindMeshMaterial per mesh; a per-face texture
    ("<baseColorTexture>\\<mesh name>.ptx" under the content root) comes from ptex_loader(path) ->
    {"res_log2", "texels", "adjfaces", "adjedges"}."""
    dropped = [] if dropped is None else dropped
    nv = int(model["total_vertex_count"])
    pos = geometry["positions"]
    if pos.shape[0] < nv:
        raise ResourceError(f"model {name}: {pos.shape[0]} positions for totalVertexCount {nv}")
    attrs = {}
    for k in ("normals", "tangents", "uvs"):
        a = geometry[k]
        if a.shape[0]:
            if a.shape[0] < nv:
                raise ResourceError(f"model {name}: {a.shape[0]} {k} for totalVertexCount {nv}")
            attrs[k] = a
    if "tangents" in attrs and "normals" not in attrs or "uvs" in attrs and len(attrs) < 3:
        raise ResourceError(f"model {name}: tangents need normals and UVs need normals and tangents")

    def drop(what):
        if strict:
            raise ValueError(f"model {name}: {what} (not rendered by the device path; strict=False drops it)")
        dropped.append(f"model {name}: {what}")

    geoms, mats, gm, keys = [], [], [], {}
    for mi, mesh in enumerate(model["meshes"]):
        mname = fixed_str(mesh["name"])
        where = f"model {name} mesh {mi} ({mname})"
        faces = _mesh_faces(geometry["indices"], mesh, nv, where)
        lo = int(faces.min()) if faces.size else 0
        hi = int(faces.max()) + 1 if faces.size else 0
        g = {"positions": np.ascontiguousarray(pos[lo:hi], np.float32)}
        g["triangles" if faces.shape[1] == 3 else "quads"] = faces - np.uint32(lo)
        if attrs:
            g["normals"] = np.ascontiguousarray(attrs["normals"][lo:hi], np.float32)
            if "tangents" in attrs:
                g["tangents"] = np.ascontiguousarray(attrs["tangents"][lo:hi], np.float32)
            if "uvs" in attrs:
                g["uvs"] = np.ascontiguousarray(attrs["uvs"][lo:hi], np.float32)
        m, _ = find_mesh_material(model, scene_material_names, scene_materials, int(mesh["material_hash"]))
        rec = material_record(m)
        tex = fixed_str(m["base_color_texture"])
        flags = int(m["flags"])
        if flags & assets.USES_PTEX:
            rel = f"{tex}\\{mname}.ptx".replace("~", os.sep).replace("\\", os.sep)
            path = os.path.join(content_root or "", "Content", rel)
            if ptex_loader is None:
                drop(f"mesh {mi}: per-face texture {path} (no ptex_loader)")
                rec[21] = float(flags & ~assets.USES_PTEX)
            else:
                t = ptex_loader(path)
                want = 1 if faces.shape[1] == 4 else 0  # Ptex mesh types: 0 triangle, 1 quad
                if t.get("mesh_type", want) != want:
                    raise ResourceError(f"{where}: texture {path} is for a {('triangle', 'quad')[t['mesh_type']]} "
                                        "mesh")
                g["ptex"] = t
        elif tex:
            drop(f"mesh {mi}: UV base-colour texture {tex}")
        key = rec.tobytes()
        if key not in keys:
            keys[key] = len(mats)
            mats.append(rec)
        geoms.append(g)
        gm.append(keys[key])
    # InitializeCurves (ModelResource.cpp:300-330): one round B-spline geometry per CurveMetaData,
    # attached after the meshes (geometry ids continue), indexing the model's shared curve vertices;
    # every curve takes the material FindMeshMaterial finds for curveModelName (:372-385) and no
    # texture (its baseColorTextureHandle is empty: a per-face texture flag would have no texture)
    if model["curves"].shape[0]:
        ncv = int(model["total_curve_vertex_count"])
        cv, cidx = geometry["curve_vertices"], geometry["curve_indices"]
        if cv.shape[0] < ncv:
            raise ResourceError(f"model {name}: {cv.shape[0]} curve vertices for totalCurveVertexCount {ncv}")
        m, _ = find_mesh_material(model, scene_material_names, scene_materials, int(model["curve_model_name"]))
        rec = material_record(m)
        rec[21] = float(int(m["flags"]) & ~assets.USES_PTEX)
        key = rec.tobytes()
        if key not in keys:
            keys[key] = len(mats)
            mats.append(rec)
        verts = np.ascontiguousarray(cv[:ncv], np.float32)
        for ci, curve in enumerate(model["curves"]):
            off, cnt = int(curve["index_offset"]), int(curve["index_count"])
            if off + cnt > cidx.shape[0]:
                raise ResourceError(f"model {name} curve {ci}: indices {off}..{off + cnt} run past the curve index "
                                    f"buffer ({cidx.shape[0]})")
            idx = np.ascontiguousarray(cidx[off:off + cnt], np.uint32)
            if cnt and int(idx.max()) + 3 >= ncv:
                raise ResourceError(f"model {name} curve {ci}: a segment's control points run past the {ncv} curve "
                                    "vertices")
            geoms.append({"curve_vertices": verts, "curves": idx})
            gm.append(keys[key])
    if not mats:
        mats.append(material_record(default_material()))
    return {"geoms": geoms, "materials": np.stack(mats), "geom_material": np.array(gm, np.uint32)}


def load_scene(root: str, scene_name: str, camera_index: int = 0, ptex_loader=None, strict: bool = True,
               reader=None) -> LoadedScene:
    """ReadSceneResource + InitializeSceneResource (SceneResource.cpp:271-312): the scene's subscenes
    (ReadSubsceneResource + InitializeSubsceneResource, SubsceneResource.cpp:102-136), their models
    (ReadModelResource + LoadModelGeometry), instances at both levels in file order (instance ids as
    Selas numbers them, SceneResource.cpp:236 and SubsceneResource.cpp:76), the light set the
    subscenes' lightSetIndex names (CreateSceneLightSets, :196-206; an index past the sets means no
    lights, as NextEventEstimation returns nothing for it, AreaLighting.cpp:357-359), the background
    intensity, the image-based light and the camera SetupSceneCamera picks. `reader(kind, name)`
    returns a file's bytes (default: asset_path under root)."""
    if reader is None:
        def reader(kind, name):
            with open(asset_path(root, kind, name), "rb") as f:
                return f.read()
    dropped = []
    scene = read_scene_resource(reader("scene", scene_name))
    models, model_keys, subs, light_sets, used = [], {}, [], [], set()
    placed = {int(i["index"]) for i in scene["subscene_instances"]}
    for si, sub_name in enumerate(scene["subscene_names"]):
        sub = read_subscene_resource(reader("subscene", fixed_str(sub_name)))
        names = [fixed_str(n) for n in sub["model_names"]]
        local = []
        for mn in names:
            md = read_model_resource(reader("model", mn))
            geo = read_model_geometry(reader("geometry", mn))
            built = build_model(md, geo, mn, sub["scene_material_names"], sub["scene_materials"], root, ptex_loader,
                                strict, dropped)
            key = (mn, built["materials"].tobytes(), built["geom_material"].tobytes())
            if key not in model_keys:
                model_keys[key] = len(models)
                models.append(built)
            local.append(model_keys[key])
        inst = []
        for mi in sub["model_instances"]:
            if int(mi["index"]) >= len(local):
                raise ResourceError(f"subscene {sub['name']}: instance of model {int(mi['index'])} of {len(local)}")
            inst.append((local[int(mi["index"])], instance_matrix(mi)))
        subs.append(inst)
        if si in placed and inst:
            light_sets.append(int(sub["light_set_index"]))
    instances = []
    for s in scene["subscene_instances"]:
        if int(s["index"]) >= len(subs):
            raise ResourceError(f"scene {scene['name']}: instance of subscene {int(s['index'])} of {len(subs)}")
        instances.append(("subscene", int(s["index"]), instance_matrix(s)))
    # one light set for the whole scene (the device holds one per scene)
    lights = np.zeros(0, LIGHT_DT)
    sets = sorted(set(light_sets))
    if len(sets) > 1:
        if strict:
            raise ValueError(f"scene {scene['name']}: subscenes use light sets {sets}; the device holds one light "
                             "set per scene (strict=False keeps the first)")
        dropped.append(f"light sets {sets[1:]}")
    if sets:
        s = light_sets[0]
        ranges = scene["light_set_ranges"]
        if s < ranges.shape[0]:
            a, n = int(ranges[s]["start"]), int(ranges[s]["count"])
            if a + n > scene["lights"].shape[0]:
                raise ResourceError(f"scene {scene['name']}: light set {s} runs past the {scene['lights'].shape[0]} lights")
            lights = scene["lights"][a:a + n].copy()
    if lights.shape[0] and np.any(lights["type"] != 0):
        raise ResourceError(f"scene {scene['name']}: light type other than QuadLight (0)")
    desc = {"models": models, "subscenes": subs, "instances": instances, "lights": lights,
            "background": np.asarray(scene["background_intensity"][:3], np.float32)}
    ibl_name = scene["ibl_name"]
    if ibl_name:
        ibl = read_ibl_resource(reader("ibl", ibl_name))
        if ibl["miss_width"] and ibl["miss_height"]:
            if strict:
                raise ValueError(f"IBL {ibl_name}: a separate miss map (the device's image-based light uses one map "
                                 "for lighting and misses; strict=False uses the light map for both)")
            dropped.append(f"IBL {ibl_name}: miss map")
        desc["ibl"] = {"rgb": ibl["light"], "rotation": float(ibl["rotation"]), "exposure": float(ibl["exposure_scale"])}
    return LoadedScene(description=desc, camera=scene_camera(scene, camera_index),
                       resources={"scene": scene}, dropped=dropped)
