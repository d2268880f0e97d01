"""ctypes binding of libselas_gpu (the C ABI declared in include/selas_gpu.h).

This module is the Python host side used by tests and benchmarks. It mirrors the names of the
C entry points one to one; the HIP code lives only in the shared library. Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libselas_gpu.so")

SGPU_OK = 0
SGPU_E_INVALID = -22
SGPU_E_NOMEM = -12
SGPU_E_STATE = -16
SGPU_E_HIP = -5
SGPU_E_RANGE = -34
INVALID_ID = 0xFFFFFFFF


class SgpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libselas_gpu error {code}: {msg}")
        self.code = code


class RaysSoA(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")]


class HitsSoA(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("t", "u", "v", "prim_id", "geom_id", "ng_x", "ng_y", "ng_z", "inst_id0",
                                          "inst_id1")]


class TraceStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "nodes_visited", "tris_tested", "bytes_node", "bytes_tri",
                                          "curve_nodes_visited", "curve_segments_tested")]


_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float
_FP = C.POINTER(C.c_float)
_U32P = C.POINTER(C.c_uint32)

# name -> (argtypes); every function returns int
_SIGNATURES = {
    "sgpu_context_create": [C.c_int, C.POINTER(_P)],
    "sgpu_context_destroy": [_P],
    "sgpu_set_tuning": [_P, _P],
    "sgpu_get_tuning": [_P, _P],
    "sgpu_get_device_info": [_P, _P],
    "sgpu_malloc": [_P, C.c_size_t, C.POINTER(_P)],
    "sgpu_free": [_P, _P],
    "sgpu_memcpy_htod": [_P, _P, C.c_size_t, _P],
    "sgpu_memcpy_dtoh": [_P, _P, C.c_size_t, _P],
    "sgpu_memset": [_P, C.c_int, C.c_size_t, _P],
    "sgpu_device_copy": [_P, _P, C.c_size_t, _P],
    "sgpu_stream_create": [C.POINTER(_P)],
    "sgpu_stream_destroy": [_P],
    "sgpu_stream_synchronize": [_P],
    "sgpu_device_synchronize": [],
    "sgpu_event_create": [C.POINTER(_P)],
    "sgpu_event_destroy": [_P],
    "sgpu_event_record": [_P, _P],
    "sgpu_event_elapsed_ms": [_P, _P, _FP],
    "sgpu_scene_create": [_P, C.POINTER(_P)],
    "sgpu_scene_add_triangles": [_P, _FP, _U32, _U32P, _U32, _FP, _U32P],
    "sgpu_scene_add_quads": [_P, _FP, _U32, _U32P, _U32, _FP, _U32P],
    "sgpu_scene_add_curves": [_P, _FP, _U32, _U32P, _U32, _FP, _U32P],
    "sgpu_scene_set_ptex": [_P, _U32, _U32, _U32, _P, _P, _P],
    "sgpu_scene_set_vertex_attributes": [_P, _U32, _P, _P, _P, _U32],
    "sgpu_ptex_eval": [_P, _U32, _P, _P, _P, _U32, _P, _P],
    "sgpu_scene_add_instance": [_P, _P, _FP, _U32P],
    "sgpu_scene_commit": [_P, _P],
    "sgpu_scene_release": [_P],
    "sgpu_scene_info": [_P, _U32P, _U32P, _U32P],
    "sgpu_scene_device_bytes": [_P, C.POINTER(C.c_uint64)],
    "sgpu_scene_check": [_P],
    "sgpu_intersect": [_P, C.POINTER(RaysSoA), C.POINTER(HitsSoA), _U32, _P],
    "sgpu_intersect_noalloc": [_P, C.POINTER(RaysSoA), C.POINTER(HitsSoA), _U32, _P],
    "sgpu_occluded": [_P, C.POINTER(RaysSoA), _P, _U32, _P],
    "sgpu_intersect_stats": [_P, C.POINTER(RaysSoA), C.POINTER(HitsSoA), _U32, C.c_int, C.POINTER(TraceStats)],
    "sgpu_scene_set_materials": [_P, _FP, _U32, _U32P, _U32],
    "sgpu_scene_set_lights": [_P, _P, _U32, _FP],
    "sgpu_scene_set_ibl": [_P, _U32, _U32, _FP, _F, _F],
    "sgpu_render_pt": [_P, _P, _P, _P],
    "sgpu_render_pt_stats": [_P, _P, _P, _P, _P],
    "sgpu_render_deferred_pt": [_P, _P, _P, _P],
    "sgpu_framebuffer_save_hdr": [C.c_char_p, _P, _U32, _U32, _U32, _P],
    "sgpu_render_vcm": [_P, _P, _P, _P, _P],
    "sgpu_render_vcm_batch": [_P, _P, _P, _U32, _U32, _P, _P],
    "sgpu_vcm_server_create": [_P, _P, _P, _P],
    "sgpu_vcm_server_submit": [_P, _P, _P, _P, _P, _P],
    "sgpu_vcm_server_step": [_P, _P, _P, _U32, _P, _P],
    "sgpu_vcm_server_stats": [_P, _P],
    "sgpu_vcm_server_footprint": [_P, _P, _P],
    "sgpu_vcm_server_destroy": [_P, _P],
    "sgpu_bsdf_eval": [_P, _P, _P, _P, _U32, _P, _P],
    "sgpu_bsdf_sample": [_P, _P, _P, _P, _U32, _U32, _P, _P],
    "sgpu_gen_cluster_soup": [_U64, _U32, _U32, _F, _F, _F, _FP],
    "sgpu_gen_camera_rays": [_FP, _FP, _FP, _F, _U32, _U32, _U32] + [_FP] * 8,
    "sgpu_gen_bounce_rays": [_U64, _FP, _U32, _U32] + [_FP] * 8,
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C selas_amd/csrc` (or __graft_entry__.build()); "
            "there is no CPU fallback for the HIP path")
    lib = C.CDLL(LIB_PATH)
    lib.sgpu_last_error.restype = C.c_char_p
    lib.sgpu_last_error.argtypes = []
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    return lib


lib = _load()


def declared_symbols():
    """Every entry point this binding declares (the C header's full list)."""
    return ["sgpu_last_error"] + list(_SIGNATURES)


def check(rc: int):
    if rc != SGPU_OK:
        raise SgpuError(rc, lib.sgpu_last_error().decode(errors="replace"))


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_U32P)


# ---- procedural inputs (host only; no GPU needed) ---------------------------------------------

def gen_cluster_soup(seed: int, n_clusters: int, tris_per_cluster: int, sigma=0.02, edge_min=1e-4, edge_max=1e-2):
    out = np.empty((n_clusters * tris_per_cluster, 9), np.float32)
    check(lib.sgpu_gen_cluster_soup(seed, n_clusters, tris_per_cluster, sigma, edge_min, edge_max, fptr(out)))
    return out


RAY_FIELDS = ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")


def _ray_arrays(n):
    return {k: np.empty(n, np.float32) for k in RAY_FIELDS}


def gen_camera_rays(pos, look_at, up, fov_degrees, width, height, spp):
    n = width * height * spp
    r = _ray_arrays(n)
    p = np.asarray(pos, np.float32)
    la = np.asarray(look_at, np.float32)
    u = np.asarray(up, np.float32)
    check(lib.sgpu_gen_camera_rays(fptr(p), fptr(la), fptr(u), fov_degrees, width, height, spp,
                                   *[fptr(r[k]) for k in RAY_FIELDS]))
    return r


def gen_bounce_rays(seed, tri9: np.ndarray, n_rays):
    r = _ray_arrays(n_rays)
    t = np.ascontiguousarray(tri9, np.float32)
    check(lib.sgpu_gen_bounce_rays(seed, fptr(t), t.shape[0], n_rays, *[fptr(r[k]) for k in RAY_FIELDS]))
    return r


# ---- device objects -------------------------------------------------------------------------------

class DeviceBuffer:
    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        p = _P()
        check(lib.sgpu_malloc(ctx.handle, nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib.sgpu_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, None))

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        check(lib.sgpu_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, None))
        return out

    def free(self):
        if self.ptr:
            check(lib.sgpu_free(self.ctx.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Tuning(C.Structure):
    _fields_ = [("vcm_wavefront_paths", C.c_uint64), ("trace_blocks_per_cu", C.c_uint32), ("trace_tree", C.c_uint32),
                ("trace_lds_nodes", C.c_uint32), ("vcm_graphs", C.c_uint32), ("vcm_order", C.c_uint32),
                ("vcm_merge_lanes", C.c_uint32), ("bvh_builder", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("compute_units", C.c_uint32), ("reserved", C.c_uint32),
                ("total_bytes", C.c_uint64), ("free_bytes", C.c_uint64)]


class Context:
    def __init__(self, device: int = 0):
        h = _P()
        check(lib.sgpu_context_create(device, C.byref(h)))
        self.handle = h.value

    def device_info(self) -> dict:
        d = DeviceInfo()
        check(lib.sgpu_get_device_info(self.handle, C.byref(d)))
        return {"name": d.name.decode(errors="replace"), "compute_units": d.compute_units,
                "total_bytes": d.total_bytes, "free_bytes": d.free_bytes}

    def tuning(self) -> dict:
        t = Tuning()
        check(lib.sgpu_get_tuning(self.handle, C.byref(t)))
        return {"vcm_wavefront_paths": t.vcm_wavefront_paths, "trace_blocks_per_cu": t.trace_blocks_per_cu,
                "trace_tree": t.trace_tree, "trace_lds_nodes": t.trace_lds_nodes, "vcm_graphs": t.vcm_graphs,
                "vcm_order": t.vcm_order, "vcm_merge_lanes": t.vcm_merge_lanes, "bvh_builder": t.bvh_builder}

    def set_tuning(self, vcm_wavefront_paths=None, trace_blocks_per_cu=None, trace_tree=None, trace_lds_nodes=None,
                   vcm_graphs=None, vcm_order=None, vcm_merge_lanes=None, bvh_builder=None):
        cur = self.tuning()
        pick = lambda new, key: cur[key] if new is None else int(new)
        t = Tuning(pick(vcm_wavefront_paths, "vcm_wavefront_paths"), pick(trace_blocks_per_cu, "trace_blocks_per_cu"),
                   pick(trace_tree, "trace_tree"), pick(trace_lds_nodes, "trace_lds_nodes"), pick(vcm_graphs, "vcm_graphs"),
                   pick(vcm_order, "vcm_order"), pick(vcm_merge_lanes, "vcm_merge_lanes"),
                   pick(bvh_builder, "bvh_builder"))
        check(lib.sgpu_set_tuning(self.handle, C.byref(t)))

    def array(self, a: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(self, max(a.nbytes, 4))
        b.upload(a)
        return b

    def empty(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, max(nbytes, 4))

    def close(self):
        if self.handle:
            check(lib.sgpu_context_destroy(self.handle))
            self.handle = None


class DeviceRays:
    """SoA rays resident in device memory."""

    def __init__(self, ctx: Context, rays: dict):
        self.n = int(rays["org_x"].shape[0])
        self.bufs = {k: ctx.array(np.ascontiguousarray(rays[k], np.float32)) for k in RAY_FIELDS}
        self.soa = RaysSoA(*[self.bufs[k].ptr for k in RAY_FIELDS])

    def set_field(self, name, values: np.ndarray):
        self.bufs[name].upload(np.ascontiguousarray(values, np.float32))


class DeviceHits:
    FIELDS = (("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim_id", np.uint32),
              ("geom_id", np.uint32), ("ng_x", np.float32), ("ng_y", np.float32), ("ng_z", np.float32),
              ("inst_id0", np.uint32), ("inst_id1", np.uint32))

    def __init__(self, ctx: Context, n: int, with_ng=True, with_inst=False):
        self.n = n
        self.bufs = {}
        ptrs = []
        for name, dt in self.FIELDS:
            if (name.startswith("ng_") and not with_ng) or (name.startswith("inst_") and not with_inst):
                ptrs.append(None)
                continue
            b = ctx.empty(4 * n)
            self.bufs[name] = (b, dt)
            ptrs.append(b.ptr)
        self.soa = HitsSoA(*ptrs)

    def to_host(self):
        return {k: b.download(dt, self.n) for k, (b, dt) in self.bufs.items()}


class Scene:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _P()
        check(lib.sgpu_scene_create(ctx.handle, C.byref(h)))
        self.handle = h.value

    def add_triangles(self, positions: np.ndarray, indices: np.ndarray, transform_colmajor=None) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        gid = C.c_uint32()
        xf = None if transform_colmajor is None else fptr(np.ascontiguousarray(transform_colmajor, np.float32))
        check(lib.sgpu_scene_add_triangles(self.handle, fptr(pos), pos.size // 3, uptr(idx), idx.size // 3, xf,
                                           C.byref(gid)))
        return gid.value

    def add_quads(self, positions: np.ndarray, quads: np.ndarray, transform_colmajor=None) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        idx = np.ascontiguousarray(quads, np.uint32).reshape(-1)
        gid = C.c_uint32()
        xf = None if transform_colmajor is None else fptr(np.ascontiguousarray(transform_colmajor, np.float32))
        check(lib.sgpu_scene_add_quads(self.handle, fptr(pos), pos.size // 3, uptr(idx), idx.size // 4, xf,
                                       C.byref(gid)))
        return gid.value

    def add_curves(self, vertices: np.ndarray, indices: np.ndarray, transform_colmajor=None) -> int:
        """Round B-spline curves: control points (n x 4: x, y, z, radius) and per segment the index of
        its first control point (sgpu_scene_add_curves)."""
        vtx = np.ascontiguousarray(vertices, np.float32).reshape(-1)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        gid = C.c_uint32()
        xf = None if transform_colmajor is None else fptr(np.ascontiguousarray(transform_colmajor, np.float32))
        check(lib.sgpu_scene_add_curves(self.handle, fptr(vtx), vtx.size // 4, uptr(idx), idx.size, xf, C.byref(gid)))
        return gid.value

    def set_vertex_attributes(self, geom: int, n_verts: int, normals=None, tangents=None, uvs=None):
        """Per-vertex normals (n x 3), tangents (n x 4) and UVs (n x 2) of geometry `geom` through
        sgpu_scene_set_vertex_attributes (before commit)."""
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (normals, tangents, uvs)]
        for a, k in zip(arrs, (3, 4, 2)):
            if a is not None and a.shape != (n_verts, k):
                raise ValueError(f"set_vertex_attributes: expected {n_verts} x {k}, got {a.shape}")
        check(lib.sgpu_scene_set_vertex_attributes(self.handle, geom, *[None if a is None else a.ctypes.data
                                                                       for a in arrs], n_verts))

    def set_ptex(self, geom: int, res_log2: int, texels: np.ndarray, adjfaces: np.ndarray, adjedges: np.ndarray):
        t = np.ascontiguousarray(texels, np.uint8)
        af = np.ascontiguousarray(adjfaces, np.int32)
        ae = np.ascontiguousarray(adjedges, np.int32)
        check(lib.sgpu_scene_set_ptex(self.handle, geom, res_log2, af.shape[0], t.ctypes.data, af.ctypes.data,
                                      ae.ctypes.data))

    def set_ibl(self, rgb, rotation: float = 0.0, exposure: float = 1.0):
        """Image-based light from an equirectangular map (height x width x 3 floats, row 0 at +y)
        through sgpu_scene_set_ibl; rgb=None removes it. The scene must be committed."""
        if rgb is None:
            check(lib.sgpu_scene_set_ibl(self.handle, 0, 0, None, 0.0, 1.0))
            return
        a = np.ascontiguousarray(rgb, np.float32)
        if a.ndim != 3 or a.shape[2] != 3:
            raise ValueError("set_ibl: rgb must be height x width x 3")
        check(lib.sgpu_scene_set_ibl(self.handle, a.shape[1], a.shape[0], fptr(a), float(rotation), float(exposure)))

    def ptex_eval(self, geom: int, faces, u, v, stream=None) -> np.ndarray:
        """Filtered texels (n x 3) of geometry geom's texture at (face, u, v), through sgpu_ptex_eval."""
        faces = np.ascontiguousarray(faces, np.int32)
        u = np.ascontiguousarray(u, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        n = faces.shape[0]
        bufs = [self.ctx.empty(4 * n) for _ in range(3)] + [self.ctx.empty(12 * n)]
        try:
            for b, a in zip(bufs, (faces, u, v)):
                b.upload(a)
            check(lib.sgpu_ptex_eval(self.handle, geom, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n, bufs[3].ptr, stream))
            synchronize()
            return bufs[3].download(np.float32, 3 * n).reshape(n, 3)
        finally:
            for b in bufs:
                b.free()

    def add_instance(self, child: "Scene", local_to_world=None) -> int:
        """Instance of the committed scene `child` (sgpu_scene_add_instance); local_to_world is a
        4 x 4 matrix acting on column vectors (default: identity). Returns the instance id."""
        m = np.eye(4, dtype=np.float64) if local_to_world is None else np.asarray(local_to_world, np.float64)
        colmajor = np.ascontiguousarray(m.T.reshape(-1), np.float32)
        iid = C.c_uint32()
        check(lib.sgpu_scene_add_instance(self.handle, child.handle, fptr(colmajor), C.byref(iid)))
        return iid.value

    def add_soup(self, tri9: np.ndarray) -> int:
        tri9 = np.ascontiguousarray(tri9, np.float32).reshape(-1, 9)
        n = tri9.shape[0]
        return self.add_triangles(tri9.reshape(-1, 3), np.arange(3 * n, dtype=np.uint32))

    def commit(self, stream=None):
        check(lib.sgpu_scene_commit(self.handle, stream))

    def info(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib.sgpu_scene_info(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return {"n_tris": a.value, "n_nodes": b.value, "max_depth": c.value}

    def device_bytes(self) -> int:
        b = C.c_uint64()
        check(lib.sgpu_scene_device_bytes(self.handle, C.byref(b)))
        return b.value

    def check(self):
        check(lib.sgpu_scene_check(self.handle))

    def intersect(self, rays: DeviceRays, hits: DeviceHits, stream=None):
        check(lib.sgpu_intersect(self.handle, C.byref(rays.soa), C.byref(hits.soa), rays.n, stream))

    def intersect_noalloc(self, rays: DeviceRays, hits: DeviceHits, stream=None):
        check(lib.sgpu_intersect_noalloc(self.handle, C.byref(rays.soa), C.byref(hits.soa), rays.n, stream))

    def occluded(self, rays: DeviceRays, out: DeviceBuffer, stream=None):
        check(lib.sgpu_occluded(self.handle, C.byref(rays.soa), out.ptr, rays.n, stream))

    def intersect_stats(self, rays: DeviceRays, any_hit=False) -> dict:
        st = TraceStats()
        check(lib.sgpu_intersect_stats(self.handle, C.byref(rays.soa), None, rays.n, int(any_hit), C.byref(st)))
        return {k: getattr(st, k) for k, _ in TraceStats._fields_}

    def release(self):
        if self.handle:
            check(lib.sgpu_scene_release(self.handle))
            self.handle = None
        for s in getattr(self, "owned", []):
            s.release()
        self.owned = []


def synchronize():
    check(lib.sgpu_device_synchronize())


class Event:
    def __init__(self):
        h = _P()
        check(lib.sgpu_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        check(lib.sgpu_event_record(self.handle, stream))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        check(lib.sgpu_event_elapsed_ms(self.handle, later.handle, C.byref(ms)))
        return ms.value


# ---- integrators ------------------------------------------------------------------------------------

class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_degrees", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32)]

    @classmethod
    def make(cls, cam: dict, width: int, height: int):
        return cls((C.c_float * 3)(*cam["position"]), (C.c_float * 3)(*cam["look_at"]),
                   (C.c_float * 3)(*cam["up"]), cam["fov"], width, height)


class PtJob(C.Structure):
    _fields_ = [("camera", Camera), ("spp", C.c_uint32), ("max_path_length", C.c_uint32), ("seed", C.c_uint64),
                ("flags", C.c_uint32)]


PT_NEE_COSINE_ONCE = 1


class DptJob(C.Structure):
    _fields_ = [("camera", Camera), ("spp_x", C.c_uint32), ("spp_y", C.c_uint32), ("max_path_length", C.c_uint32),
                ("seed", C.c_uint64), ("flags", C.c_uint32)]


DPT_SORT_HITS = 1
DPT_SORT_RAYS = 2


def build_render_scene(ctx: Context, sc: dict) -> Scene:
    """A committed device scene with materials and lights from a workloads.*_scene dict, or an
    instanced scene from a dict with "models" (scene dicts of geometries, materials and textures),
    optional "subscenes" (lists of (model index, 4 x 4 local-to-world)) and "instances" (("model" |
    "subscene", index, 4 x 4 local-to-world)); the returned scene releases the scenes it places."""
    if "models" in sc:
        models = [build_render_scene(ctx, dict(m, lights=np.zeros(0, sc["lights"].dtype),
                                               background=np.zeros(3, np.float32)))
                  for m in sc["models"]]
        subs = []
        for desc in sc.get("subscenes", []):
            sub = Scene(ctx)
            for midx, m in desc:
                sub.add_instance(models[midx], m)
            sub.commit()
            subs.append(sub)
        top = Scene(ctx)
        for kind, idx, m in sc["instances"]:
            top.add_instance(models[idx] if kind == "model" else subs[idx], m)
        top.commit()
        top.owned = subs[::-1] + models[::-1]
        lights = np.ascontiguousarray(sc["lights"])
        bg = np.ascontiguousarray(sc["background"], np.float32)
        check(lib.sgpu_scene_set_lights(top.handle, lights.ctypes.data if lights.size else None, lights.shape[0],
                                        fptr(bg)))
        ibl = sc.get("ibl")
        if ibl is not None:
            top.set_ibl(ibl["rgb"], ibl.get("rotation", 0.0), ibl.get("exposure", 1.0))
        return top
    scene = Scene(ctx)
    for g in sc["geoms"]:
        if isinstance(g, dict):
            if "curves" in g:
                scene.add_curves(g["curve_vertices"], g["curves"])
                continue
            if "triangles" in g:
                gid = scene.add_triangles(g["positions"], g["triangles"])
            else:
                gid = scene.add_quads(g["positions"], g["quads"])
            if g.get("ptex") is not None:
                p = g["ptex"]
                scene.set_ptex(gid, p["res_log2"], p["texels"], p["adjfaces"], p["adjedges"])
            if g.get("normals") is not None:
                scene.set_vertex_attributes(gid, np.asarray(g["positions"]).reshape(-1, 3).shape[0], g["normals"],
                                            g.get("tangents"), g.get("uvs"))
        else:
            scene.add_soup(g)
    scene.commit()
    mats = np.ascontiguousarray(sc["materials"], np.float32)
    gm = np.ascontiguousarray(sc["geom_material"], np.uint32)
    check(lib.sgpu_scene_set_materials(scene.handle, fptr(mats), mats.shape[0], uptr(gm), gm.shape[0]))
    lights = np.ascontiguousarray(sc["lights"])
    bg = np.ascontiguousarray(sc["background"], np.float32)
    check(lib.sgpu_scene_set_lights(scene.handle, lights.ctypes.data if lights.size else None, lights.shape[0],
                                    fptr(bg)))
    ibl = sc.get("ibl")
    if ibl is not None:
        scene.set_ibl(ibl["rgb"], ibl.get("rotation", 0.0), ibl.get("exposure", 1.0))
    return scene


VCM_KERNEL_TIMES = 1
VCM_COUNT_TRAVERSAL = 2
VCM_ONE_SLICE_PER_WAVEFRONT = 4
VCM_SHORTEST_FIRST = 8  # servers only
VCM_COUNT_MERGE = 16

# sgpu_tuning.vcm_order bits
ORDER_MERGE_CELLS = 1
ORDER_SHADE_MATERIAL = 2


class VcmJob(C.Structure):
    _fields_ = [("camera", Camera), ("iterations", C.c_uint32), ("iteration_base", C.c_uint32),
                ("max_path_length", C.c_uint32), ("base_radius", C.c_float), ("radius_alpha", C.c_float),
                ("flags", C.c_uint32), ("seed", C.c_uint64)]


class VcmStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("rays", "closest_rays", "shadow_rays", "closest_launches",
                                          "light_vertices", "samples", "wavefronts", "nodes_visited",
                                          "tris_tested")] + \
               [(k, C.c_float) for k in ("closest_ms", "shadow_ms", "light_ms", "camera_ms", "grid_ms",
                                         "merge_ms")] + \
               [(k, C.c_uint64) for k in ("merge_launches", "merge_entries", "merge_queries", "merge_visited",
                                          "merge_near")]


def render_vcm(scene: Scene, cam: dict, width, height, iterations, max_path_length, base_radius, radius_alpha, seed,
               iteration_base=0, stream=None, return_stats=False, flags=0, framebuffer=None):
    """VCM image (height x width x 3, radiance averaged over iterations) through sgpu_render_vcm.
    With `framebuffer` (a device buffer of 12 x width x height bytes) the image stays on the device
    and None is returned in its place."""
    ctx = scene.ctx
    fb = framebuffer if framebuffer is not None else ctx.empty(3 * width * height * 4)
    job = VcmJob(Camera.make(cam, width, height), iterations, iteration_base, max_path_length, base_radius,
                 radius_alpha, flags, seed)
    st = VcmStats()
    try:
        check(lib.sgpu_render_vcm(scene.handle, C.byref(job), fb.ptr, C.byref(st), stream))
        out = None if framebuffer is not None else fb.download(np.float32, 3 * width * height).reshape(height, width, 3)
    finally:
        if framebuffer is None:
            fb.free()
    if return_stats:
        return out, {k: getattr(st, k) for k, _ in VcmStats._fields_}
    return out


class VcmServerConfig(C.Structure):
    _fields_ = [("wavefront_paths", C.c_uint64), ("max_path_length", C.c_uint32), ("flags", C.c_uint32)]


def vcm_job_struct(j: dict) -> VcmJob:
    return VcmJob(Camera.make(j["camera"], j["width"], j["height"]), j["iterations"], j.get("iteration_base", 0),
                  j["max_path_length"], j["base_radius"], j["radius_alpha"], 0, j["seed"])


class VcmServer:
    """A persistent VCM executor (sgpu_vcm_server_*): submit jobs at any time, advance one wavefront
    per step; a job's framebuffer (DeviceBuffer, H x W x 3 floats) is final once its id is retired."""

    def __init__(self, ctx: Context, wavefront_paths=1 << 20, max_path_length=10, flags=0, stream=None):
        self.ctx = ctx
        self.stream = stream
        self.cfg = VcmServerConfig(wavefront_paths, max_path_length, flags)
        h = _P()
        check(lib.sgpu_vcm_server_create(ctx.handle, C.byref(self.cfg), stream, C.byref(h)))
        self.handle = h.value
        self._done = (C.c_uint64 * 4096)()

    @staticmethod
    def footprint(ctx: Context, wavefront_paths=1 << 20, max_path_length=10, flags=0) -> int:
        cfg = VcmServerConfig(wavefront_paths, max_path_length, flags)
        b = C.c_uint64()
        check(lib.sgpu_vcm_server_footprint(ctx.handle, C.byref(cfg), C.byref(b)))
        return b.value

    def submit(self, scene: Scene, job: dict, framebuffer: DeviceBuffer) -> int:
        js = vcm_job_struct(job)
        jid = C.c_uint64()
        check(lib.sgpu_vcm_server_submit(self.handle, scene.handle, C.byref(js), framebuffer.ptr, C.byref(jid),
                                         self.stream))
        return jid.value

    def step(self):
        """Runs one wavefront; returns (ids retired, jobs still admitted)."""
        nd, na = C.c_uint32(), C.c_uint32()
        check(lib.sgpu_vcm_server_step(self.handle, self.stream, self._done, 4096, C.byref(nd), C.byref(na)))
        return [self._done[k] for k in range(min(nd.value, 4096))], na.value

    def stats(self) -> dict:
        st = VcmStats()
        check(lib.sgpu_vcm_server_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in VcmStats._fields_}

    def close(self):
        if self.handle:
            check(lib.sgpu_vcm_server_destroy(self.handle, self.stream))
            self.handle = None


def render_vcm_batch(jobs, flags=0, stream=None, return_stats=False, keep_on_device=False, framebuffers=None):
    """Renders a list of jobs with sgpu_render_vcm_batch. Each job is a dict: scene (Scene), camera
    (dict), width, height, iterations, max_path_length, base_radius, radius_alpha, seed and optionally
    iteration_base. Returns the images (height x width x 3 each) or, with keep_on_device, the device
    framebuffers; with `framebuffers` (DeviceBuffers, one per job) the images stay in those and
    nothing is returned but the stats."""
    n = len(jobs)
    ctx = jobs[0]["scene"].ctx
    arr_j = (VcmJob * n)()
    scenes = (_P * n)()
    if framebuffers is not None:
        assert len(framebuffers) == n
        keep_on_device = True
    fbs = framebuffers if framebuffers is not None else [ctx.empty(3 * j["width"] * j["height"] * 4) for j in jobs]
    fb_ptrs = (_P * n)(*[b.ptr for b in fbs])
    for k, j in enumerate(jobs):
        scenes[k] = j["scene"].handle
        arr_j[k] = VcmJob(Camera.make(j["camera"], j["width"], j["height"]), j["iterations"],
                          j.get("iteration_base", 0), j["max_path_length"], j["base_radius"], j["radius_alpha"], 0,
                          j["seed"])
    st = VcmStats()
    try:
        check(lib.sgpu_render_vcm_batch(scenes, arr_j, fb_ptrs, n, flags, C.byref(st), stream))
        if keep_on_device:
            out = None if framebuffers is not None else fbs
        else:
            out = [b.download(np.float32, 3 * j["width"] * j["height"]).reshape(j["height"], j["width"], 3)
                   for b, j in zip(fbs, jobs)]
    finally:
        if not keep_on_device:
            for b in fbs:
                b.free()
    if return_stats:
        return out, {k: getattr(st, k) for k, _ in VcmStats._fields_}
    return out


def save_hdr(ctx: Context, path_prefix: str, framebuffer, width: int, height: int, layers: int = 1, stream=None):
    """sgpu_framebuffer_save_hdr: writes <path_prefix>_<layer>.hdr for each layer of a device
    framebuffer (a DeviceBuffer) or of a host array (layers x height x width x 3, uploaded first)."""
    buf, owned = framebuffer, None
    if isinstance(framebuffer, np.ndarray):
        a = np.ascontiguousarray(framebuffer, np.float32)
        if a.size != layers * height * width * 3:
            raise ValueError("save_hdr: the array is not layers x height x width x 3")
        owned = ctx.empty(a.nbytes)
        owned.upload(a)
        buf = owned
    try:
        check(lib.sgpu_framebuffer_save_hdr(path_prefix.encode(), buf.ptr, width, height, layers, stream))
    finally:
        if owned is not None:
            owned.free()


def render_deferred_pt(scene: Scene, cam: dict, width, height, spp_x, spp_y, max_path_length, seed,
                       flags=0, framebuffer=None, stream=None):
    """sgpu_render_deferred_pt: height x width x 3 (or into `framebuffer`, a DeviceBuffer, and None)."""
    ctx = scene.ctx
    fb = framebuffer if framebuffer is not None else ctx.empty(3 * width * height * 4)
    job = DptJob(Camera.make(cam, width, height), spp_x, spp_y, max_path_length, seed, flags)
    try:
        check(lib.sgpu_render_deferred_pt(scene.handle, C.byref(job), fb.ptr, stream))
        if framebuffer is not None:
            return None
        return fb.download(np.float32, 3 * width * height).reshape(height, width, 3)
    finally:
        if framebuffer is None:
            fb.free()


class PtStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("closest_rays", "shadow_rays", "hits", "bounces", "nodes_visited",
                                          "tris_tested")] + \
               [(k, C.c_float) for k in ("closest_ms", "shadow_ms", "shade_ms", "other_ms")]


PT_COUNT_TRAVERSAL = 2


def render_pt(scene: Scene, cam: dict, width, height, spp, max_path_length, seed, flags=0, stream=None,
              return_stats=False, framebuffer=None):
    """Path-traced layers (2 x height x width x 3) through sgpu_render_pt, or with return_stats through
    sgpu_render_pt_stats, returning (layers, stats dict). With `framebuffer` (a device buffer of
    24 x width x height bytes) the layers stay on the device and None is returned in their place."""
    ctx = scene.ctx
    fb = framebuffer if framebuffer is not None else ctx.empty(2 * 3 * width * height * 4)
    job = PtJob(Camera.make(cam, width, height), spp, max_path_length, seed, flags)
    st = PtStats()
    try:
        if return_stats:
            check(lib.sgpu_render_pt_stats(scene.handle, C.byref(job), fb.ptr, C.byref(st), stream))
        else:
            check(lib.sgpu_render_pt(scene.handle, C.byref(job), fb.ptr, stream))
        out = None if framebuffer is not None else \
            fb.download(np.float32, 2 * 3 * width * height).reshape(2, height, width, 3)
    finally:
        if framebuffer is None:
            fb.free()
    if return_stats:
        return out, {k: getattr(st, k) for k, _ in PtStats._fields_}
    return out
