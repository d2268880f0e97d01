"""ORACLE -- test infrastructure only.

ctypes binding of oracle/liboracle.so, the CPU restatement of Selas's hot path used as the
checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product
(selas_amd / libselas_gpu) never imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tnear", "tfar")]


class _Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("t", "prim", "u", "v")]


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    U32 = C.c_uint32
    lib.orc_closest_brute_f32.argtypes = [P, U32, C.POINTER(_Rays), U32, C.POINTER(_Hits), C.c_int]
    lib.orc_occluded_brute_f32.argtypes = [P, U32, C.POINTER(_Rays), U32, P, C.c_int]
    lib.orc_closest_brute_f64.argtypes = [P, U32, C.POINTER(_Rays), U32, P, P, P, P, C.c_int]
    lib.orc_bvh_build.argtypes = [P, U32]
    lib.orc_bvh_build.restype = P
    lib.orc_bvh_free.argtypes = [P]
    lib.orc_bvh_closest.argtypes = [P, C.POINTER(_Rays), U32, C.POINTER(_Hits), C.c_int]
    lib.orc_bvh_occluded.argtypes = [P, C.POINTER(_Rays), U32, P, C.c_int]
    lib.orc_bvh_node_count.argtypes = [P]
    lib.orc_bvh_node_count.restype = U32
    lib.orc_tri_test.argtypes = [P, P, P, P, P, C.c_float, P, P, P]
    lib.orc_tri_test.restype = C.c_int
    return lib


lib = _load()

RAY_KEYS = ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")


def _rays(r: dict):
    arrs = [np.ascontiguousarray(r[k], np.float32) for k in RAY_KEYS]
    return _Rays(*[a.ctypes.data for a in arrs]), arrs, int(arrs[0].shape[0])


def _threads(threads):
    return threads if threads else (os.cpu_count() or 1)


def closest_brute_f32(tri9, rays, threads=0):
    tri = np.ascontiguousarray(tri9, np.float32)
    rs, keep, n = _rays(rays)
    out = {"t": np.empty(n, np.float32), "prim": np.empty(n, np.uint32), "u": np.empty(n, np.float32),
           "v": np.empty(n, np.float32)}
    h = _Hits(*[out[k].ctypes.data for k in ("t", "prim", "u", "v")])
    lib.orc_closest_brute_f32(tri.ctypes.data, tri.shape[0], C.byref(rs), n, C.byref(h), _threads(threads))
    return out


def occluded_brute_f32(tri9, rays, threads=0):
    tri = np.ascontiguousarray(tri9, np.float32)
    rs, keep, n = _rays(rays)
    occ = np.empty(n, np.uint32)
    lib.orc_occluded_brute_f32(tri.ctypes.data, tri.shape[0], C.byref(rs), n, occ.ctypes.data, _threads(threads))
    return occ


def closest_brute_f64(tri9, rays, threads=0):
    tri = np.ascontiguousarray(tri9, np.float32)
    rs, keep, n = _rays(rays)
    prim = np.empty(n, np.uint32)
    t1, t2, margin = (np.empty(n, np.float64) for _ in range(3))
    lib.orc_closest_brute_f64(tri.ctypes.data, tri.shape[0], C.byref(rs), n, prim.ctypes.data, t1.ctypes.data,
                              t2.ctypes.data, margin.ctypes.data, _threads(threads))
    return {"prim": prim, "t": t1, "t2": t2, "margin": margin}


class CpuBvh:
    """The oracle's own SAH tree over a triangle soup (shares nothing with the device LBVH)."""

    def __init__(self, tri9):
        self._tri = np.ascontiguousarray(tri9, np.float32)
        self.handle = lib.orc_bvh_build(self._tri.ctypes.data, self._tri.shape[0])

    def closest(self, rays, threads=0):
        rs, keep, n = _rays(rays)
        out = {"t": np.empty(n, np.float32), "prim": np.empty(n, np.uint32), "u": np.empty(n, np.float32),
               "v": np.empty(n, np.float32)}
        h = _Hits(*[out[k].ctypes.data for k in ("t", "prim", "u", "v")])
        lib.orc_bvh_closest(self.handle, C.byref(rs), n, C.byref(h), _threads(threads))
        return out

    def occluded(self, rays, threads=0):
        rs, keep, n = _rays(rays)
        occ = np.empty(n, np.uint32)
        lib.orc_bvh_occluded(self.handle, C.byref(rs), n, occ.ctypes.data, _threads(threads))
        return occ

    def __del__(self):
        if getattr(self, "handle", None):
            lib.orc_bvh_free(self.handle)
            self.handle = None


def tri_test(v0, e1, e2, o, d, tnear):
    a = [np.ascontiguousarray(x, np.float32) for x in (v0, e1, e2, o, d)]
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = lib.orc_tri_test(*[x.ctypes.data for x in a], C.c_float(tnear), C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


# ---- shading (orc_shade.c) -------------------------------------------------------------------------

def _f(a):
    return np.ascontiguousarray(a, np.float32)


lib.orc_bsdf_eval_batch.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_void_p]
lib.orc_bsdf_sample_batch.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_uint32, C.c_void_p]
lib.orc_ggx.argtypes = [C.c_void_p] * 3 + [C.c_float] * 4 + [C.c_void_p]
lib.orc_fresnel_dielectric.argtypes = [C.c_float] * 3
lib.orc_fresnel_dielectric.restype = C.c_float
lib.orc_schlick_weight.argtypes = [C.c_float]
lib.orc_schlick_weight.restype = C.c_float
lib.orc_schlick_r0.argtypes = [C.c_float]
lib.orc_schlick_r0.restype = C.c_float
lib.orc_schlick.argtypes = [C.c_float] * 2
lib.orc_schlick.restype = C.c_float
lib.orc_cmj.argtypes = [C.c_int32] * 4 + [C.c_void_p]
lib.orc_rect_light.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
lib.orc_ray_quad.argtypes = [C.c_void_p]
lib.orc_ray_quad.restype = C.c_int
lib.orc_concentric_disc.argtypes = [C.c_float, C.c_float, C.c_void_p]
lib.orc_make_frame.argtypes = [C.c_void_p] * 3


class _Camera(C.Structure):
    _fields_ = [(k, C.c_float * 3) for k in ("x", "y", "z", "pos", "forward")] + [
        ("width", C.c_uint32), ("height", C.c_uint32), ("vipd", C.c_float)]


lib.orc_camera_init.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(_Camera)]
lib.orc_camera_cmj_dir.argtypes = [C.POINTER(_Camera)] + [C.c_int32] * 6 + [C.c_void_p]


lib.orc_set_libm.argtypes = [C.c_int]


def set_libm_single(single: bool):
    """Evaluate cos/sin/acos/pow with the single-precision libm, as the reference's code does
    (True), or in double rounded to float, as the device code does (False, the default)."""
    lib.orc_set_libm(1 if single else 0)


def bsdf_eval(blocks, n, v, l):
    blocks, n, v, l = _f(blocks), _f(n), _f(v), _f(l)
    out = np.zeros((blocks.shape[0], 5), np.float32)
    lib.orc_bsdf_eval_batch(blocks.ctypes.data, n.ctypes.data, v.ctypes.data, l.ctypes.data, blocks.shape[0],
                            out.ctypes.data)
    return out


def bsdf_sample(blocks, n, v, uniforms):
    blocks, n, v, uniforms = _f(blocks), _f(n), _f(v), _f(uniforms)
    out = np.zeros((blocks.shape[0], 14), np.float32)
    lib.orc_bsdf_sample_batch(blocks.ctypes.data, n.ctypes.data, v.ctypes.data, uniforms.ctypes.data,
                              uniforms.shape[1], blocks.shape[0], out.ctypes.data)
    return out


def ggx(wo, wi, wm, ax, ay, u1, u2):
    out = np.zeros(8, np.float32)
    lib.orc_ggx(_f(wo).ctypes.data, _f(wi).ctypes.data, _f(wm).ctypes.data, ax, ay, u1, u2, out.ctypes.data)
    return out


def cmj(s, m, n, p):
    out = np.zeros(2, np.float32)
    lib.orc_cmj(int(s), int(m), int(n), int(p), out.ctypes.data)
    return out


def camera_rays(cam10, width, height, px):
    c = _Camera()
    cam = _f(cam10)
    lib.orc_camera_init(cam.ctypes.data, width, height, C.byref(c))
    out = np.zeros((px.shape[0], 3), np.float32)
    for i, q in enumerate(px):
        lib.orc_camera_cmj_dir(C.byref(c), *[int(x) for x in q], out[i].ctypes.data)
    basis = np.array(list(c.x) + list(c.y) + list(c.z) + [c.vipd], np.float32)
    return out, basis


def rect_light(q12, u, v):
    out = np.zeros(5, np.float32)
    lib.orc_rect_light(_f(q12).ctypes.data, u, v, out.ctypes.data)
    return out


def ray_quad(q18):
    return lib.orc_ray_quad(_f(q18).ctypes.data)


def concentric_disc(r0, r1):
    out = np.zeros(2, np.float32)
    lib.orc_concentric_disc(r0, r1, out.ctypes.data)
    return out


# ---- integrators (orc_pt.c) --------------------------------------------------------------------------

class _Ptex(C.Structure):
    _fields_ = [("res_log2", C.c_int), ("nfaces", C.c_uint32), ("texels", C.c_void_p), ("adjfaces", C.c_void_p),
                ("adjedges", C.c_void_p), ("tri", C.c_int)]


class _Scene(C.Structure):
    _fields_ = [("bvh", C.c_void_p), ("tri9", C.c_void_p), ("tri_geom", C.c_void_p), ("materials", C.c_void_p),
                ("geom_material", C.c_void_p), ("n_geoms", C.c_uint32), ("lights", C.c_void_p),
                ("n_lights", C.c_uint32), ("background", C.c_float * 3), ("geom_first", C.c_void_p),
                ("geom_tex", C.c_void_p), ("ibl", C.c_void_p), ("inst", C.c_void_p), ("geom_attr", C.c_void_p),
                ("curve_cp", C.c_void_p), ("curve_box", C.c_void_p), ("curve_geom", C.c_void_p),
                ("curve_prim", C.c_void_p), ("n_curves", C.c_uint32)]


def curve_segments(g):
    """Segments (n, 16) of a curve geometry entry {"curve_vertices": (m, 4), "curves": first control
    point of each segment}: the four control points (x, y, z, radius) of each."""
    v = np.asarray(g["curve_vertices"], np.float32).reshape(-1, 4)
    idx = np.asarray(g["curves"], np.int64).reshape(-1)
    return np.ascontiguousarray(np.stack([v[idx + j] for j in range(4)], 1).reshape(-1, 16), np.float32)


def _curve_arrays(geoms):
    """(cp, boxes, geom, prim) over the curve geometries of a geometry list (None if there are none)."""
    cps, gs, ps = [], [], []
    for i, g in enumerate(geoms):
        if isinstance(g, dict) and "curves" in g:
            cp = curve_segments(g)
            cps.append(cp)
            gs.append(np.full(cp.shape[0], i, np.uint32))
            ps.append(np.arange(cp.shape[0], dtype=np.uint32))
    if not cps:
        return None
    cp = np.ascontiguousarray(np.concatenate(cps), np.float32)
    box = np.zeros((cp.shape[0], 6), np.float32)
    lib.orc_curve_boxes(cp.ctypes.data, cp.shape[0], box.ctypes.data)
    return cp, box, np.ascontiguousarray(np.concatenate(gs)), np.ascontiguousarray(np.concatenate(ps))


class _Attr(C.Structure):
    _fields_ = [("n", C.c_void_p), ("t", C.c_void_p), ("uv", C.c_void_p), ("corners", C.c_uint32)]


lib.orc_interp_attr.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_void_p]


def geom_corner_attrs(g):
    """Corner arrays (primitive order, the primitive's own vertex order: 3 per triangle, 4 per quad)
    of a geometry dict's per-vertex "normals", "tangents" and "uvs" (None where absent), and the
    corner count; (None, None, None, 0) for a geometry without attributes."""
    if not isinstance(g, dict) or g.get("normals") is None:
        return None, None, None, 0
    idx = np.asarray(g["triangles"] if "triangles" in g else g["quads"], np.int64)
    out = []
    for key, k in (("normals", 3), ("tangents", 4), ("uvs", 2)):
        a = g.get(key)
        out.append(None if a is None else np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, k)[idx]
                                                               .reshape(-1), np.float32))
    return out[0], out[1], out[2], idx.shape[1]


def interp_attr(corner_values, K, corners, prim, u, v):
    """The oracle's rtcInterpolate0 of a K-float corner array at primitive prim, (u, v)."""
    a = _f(corner_values)
    out = np.zeros(K, np.float32)
    lib.orc_interp_attr(a.ctypes.data, K, corners, prim, float(u), float(v), out.ctypes.data)
    return out


class _AttrTable:
    """Per-geometry attribute records (a list of geometry dicts or arrays, in the scene's list
    order) with the corner arrays they point to; `ptr` is NULL when no geometry has attributes."""

    def __init__(self, geoms):
        self.keep = []
        self.table = (_Attr * max(len(geoms), 1))()
        any_attr = False
        for i, g in enumerate(geoms):
            n, t, uv, corners = geom_corner_attrs(g)
            if n is None:
                continue
            any_attr = True
            self.keep += [a for a in (n, t, uv) if a is not None]
            self.table[i] = _Attr(n.ctypes.data, None if t is None else t.ctypes.data,
                                  None if uv is None else uv.ctypes.data, corners)
        self.ptr = C.cast(self.table, C.c_void_p) if any_attr else None


# ---- image-based light (orc_ibl.c) ---------------------------------------------------------------

class _Ibl(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("rgb", C.c_void_p), ("mdf", C.c_void_p), ("cdf", C.c_void_p),
                ("rotation", C.c_float), ("exposure", C.c_float), ("center", C.c_float * 3), ("radius", C.c_float)]


lib.orc_ibl_tables.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
lib.orc_ibl_rotation.argtypes = [C.c_float]
lib.orc_ibl_rotation.restype = C.c_float
lib.orc_bounding_sphere.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
lib.orc_ibl_sample_distribution.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]
lib.orc_ibl_sample_distribution.restype = C.c_uint32
lib.orc_ibl_sample.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
lib.orc_ibl_lookup.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
lib.orc_ibl_lookup.restype = C.c_float
lib.orc_ibl_emit.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p]
lib.orc_ibl_direct.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
lib.orc_ibl_position_pdf.argtypes = [C.c_void_p]
lib.orc_ibl_position_pdf.restype = C.c_float


def ibl_tables(rgb):
    """(marginal CDF (h), conditional CDFs (h x w)) of an h x w x 3 map (BuildImageBasedLight.cpp)."""
    a = _f(rgb)
    h, w = a.shape[0], a.shape[1]
    mdf = np.zeros(h, np.float32)
    cdf = np.zeros((h, w), np.float32)
    lib.orc_ibl_tables(w, h, a.ctypes.data, mdf.ctypes.data, cdf.ctypes.data)
    return mdf, cdf


def bounding_sphere(tri9):
    t = _f(tri9)
    out = np.zeros(4, np.float32)
    lib.orc_bounding_sphere(t.ctypes.data, t.size // 9, out.ctypes.data)
    return out


class Ibl:
    """The oracle's image-based light: tables, rotation and the emission sphere of a triangle set."""

    def __init__(self, rgb, rotation=0.0, exposure=1.0, tri9=None, sphere=None):
        self.rgb = _f(rgb)
        self.mdf, self.cdf = ibl_tables(self.rgb)
        sph = bounding_sphere(tri9) if tri9 is not None else np.asarray(sphere, np.float32)
        self.s = _Ibl(self.rgb.shape[1], self.rgb.shape[0], self.rgb.ctypes.data, self.mdf.ctypes.data,
                      self.cdf.ctypes.data, lib.orc_ibl_rotation(rotation), exposure,
                      (C.c_float * 3)(*[float(x) for x in sph[:3]]), float(sph[3]))

    def sample(self, r0, r1):
        """Ibl: (theta, phi, x, y, pdf)"""
        out = np.zeros(6, np.float32)
        lib.orc_ibl_sample(C.byref(self.s), r0, r1, out.ctypes.data)
        return out[:5]

    def lookup(self, wi):
        """SampleIbl(wi): (radiance, pdf)"""
        w = _f(wi)
        rgb = np.zeros(3, np.float32)
        pdf = lib.orc_ibl_lookup(C.byref(self.s), w.ctypes.data, rgb.ctypes.data)
        return rgb, np.float32(pdf)

    def emit(self, r0, r1, r2, r3):
        """EmitIblLightSample: (position, direction, radiance, emission pdf, direction pdf)"""
        out = np.zeros(14, np.float32)
        lib.orc_ibl_emit(C.byref(self.s), r0, r1, r2, r3, out.ctypes.data)
        return out[0:3], out[3:6], out[6:9], out[9], out[10]

    def direct(self, r0, r1):
        """DirectIblLightSample: (direction, radiance, pdf)"""
        out = np.zeros(8, np.float32)
        lib.orc_ibl_direct(C.byref(self.s), r0, r1, out.ctypes.data)
        return out[0:3], out[3:6], out[6]

    def position_pdf(self):
        return lib.orc_ibl_position_pdf(C.byref(self.s))


class _PtJob(C.Structure):
    _fields_ = [("cam10", C.c_float * 10), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_path_length", C.c_uint32), ("seed", C.c_uint64), ("flags", C.c_uint32)]


lib.orc_render_pt.argtypes = [C.POINTER(_Scene), C.POINTER(_PtJob), C.c_void_p, C.c_int]


class RenderScene:
    """The oracle's view of a workloads.*_scene dict: flattened triangles (quad geometries split as
    Embree splits them), the oracle SAH tree, and the per-face textures of quad geometries."""

    def __init__(self, sc: dict):
        from selas_amd.workloads import geom_tri9

        tris = [geom_tri9(g) for g in sc["geoms"]]
        self.tri9 = np.ascontiguousarray(np.concatenate(tris), np.float32)
        self.tri_geom = np.concatenate([np.full(t.shape[0], i, np.uint32) for i, t in enumerate(tris)])
        self.geom_first = np.cumsum([0] + [t.shape[0] for t in tris[:-1]]).astype(np.uint32)
        self.materials = np.ascontiguousarray(sc["materials"], np.float32)
        self.geom_material = np.ascontiguousarray(sc["geom_material"], np.uint32)
        self.lights = np.ascontiguousarray(sc["lights"])
        self.textures = []
        self.geom_tex = (_Ptex * len(tris))()
        for i, g in enumerate(sc["geoms"]):
            if isinstance(g, dict) and g.get("ptex") is not None:
                p = g["ptex"]
                t = PtexTexture(p["texels"], p["res_log2"], p["adjfaces"], p["adjedges"], tri="triangles" in g)
                self.textures.append(t)
                self.geom_tex[i] = t.s
        self.bvh = CpuBvh(self.tri9)
        self.attrs = _AttrTable(sc["geoms"])
        ibl = sc.get("ibl")
        self.curves = _curve_arrays(sc["geoms"])
        box9 = self.tri9
        if self.curves is not None:
            # the library's emission sphere bounds the curves' boxes too: two degenerate triangles at
            # the corners of their union
            cb = self.curves[1]
            lo, hi = cb[:, :3].min(0), cb[:, 3:].max(0)
            box9 = np.concatenate([self.tri9, np.tile(lo, 3)[None], np.tile(hi, 3)[None]]).astype(np.float32)
        self.ibl = None if ibl is None else Ibl(ibl["rgb"], ibl.get("rotation", 0.0), ibl.get("exposure", 1.0),
                                                tri9=box9)
        self.s = _Scene(self.bvh.handle, self.tri9.ctypes.data, self.tri_geom.ctypes.data, self.materials.ctypes.data,
                        self.geom_material.ctypes.data, self.geom_material.shape[0],
                        self.lights.ctypes.data if self.lights.size else None, self.lights.shape[0],
                        (C.c_float * 3)(*[float(x) for x in sc["background"]]), self.geom_first.ctypes.data,
                        C.cast(self.geom_tex, C.c_void_p),
                        C.cast(C.pointer(self.ibl.s), C.c_void_p) if self.ibl is not None else None, None,
                        self.attrs.ptr)
        if self.curves is not None:
            cp, box, cg, cpr = self.curves
            self.s.curve_cp, self.s.curve_box = cp.ctypes.data, box.ctypes.data
            self.s.curve_geom, self.s.curve_prim, self.s.n_curves = cg.ctypes.data, cpr.ctypes.data, cp.shape[0]


def render_pt(rs: RenderScene, cam: dict, width, height, spp, max_path_length, seed, flags=0, threads=0):
    cam10 = list(cam["position"]) + list(cam["look_at"]) + list(cam["up"]) + [cam["fov"]]
    job = _PtJob((C.c_float * 10)(*cam10), width, height, spp, max_path_length, seed, flags)
    fb = np.zeros((2, height, width, 3), np.float32)
    lib.orc_render_pt(C.byref(rs.s), C.byref(job), fb.ctypes.data, _threads(threads))
    return fb


class _DptJob(C.Structure):
    _fields_ = [("cam10", C.c_float * 10), ("width", C.c_uint32), ("height", C.c_uint32), ("spp_x", C.c_uint32),
                ("spp_y", C.c_uint32), ("max_path_length", C.c_uint32), ("seed", C.c_uint64)]


lib.orc_render_dpt.argtypes = [C.POINTER(_Scene), C.POINTER(_DptJob), C.c_void_p, C.c_int]


def render_dpt(rs: RenderScene, cam: dict, width, height, spp_x, spp_y, max_path_length, seed, threads=0):
    """The deferred path tracer's estimator (orc_pt.c, orc_render_dpt): height x width x 3."""
    cam10 = list(cam["position"]) + list(cam["look_at"]) + list(cam["up"]) + [cam["fov"]]
    job = _DptJob((C.c_float * 10)(*cam10), width, height, spp_x, spp_y, max_path_length, seed)
    fb = np.zeros((height, width, 3), np.float32)
    lib.orc_render_dpt(C.byref(rs.s), C.byref(job), fb.ctypes.data, _threads(threads))
    return fb


class _VcmJob(C.Structure):
    _fields_ = [("cam10", C.c_float * 10), ("width", C.c_uint32), ("height", C.c_uint32),
                ("iterations", C.c_uint32), ("iteration_base", C.c_uint32), ("max_path_length", C.c_uint32),
                ("base_radius", C.c_float), ("radius_alpha", C.c_float), ("seed", C.c_uint64)]


class _HashGrid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("hi", C.c_float * 3), ("radius", C.c_float), ("radius_sq", C.c_float),
                ("inv_cell", C.c_float), ("cell_count", C.c_uint32), ("count", C.c_uint32),
                ("cell_ends", C.c_void_p), ("indices", C.c_void_p)]


lib.orc_render_vcm.argtypes = [C.POINTER(_Scene), C.POINTER(_VcmJob), C.c_void_p, C.c_void_p, C.c_int]
lib.orc_vcm_constants.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_void_p]
lib.orc_hash_grid_build.argtypes = [C.POINTER(_HashGrid), C.c_void_p, C.c_uint32, C.c_uint32, C.c_float]
lib.orc_hash_grid_free.argtypes = [C.POINTER(_HashGrid)]
lib.orc_hash_grid_query.argtypes = [C.POINTER(_HashGrid), C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
lib.orc_hash_grid_query.restype = C.c_uint32


def render_vcm(rs: RenderScene, cam: dict, width, height, iterations, max_path_length, base_radius,
               radius_alpha=0.75, seed=1, iteration_base=0, threads=0):
    cam10 = list(cam["position"]) + list(cam["look_at"]) + list(cam["up"]) + [cam["fov"]]
    job = _VcmJob((C.c_float * 10)(*cam10), width, height, iterations, iteration_base, max_path_length,
                  base_radius, radius_alpha, seed)
    fb = np.zeros((height, width, 3), np.float32)
    stats = np.zeros(2, np.float32)
    lib.orc_render_vcm(C.byref(rs.s), C.byref(job), fb.ctypes.data, stats.ctypes.data, _threads(threads))
    return fb, {"light_vertices_per_iteration": float(stats[0]), "rays": float(stats[1])}


def vcm_constants(vm_count, vc_count, base_radius, alpha, iteration):
    out = np.zeros(6, np.float32)
    lib.orc_vcm_constants(vm_count, vc_count, base_radius, alpha, iteration, out.ctypes.data)
    return {"radius": out[0], "radius_sq": out[1], "vm_normalization": out[2], "vm_weight": out[3],
            "vc_weight": out[4]}


def hash_grid(points, cell_count, radius, queries, max_found=512):
    """BuildHashGrid + SearchHashGrid restated: cell ends, point order, box and per-query finds."""
    pts = _f(points)
    g = _HashGrid()
    lib.orc_hash_grid_build(C.byref(g), pts.ctypes.data, pts.shape[0], cell_count, radius)
    ends = np.ctypeslib.as_array(C.cast(g.cell_ends, C.POINTER(C.c_uint32)), (cell_count,)).copy()
    order = np.ctypeslib.as_array(C.cast(g.indices, C.POINTER(C.c_uint32)), (max(1, pts.shape[0]),)).copy()
    box = np.array(list(g.lo) + list(g.hi), np.float32)
    q = _f(queries)
    found = np.zeros((q.shape[0], max_found), np.uint32)
    counts = np.zeros(q.shape[0], np.uint32)
    for i in range(q.shape[0]):
        counts[i] = lib.orc_hash_grid_query(C.byref(g), pts.ctypes.data, q[i].ctypes.data, found[i].ctypes.data,
                                            max_found)
    lib.orc_hash_grid_free(C.byref(g))
    return {"cell_ends": ends, "cell_indices": order[:pts.shape[0]], "box": box, "found": found,
            "found_count": counts}


class _Uniforms(C.Structure):
    _fields_ = [("values", C.c_void_p), ("count", C.c_uint32), ("pos", C.c_uint32), ("pcg_state", C.c_void_p),
                ("pcg_inc", C.c_uint64)]


lib.orc_nee.argtypes = [C.POINTER(_Uniforms), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]


def nee(lights, position, normal, uniforms, cosine_once=False):
    """NextEventEstimation restated (orc_nee): one light sample per row of `uniforms` (3 used).
    Returns (k, 8): direction, radiance, distance, pdfW."""
    lights = np.ascontiguousarray(lights)
    pos = _f(position)
    nrm = _f(normal)
    u = _f(uniforms)
    out = np.zeros((u.shape[0], 8), np.float32)
    for i in range(u.shape[0]):
        un = _Uniforms(u[i].ctypes.data, u.shape[1], 0, None, 0)
        lib.orc_nee(C.byref(un), lights.ctypes.data, lights.shape[0], pos.ctypes.data, nrm.ctypes.data,
                    int(cosine_once), out[i].ctypes.data)
    return out


lib.orc_ptex_eval_batch.argtypes = [C.POINTER(_Ptex), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]


class PtexTexture:
    """Per-face texture data for the oracle's filter: texels (nfaces, res, res, 3) uint8,
    adjfaces / adjedges (nfaces, 4); tri: a triangle mesh (Ptex's triangle filter)."""

    def __init__(self, texels, res_log2, adjfaces, adjedges, tri=False):
        self.texels = np.ascontiguousarray(texels, np.uint8)
        self.adjfaces = np.ascontiguousarray(adjfaces, np.int32)
        self.adjedges = np.ascontiguousarray(adjedges, np.int32)
        self.s = _Ptex(int(res_log2), self.texels.shape[0], self.texels.ctypes.data, self.adjfaces.ctypes.data,
                       self.adjedges.ctypes.data, int(bool(tri)))

    def eval(self, faces, u, v):
        faces = np.ascontiguousarray(faces, np.int32)
        u = np.ascontiguousarray(u, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        out = np.zeros((faces.shape[0], 3), np.float32)
        lib.orc_ptex_eval_batch(C.byref(self.s), faces.shape[0], faces.ctypes.data, u.ctypes.data, v.ctypes.data,
                                out.ctypes.data)
        return out


# ---- instanced scenes (orc_trace.c) -------------------------------------------------------------

class _Inst(C.Structure):
    _fields_ = [("w2l", C.c_float * 12), ("blas", C.c_void_p)]


lib.orc_inst_closest.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(_Rays), C.c_uint32, C.POINTER(_Hits), C.c_void_p,
                                 C.c_int]
lib.orc_inst_occluded.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(_Rays), C.c_uint32, C.c_void_p, C.c_int]


class InstancedScene:
    """Composed instances: (world-to-local float rows of 4 (3 x 4), CpuBvh of the placed scene)."""

    def __init__(self, insts):
        self.blas = [b for _, b in insts]
        self.arr = (_Inst * max(len(insts), 1))()
        for k, (w2l, b) in enumerate(insts):
            self.arr[k] = _Inst((C.c_float * 12)(*[float(x) for x in np.asarray(w2l, np.float32).reshape(-1)]),
                                b.handle)
        self.n = len(insts)

    def closest(self, rays, threads=0):
        """Closest hits of a ray dict (RAY_KEYS): t, u, v, prim (in the placed scene), inst."""
        rs, keep, n = _rays(rays)
        out = {"t": np.zeros(n, np.float32), "u": np.zeros(n, np.float32), "v": np.zeros(n, np.float32),
               "prim": np.zeros(n, np.uint32), "inst": np.zeros(n, np.uint32)}
        h = _Hits(*[out[k].ctypes.data for k in ("t", "prim", "u", "v")])
        lib.orc_inst_closest(C.cast(self.arr, C.c_void_p), self.n, C.byref(rs), n, C.byref(h), out["inst"].ctypes.data,
                             _threads(threads))
        return out

    def occluded(self, rays, threads=0):
        rs, keep, n = _rays(rays)
        occ = np.zeros(n, np.uint32)
        lib.orc_inst_occluded(C.cast(self.arr, C.c_void_p), self.n, C.byref(rs), n, occ.ctypes.data, _threads(threads))
        return occ


# ---- instanced render scenes ---------------------------------------------------------------------
# The composition of sgpu_scene_add_instance's levels as the library performs it: row-major 3 x 4
# affine transforms from column-major float input, products and inverses in double with the same
# operations in the same order (Python floats are IEEE doubles), rounded once to float.

def rows_from_colmajor(m_colmajor):
    """Column-major 4 x 4 floats -> row-major 3 x 4 doubles (l2w[4 r + c] = m[4 c + r])."""
    m = np.asarray(m_colmajor, np.float32).reshape(-1)
    return [float(m[4 * c + r]) for r in range(3) for c in range(4)]


def affine_mul(a, b):
    r = [0.0] * 12
    for i in range(3):
        for j in range(4):
            v = a[4 * i + 0] * b[0 + j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j]
            if j == 3:
                v += a[4 * i + 3]
            r[4 * i + j] = v
    return r


def affine_inverse(m):
    a, b, c, d, e, f, g, h, k = m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]
    A = e * k - f * h
    B = -(d * k - f * g)
    Cc = d * h - e * g
    det = a * A + b * B + c * Cc
    idet = 1.0 / det
    inv = [A * idet, -(b * k - c * h) * idet, (b * f - c * e) * idet,
           B * idet, (a * k - c * g) * idet, -(a * f - c * d) * idet,
           Cc * idet, -(a * h - b * g) * idet, (a * e - b * d) * idet]
    r = [0.0] * 12
    for i in range(3):
        r[4 * i + 0] = inv[3 * i + 0]
        r[4 * i + 1] = inv[3 * i + 1]
        r[4 * i + 2] = inv[3 * i + 2]
        r[4 * i + 3] = -(inv[3 * i + 0] * m[3] + inv[3 * i + 1] * m[7] + inv[3 * i + 2] * m[11])
    return r


def colmajor(m4):
    """A 4 x 4 matrix acting on column vectors -> column-major float32 (what the C ABI takes)."""
    return np.ascontiguousarray(np.asarray(m4, np.float64).T.reshape(-1), np.float32)


def compose_instances(sc):
    """Composed instances of an instanced scene dict, in composition order: (model index, l2w rows
    (double, 12), id0, id1)."""
    out = []
    for i, (kind, idx, m) in enumerate(sc["instances"]):
        l2w = rows_from_colmajor(colmajor(m))
        if kind == "model":
            out.append((idx, l2w, i, 0xFFFFFFFF))
        else:
            for j, (midx, m2) in enumerate(sc["subscenes"][idx]):
                out.append((midx, affine_mul(l2w, rows_from_colmajor(colmajor(m2))), i, j))
    return out


class _IScene(C.Structure):
    _fields_ = [("insts", C.c_void_p), ("n_inst", C.c_uint32), ("tri9", C.c_void_p), ("tri_geom", C.c_void_p),
                ("geom_first", C.c_void_p), ("geom_base", C.c_void_p), ("l2w", C.c_void_p), ("n_comp", C.c_uint32),
                ("curve_cp", C.c_void_p), ("curve_box", C.c_void_p), ("curve_geom", C.c_void_p),
                ("curve_prim", C.c_void_p), ("n_curves", C.c_void_p)]


class InstancedRenderScene:
    """The oracle's view of an instanced scene dict: "models" (scene dicts of geometries with their
    materials and textures), optional "subscenes" (lists of (model index, 4 x 4 local-to-world)),
    "instances" (("model" | "subscene", index, 4 x 4 local-to-world)), and the scene-level lights,
    background and optional "ibl". The placed scenes' geometries are numbered as one list in
    first-appearance order, as the library numbers them."""

    def __init__(self, sc: dict):
        from selas_amd.workloads import geom_tri9

        comp = compose_instances(sc)
        self.models = []
        for mdesc in sc["models"]:
            tris = [geom_tri9(g) for g in mdesc["geoms"]]
            tri9 = np.ascontiguousarray(np.concatenate(tris), np.float32)
            mm = {"tri9": tri9,
                  "tri_geom": np.concatenate([np.full(t.shape[0], i, np.uint32) for i, t in enumerate(tris)]),
                  "geom_first": np.cumsum([0] + [t.shape[0] for t in tris[:-1]]).astype(np.uint32),
                  "materials": np.ascontiguousarray(mdesc["materials"], np.float32),
                  "geom_material": np.ascontiguousarray(mdesc["geom_material"], np.uint32),
                  "n_geoms": len(tris), "bvh": CpuBvh(tri9) if tri9.shape[0] else None, "geoms": mdesc["geoms"]}
            pts = tri9.reshape(-1, 3)
            mm["box"] = (pts.min(0), pts.max(0)) if pts.shape[0] else None
            mm["curves"] = _curve_arrays(mdesc["geoms"])
            if mm["box"] is not None and mm["curves"] is not None:
                # the placed scene's box takes in its curves' boxes (the library's instance box for the
                # scene's bounds is this box through the transform)
                cb = mm["curves"][1]
                mm["box"] = (np.minimum(mm["box"][0], cb[:, :3].min(0)), np.maximum(mm["box"][1], cb[:, 3:].max(0)))
            self.models.append(mm)
        # composed instances: those of scenes with triangles (the library's top tree), then those of
        # scenes of curves only, each in composition order
        n_tri_comp = len([c for c in comp if self.models[c[0]]["tri9"].shape[0] > 0])
        comp = ([c for c in comp if self.models[c[0]]["tri9"].shape[0] > 0] +
                [c for c in comp if self.models[c[0]]["tri9"].shape[0] == 0 and self.models[c[0]]["curves"] is not None])
        order = []
        for c in comp:
            if c[0] not in order:
                order.append(c[0])
        g0, m0, ng, nm = {}, {}, 0, 0
        for mi in order:
            g0[mi], m0[mi] = ng, nm
            ng += self.models[mi]["n_geoms"]
            nm += self.models[mi]["materials"].shape[0]
        mats = [self.models[mi]["materials"] for mi in order]
        self.materials = np.ascontiguousarray(np.concatenate(mats) if mats else np.zeros((1, 22)), np.float32)
        gm = [self.models[mi]["geom_material"][:self.models[mi]["n_geoms"]] + m0[mi] for mi in order]
        self.geom_material = np.ascontiguousarray(np.concatenate(gm) if gm else np.zeros(1), np.uint32)
        self.attrs = _AttrTable([g for mi in order for g in self.models[mi]["geoms"]])
        self.textures = []
        self.geom_tex = (_Ptex * max(ng, 1))()
        for mi in order:
            for gi, g in enumerate(self.models[mi]["geoms"]):
                if isinstance(g, dict) and g.get("ptex") is not None:
                    p = g["ptex"]
                    t = PtexTexture(p["texels"], p["res_log2"], p["adjfaces"], p["adjedges"], tri="triangles" in g)
                    self.textures.append(t)
                    self.geom_tex[g0[mi] + gi] = t.s
        n = len(comp)
        self.curve_ptrs = [(C.c_void_p * max(n, 1))() for _ in range(4)]
        self.n_curves = np.zeros(max(n, 1), np.uint32)
        self.inst_arr = (_Inst * max(n, 1))()
        self.tri9_ptrs = (C.c_void_p * max(n, 1))()
        self.tri_geom_ptrs = (C.c_void_p * max(n, 1))()
        self.geom_first_ptrs = (C.c_void_p * max(n, 1))()
        self.geom_base = np.zeros(max(n, 1), np.uint32)
        self.l2w = np.zeros((max(n, 1), 9), np.float32)
        lo = np.full(3, np.inf, np.float32)
        hi = np.full(3, -np.inf, np.float32)
        for k, (mi, l2w, _, _) in enumerate(comp):
            mm = self.models[mi]
            w2l = np.array(affine_inverse(l2w), np.float32)
            self.inst_arr[k] = _Inst((C.c_float * 12)(*[float(x) for x in w2l]),
                                     mm["bvh"].handle if mm["bvh"] is not None else None)
            if mm["curves"] is not None:
                for a, arr in zip(self.curve_ptrs, mm["curves"]):
                    a[k] = arr.ctypes.data
                self.n_curves[k] = mm["curves"][0].shape[0]
            self.tri9_ptrs[k] = mm["tri9"].ctypes.data
            self.tri_geom_ptrs[k] = mm["tri_geom"].ctypes.data
            self.geom_first_ptrs[k] = mm["geom_first"].ctypes.data
            self.geom_base[k] = g0[mi]
            self.l2w[k] = [l2w[4 * r + c] for r in range(3) for c in range(3)]
            if mm["box"] is None:
                continue
            # the instance's world box: the placed scene's box corners through l2w in double,
            # rounded outward to float (the library's top-tree boxes; they bound the emission sphere)
            blo, bhi = mm["box"]
            wlo, whi = [np.inf] * 3, [-np.inf] * 3
            for cidx in range(8):
                pnt = [float(bhi[0] if cidx & 1 else blo[0]), float(bhi[1] if cidx & 2 else blo[1]),
                       float(bhi[2] if cidx & 4 else blo[2])]
                for r in range(3):
                    wv = l2w[4 * r] * pnt[0] + l2w[4 * r + 1] * pnt[1] + l2w[4 * r + 2] * pnt[2] + l2w[4 * r + 3]
                    wlo[r] = min(wlo[r], wv)
                    whi[r] = max(whi[r], wv)
            for r in range(3):
                lo[r] = min(lo[r], np.nextafter(np.float32(wlo[r]), np.float32(-np.inf)))
                hi[r] = max(hi[r], np.nextafter(np.float32(whi[r]), np.float32(np.inf)))
        for k, (mi, l2w, _, _) in enumerate(comp):
            mm = self.models[mi]
            if mm["curves"] is None:
                continue
            # the curves carried into world space as the library carries them (the composed transform in
            # double, each coordinate rounded once; radii by the transform's scale), then their boxes
            n2 = [l2w[c] * l2w[c] + l2w[4 + c] * l2w[4 + c] + l2w[8 + c] * l2w[8 + c] for c in range(3)]
            s2 = 0.0
            for c in range(3):
                s2 += n2[c] / 3.0
            sc_ = np.sqrt(s2)
            src = mm["curves"][0].reshape(-1, 4).astype(np.float64)
            w = np.empty((src.shape[0], 4), np.float32)
            for r in range(3):
                w[:, r] = ((l2w[4 * r] * src[:, 0] + l2w[4 * r + 1] * src[:, 1]) + l2w[4 * r + 2] * src[:, 2]) + l2w[4 * r + 3]
            w[:, 3] = src[:, 3] * sc_
            w = w.reshape(-1, 4, 4)
            rmax = w[:, :, 3].max(1)
            clo = (w[:, :, :3].min(1) - rmax[:, None]).astype(np.float32)
            chi = (w[:, :, :3].max(1) + rmax[:, None]).astype(np.float32)
            lo = np.minimum(lo, clo.min(0))
            hi = np.maximum(hi, chi.max(0))
        self.iscene = _IScene(C.cast(self.inst_arr, C.c_void_p), n_tri_comp, C.cast(self.tri9_ptrs, C.c_void_p),
                              C.cast(self.tri_geom_ptrs, C.c_void_p), C.cast(self.geom_first_ptrs, C.c_void_p),
                              self.geom_base.ctypes.data, self.l2w.ctypes.data, n,
                              *[C.cast(a, C.c_void_p) for a in self.curve_ptrs],
                              self.n_curves.ctypes.data if self.n_curves.any() else None)
        half = np.float32(0.5)
        d = hi - lo
        self.sphere = np.array([(lo[0] + hi[0]) * half, (lo[1] + hi[1]) * half, (lo[2] + hi[2]) * half,
                                half * np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))], np.float32)
        ibl = sc.get("ibl")
        self.ibl = None if ibl is None else Ibl(ibl["rgb"], ibl.get("rotation", 0.0), ibl.get("exposure", 1.0),
                                                sphere=self.sphere)
        self.lights = np.ascontiguousarray(sc["lights"])
        self.s = _Scene(None, None, None, self.materials.ctypes.data, self.geom_material.ctypes.data, max(ng, 1),
                        self.lights.ctypes.data if self.lights.size else None, self.lights.shape[0],
                        (C.c_float * 3)(*[float(x) for x in sc["background"]]), None,
                        C.cast(self.geom_tex, C.c_void_p),
                        C.cast(C.pointer(self.ibl.s), C.c_void_p) if self.ibl is not None else None,
                        C.cast(C.pointer(self.iscene), C.c_void_p), self.attrs.ptr)


# ---- round B-spline curves (orc_curve.c) ---------------------------------------------------------

_D = C.POINTER(C.c_double)
lib.orc_curve_hit_dense.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, _D, _D, C.c_void_p]
lib.orc_curve_hit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_int, _D, _D, C.c_void_p]
lib.orc_curve_clearance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
lib.orc_curve_clearance.restype = C.c_double
lib.orc_curves_closest.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(_Rays), C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int, C.c_int]
lib.orc_curves_occluded.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(_Rays), C.c_uint32, C.c_void_p, C.c_int]
lib.orc_curve_boxes.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]


def curve_hit(cp, o, d, tnear=0.0, tfar=3.4e38, dense=False, any_hit=False):
    """One segment (4 x (x, y, z, r) control points): (t, u, ng) in double, or None on a miss."""
    cp = np.ascontiguousarray(cp, np.float32).reshape(16)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    t, u = C.c_double(), C.c_double()
    ng = np.zeros(3, np.float64)
    if dense:
        hit = lib.orc_curve_hit_dense(cp.ctypes.data, o.ctypes.data, d.ctypes.data, tnear, tfar, C.byref(t), C.byref(u),
                                      ng.ctypes.data)
    else:
        hit = lib.orc_curve_hit(cp.ctypes.data, o.ctypes.data, d.ctypes.data, tnear, tfar, int(any_hit), C.byref(t),
                                C.byref(u), ng.ctypes.data)
    return (t.value, u.value, ng) if hit else None


def curve_clearance(cp, o, d):
    cp = np.ascontiguousarray(cp, np.float32).reshape(16)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    return lib.orc_curve_clearance(cp.ctypes.data, o.ctypes.data, d.ctypes.data)


def curves_closest(cps, rays, dense=False, threads=0):
    """Closest hits of a ray dict (RAY_KEYS) on curve segments cps (nseg, 16): t, u, prim (segment,
    0xFFFFFFFF on a miss), ng (n, 3)."""
    cps = np.ascontiguousarray(cps, np.float32).reshape(-1, 16)
    rs, keep, n = _rays(rays)
    out = {"t": np.zeros(n, np.float32), "u": np.zeros(n, np.float32), "prim": np.zeros(n, np.uint32),
           "ng": np.zeros((n, 3), np.float32)}
    lib.orc_curves_closest(cps.ctypes.data, cps.shape[0], C.byref(rs), n, out["t"].ctypes.data, out["u"].ctypes.data,
                           out["prim"].ctypes.data, out["ng"].ctypes.data, int(dense), _threads(threads))
    return out


def curves_occluded(cps, rays, threads=0):
    cps = np.ascontiguousarray(cps, np.float32).reshape(-1, 16)
    rs, keep, n = _rays(rays)
    occ = np.zeros(n, np.uint32)
    lib.orc_curves_occluded(cps.ctypes.data, cps.shape[0], C.byref(rs), n, occ.ctypes.data, _threads(threads))
    return occ
