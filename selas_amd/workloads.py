"""Synthetic workloads, defined concretely and rebuilt from seeds.

BVH traversal microbenchmark (the 2M-triangle / 16M-ray workload):
  * 2,097,152 triangles = 256 Gaussian clusters x 8,192; cluster centres uniform in [0,1]^3,
    triangle centroids normal around them with sigma 0.02, two edge lengths log-uniform in
    [1e-4, 1e-2], random orientation (sgpu_gen_cluster_soup).
  * 16,777,216 rays: 8,388,608 primary rays of a 1024x1024 pinhole camera looking at the unit box,
    the first 8 samples of each pixel's 4x4 correlated multi-jitter set (Selas's JitteredCameraRay
    with CorrelatedMultiJitter, p = pixel index), followed by 8,388,608 cosine-distributed rays
    leaving uniformly chosen points of uniformly chosen triangles (sgpu_gen_bounce_rays).
  * closest hit with tnear = 0, tfar = +inf; any hit on the same rays with tfar = 0.5 * t_closest
    for a seeded half (never occluded) and tfar = +inf for the other half (occluded exactly when a
    closest hit exists).
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import native as N

CAMERA_POS = (0.5, 0.5, -1.2)
CAMERA_LOOK_AT = (0.5, 0.5, 0.5)
CAMERA_UP = (0.0, 1.0, 0.0)
CAMERA_FOV = 50.0


def traversal_scene(scale: float = 1.0, seed: int = 2024):
    """Triangle soup (n, 9) of the traversal workload; scale < 1 keeps the cluster count and
    shrinks the triangles per cluster."""
    per = max(1, int(round(8192 * scale)))
    return N.gen_cluster_soup(seed, 256, per, 0.02, 1e-4, 1e-2)


def traversal_rays(tri9: np.ndarray, width: int = 1024, height: int = 1024, spp: int = 8, n_bounce=None,
                   seed: int = 77):
    cam = N.gen_camera_rays(CAMERA_POS, CAMERA_LOOK_AT, CAMERA_UP, CAMERA_FOV, width, height, spp)
    nb = width * height * spp if n_bounce is None else n_bounce
    bounce = N.gen_bounce_rays(seed, tri9, nb)
    return {k: np.concatenate([cam[k], bounce[k]]) for k in N.RAY_FIELDS}


def icosphere(subdivisions: int, centre, radius, mesh: bool = False):
    """Icosahedron subdivided `subdivisions` times (20 * 4^n faces), vertices on the sphere,
    faces wound so that (v1 - v0) x (v2 - v0) points outward. Returns (n, 9) float32, or with
    `mesh` the unit-sphere vertices and the faces (vertex indices)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts)
    f = np.array(faces)
    tri = v[f]  # (n, 3, 3)
    ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    flip = np.sum(ng * tri.mean(axis=1), axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    tri = v[f] * radius + np.asarray(centre)
    if mesh:
        return v, f
    return tri.reshape(-1, 9).astype(np.float32)


def icosphere_mesh(subdivisions: int, centre, radius):
    """The icosphere as an indexed triangle mesh with per-vertex shading attributes: unit radial
    normals, tangents (the unit direction of cross(a, n) with a = +y, or +x near the poles; w = +1
    where x >= 0, else -1) and longitude/latitude UVs."""
    v, f = icosphere(subdivisions, centre, radius, mesh=True)
    pos = (v * radius + np.asarray(centre)).astype(np.float32)
    n = v / np.linalg.norm(v, axis=1, keepdims=True)
    a = np.where(np.abs(n[:, 1:2]) > 0.99, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    t = np.cross(a, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    w = np.where(n[:, 0] >= 0.0, 1.0, -1.0)
    uv = np.stack([np.arctan2(n[:, 2], n[:, 0]) / (2 * np.pi) + 0.5, np.arccos(np.clip(n[:, 1], -1, 1)) / np.pi], 1)
    return {"positions": pos, "triangles": f.astype(np.uint32), "normals": n.astype(np.float32),
            "tangents": np.concatenate([t, w[:, None]], 1).astype(np.float32), "uvs": uv.astype(np.float32)}


def attributed_panel(nu: int, nv: int, origin, eu, ev, seed: int):
    """A gently curved quad grid of nu x nv quads spanning origin + [0,1] eu + [0,1] ev (normal
    eu x ev), with per-vertex normals tilted by a smooth field, tangents along eu (w alternating
    by column) and UVs: a quad geometry whose shading frame differs from its faces'."""
    rng = np.random.default_rng(seed)
    origin, eu, ev = (np.asarray(x, np.float64) for x in (origin, eu, ev))
    nrm = np.cross(eu, ev)
    nrm /= np.linalg.norm(nrm)
    us, vs = np.meshgrid(np.linspace(0, 1, nu + 1), np.linspace(0, 1, nv + 1), indexing="ij")
    bump = 0.05 * np.sin(np.pi * us) * np.sin(np.pi * vs)
    pos = origin + us[..., None] * eu + vs[..., None] * ev + bump[..., None] * nrm
    ph = rng.uniform(0, 2 * np.pi, 2)
    tilt_u = 0.3 * np.sin(2 * np.pi * us + ph[0])
    tilt_v = 0.3 * np.cos(2 * np.pi * vs + ph[1])
    eun = eu / np.linalg.norm(eu)
    evn = ev / np.linalg.norm(ev)
    n = nrm + tilt_u[..., None] * eun + tilt_v[..., None] * evn
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    t = np.broadcast_to(eun, n.shape) + 0.2 * tilt_v[..., None] * nrm
    w = np.where((np.arange(nu + 1)[:, None] + 0 * np.arange(nv + 1)[None, :]) % 2 == 0, 1.0, -1.0)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i, j = i.reshape(-1), j.reshape(-1)
    vid = lambda a, b: a * (nv + 1) + b
    quads = np.stack([vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)], 1)
    return {"positions": pos.reshape(-1, 3).astype(np.float32), "quads": quads.astype(np.uint32),
            "normals": n.reshape(-1, 3).astype(np.float32),
            "tangents": np.concatenate([t.reshape(-1, 3), w.reshape(-1, 1)], 1).astype(np.float32),
            "uvs": np.stack([us.reshape(-1), vs.reshape(-1)], 1).astype(np.float32)}


def attributed_room_scene(seed: int = 1):
    """room_scene with shading attributes: the icosphere as a smooth mesh (vertex normals, tangents,
    UVs; anisotropic metallic Disney) and a curved quad panel on the left wall with tilted vertex
    normals and tangents (anisotropic Disney, dielectric)."""
    sc = room_scene(seed)
    geoms = sc["geoms"][:-1] + [icosphere_mesh(3, (0.5, 0.3, 0.55), 0.25),
                                attributed_panel(6, 5, (0.02, 0.15, 0.3), (0.0, 0.6, 0.0), (0.0, 0.0, 0.5), seed)]
    mats = list(sc["materials"][:-1]) + [material(1, (0.8, 0.7, 0.6), 1.0, 0.35, anisotropic=0.7),
                                         material(0, (0.3, 0.6, 0.8), 0.0, 0.4, anisotropic=0.5, sheen=0.3)]
    return dict(sc, geoms=geoms, materials=np.stack(mats), geom_material=np.arange(len(geoms), dtype=np.uint32))


def material(shader, base_color, metallic=0.3, roughness=0.5, ior=1.5, **kw):
    """A 22-float material record (raw MaterialResourceData values; defaults are Selas's
    ImportMaterial defaults, Source/Core/BuildCommon/ImportMaterial.cpp:36-51)."""
    m = np.zeros(22, np.float32)
    m[0:3] = base_color
    m[3:6] = kw.get("transmittance_color", (1.0, 1.0, 1.0))
    m[6] = kw.get("sheen", 0.0)
    m[7] = kw.get("sheen_tint", 0.0)
    m[8] = kw.get("clearcoat", 0.0)
    m[9] = kw.get("clearcoat_gloss", 0.0)
    m[10] = metallic
    m[11] = kw.get("spec_trans", 0.0)
    m[12] = kw.get("diff_trans", 0.0)
    m[13] = kw.get("flatness", 0.0)
    m[14] = kw.get("anisotropic", 0.0)
    m[16] = kw.get("specular_tint", 0.0)
    m[17] = roughness
    m[18] = kw.get("scatter_distance", 0.0)
    m[19] = ior
    m[20] = shader
    m[21] = kw.get("flags", 0)
    return m


LIGHT_DTYPE = np.dtype([("type", np.uint32), ("position", np.float32, 3), ("direction", np.float32, 3),
                        ("x", np.float32, 3), ("z", np.float32, 3), ("radiance", np.float32, 3)])


def room_scene(seed: int = 1):
    """Closed-room path tracing scene: the unit box [0,1]^3 as 12 axis-aligned quads (each of the
    6 faces split in two halves; 24 triangles, normals inward), Lambert walls with albedo uniform in
    [0.2, 0.8] per channel per quad, an icosphere subdivided 5 times (20,480 triangles, centre
    (0.5, 0.3, 0.55), radius 0.25) with the Disney BSDF (solid, metallic 1, roughness 0.3, base
    colour 0.8), one 0.3 x 0.3 quad light of radiance 15 facing down just below the ceiling (not
    geometry, as in Selas), black background. Camera inside the room at z = 0.02 looking at +z."""
    rng = np.random.default_rng(seed)
    quads = []
    for axis in range(3):
        for side in (0.0, 1.0):
            for half in (0, 1):
                u_ax, v_ax = [a for a in range(3) if a != axis]
                lo = np.zeros(3)
                lo[axis] = side
                lo[u_ax] = 0.5 * half
                eu = np.zeros(3)
                eu[u_ax] = 0.5
                ev = np.zeros(3)
                ev[v_ax] = 1.0
                p0, p1, p2, p3 = lo, lo + eu, lo + eu + ev, lo + ev
                tris = np.array([[p0, p1, p2], [p0, p2, p3]])
                inward = np.zeros(3)
                inward[axis] = 1.0 if side == 0.0 else -1.0
                ng = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
                flip = np.sum(ng * inward, axis=1) < 0
                tris[flip] = tris[flip][:, [0, 2, 1]]
                quads.append(tris.reshape(-1, 9).astype(np.float32))
    sphere = icosphere(5, (0.5, 0.3, 0.55), 0.25)
    geoms = quads + [sphere]
    mats = [material(2, rng.uniform(0.2, 0.8, 3)) for _ in quads] + [material(1, (0.8, 0.8, 0.8), 1.0, 0.3)]
    light = np.zeros(1, LIGHT_DTYPE)
    light[0] = (0, (0.5, 0.999, 0.5), (0.0, -1.0, 0.0), (0.3, 0.0, 0.0), (0.0, 0.0, 0.3), (15.0, 15.0, 15.0))
    return {
        "geoms": geoms,
        "materials": np.stack(mats),
        "geom_material": np.arange(len(geoms), dtype=np.uint32),
        "lights": light,
        "background": np.zeros(3, np.float32),
        "camera": {"position": (0.5, 0.5, 0.02), "look_at": (0.5, 0.45, 1.0), "up": (0.0, 1.0, 0.0), "fov": 70.0},
    }


def value_noise_2d(rng, lattice, x, z):
    """Smooth value noise on [0,1]^2: random lattice values, smoothstep-weighted bilinear blend."""
    g = rng.random((lattice + 1, lattice + 1))
    fx, fz = x * lattice, z * lattice
    ix = np.clip(np.floor(fx).astype(int), 0, lattice - 1)
    iz = np.clip(np.floor(fz).astype(int), 0, lattice - 1)
    tx, tz = fx - ix, fz - iz
    sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
    a = g[ix, iz] * (1 - sx) + g[ix + 1, iz] * sx
    b = g[ix, iz + 1] * (1 - sx) + g[ix + 1, iz + 1] * sx
    return a * (1 - sz) + b * sz


def quad_grid_mesh(nu: int, nv: int, width: float, depth: float, heights: np.ndarray):
    """A displaced grid of nu x nv quads; each quad (v0, v1, v2, v3) is split as Embree splits
    quads: triangles (v0, v1, v3) and (v2, v3, v1). Returns (2 nu nv, 9) float32, quad-major."""
    xs = np.linspace(0.0, width, nu + 1)
    zs = np.linspace(0.0, depth, nv + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    P = np.stack([X, heights, Z], axis=-1)  # (nu+1, nv+1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i, j = i.T.reshape(-1), j.T.reshape(-1)  # quad index = j * nu + i
    v0, v1, v2, v3 = P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]
    ta = np.stack([v0, v1, v3], 1)
    tb = np.stack([v2, v3, v1], 1)
    tris = np.stack([ta, tb], 1).reshape(-1, 9)
    return tris


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _frame_from_normal(n, twist):
    """Rotation taking +y to the unit vector n, preceded by a rotation of `twist` radians about y."""
    n = np.asarray(n, np.float64)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 0.0, 1.0])
    t = np.cross(a, n)
    t /= np.linalg.norm(t)
    b = np.cross(t, n)
    M = np.stack([t, n, b], axis=1)  # columns: images of x, y, z
    c, s_ = np.cos(twist), np.sin(twist)
    Ry = np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]])
    return M @ Ry


def _smooth_interp_matrix(n_samples, lattice, n_units):
    """(n_samples x lattice+1) weights of a smoothstep-blended linear interpolation of lattice
    values at sample positions (k + 0.5) / n_samples * lattice, k = 0..n_samples-1."""
    x = (np.arange(n_samples) + 0.5) / n_samples * lattice
    ix = np.clip(np.floor(x).astype(int), 0, lattice - 1)
    t = x - ix
    s = t * t * (3 - 2 * t)
    W = np.zeros((n_samples, lattice + 1))
    W[np.arange(n_samples), ix] = 1 - s
    W[np.arange(n_samples), ix + 1] += s
    return W


def quad_adjacency(quads):
    """Ptex adjacency of a quad mesh ((n, 4) vertex indices, edge e from corner e to corner e+1):
    the face across each edge (-1 at a border) and that face's edge id of the shared edge."""
    quads = np.asarray(quads, np.int64)
    n = quads.shape[0]
    a = quads.reshape(-1)
    b = np.roll(quads, -1, axis=1).reshape(-1)
    key = a * (a.max() + 1 + b.max() + 1) + b
    rkey = b * (a.max() + 1 + b.max() + 1) + a
    order = np.argsort(key)
    pos = np.searchsorted(key[order], rkey)
    pos = np.clip(pos, 0, key.size - 1)
    match = key[order][pos] == rkey
    other = order[pos]
    adjfaces = np.where(match, other // 4, -1).reshape(n, 4).astype(np.int32)
    adjedges = np.where(match, other % 4, 0).reshape(n, 4).astype(np.int32)
    return adjfaces, adjedges


def grid_ptex(rng, nu, nv, res_log2, tint, lattice=(32, 16)):
    """Per-face RGB8 value-noise texture of an nu x nv grid mesh whose quad (i, j) is face
    j * nu + i with Ptex u along j and v along i (the quads of vcm_scene): a smooth field over the
    whole mesh sampled at each face's texel centres, so neighbouring faces continue each other.
    Channel c is round(255 tint_c (0.35 + 0.65 F_c)) with F_c = Wi L_c Wj^T in [0, 1]; since the
    rows of Wi and Wj sum to 1 the affine map is applied to the lattice values L_c."""
    res = 1 << res_log2
    Wi = _smooth_interp_matrix(nu * res, lattice[0], nu)  # along i (= face v)
    Wj = _smooth_interp_matrix(nv * res, lattice[1], nv)  # along j (= face u)
    nb = lattice[1] + 1
    # T[(i, tv), (c, b)] = (Wi L_c)[i*res + tv, b]
    T = np.concatenate([Wi @ (255 * tint[c] * (0.35 + 0.65 * rng.random((lattice[0] + 1, nb))))
                        for c in range(3)], axis=1).astype(np.float32)
    # block-diagonal right factor: Bm[(c, b), (j, tu, c)] = Wj[j*res + tu, b]
    Bm = np.zeros((3, nb, nv, res, 3), np.float32)
    Wj3 = Wj.reshape(nv, res, nb).transpose(2, 0, 1)
    for c in range(3):
        Bm[c, :, :, :, c] = Wj3
    out = T @ Bm.reshape(3 * nb, nv * res * 3)  # [(i, tv), (j, tu, c)]
    np.rint(out, out=out)  # in [0, 255] since tint <= 1
    tex = out.astype(np.uint8).reshape(nu * res, nv, res * 3).transpose(1, 0, 2)  # [j, (i, tv), (tu, c)]
    return np.ascontiguousarray(tex).reshape(nv * nu, res, res, 3)


def geom_tri9(g):
    """Triangles (n, 9) of a geometry entry: a triangle soup as is, an indexed triangle mesh
    ({"positions", "triangles"}) in order, a quad geometry ({"positions", "quads"}) split as Embree
    splits quads, (v0, v1, v3) then (v2, v3, v1) for each quad in order."""
    if isinstance(g, np.ndarray):
        return g
    if "curves" in g:  # a curve geometry holds no triangles
        return np.zeros((0, 9), np.float32)
    p = np.asarray(g["positions"], np.float32)
    if "triangles" in g:
        t = np.asarray(g["triangles"], np.int64)
        return np.ascontiguousarray(p[t].reshape(-1, 9), np.float32)
    q = np.asarray(g["quads"], np.int64)
    ta = np.stack([p[q[:, 0]], p[q[:, 1]], p[q[:, 3]]], 1)
    tb = np.stack([p[q[:, 2]], p[q[:, 3]], p[q[:, 1]]], 1)
    return np.stack([ta, tb], 1).reshape(-1, 9).astype(np.float32)


MAT_TRANSPARENT = 1
MAT_USES_PTEX = 16


def sky_map(h: int, w: int) -> np.ndarray:
    """An equirectangular sky (h x w x 3 float32 radiance): brighter toward +y, with a broad bright
    lobe around (0.5, 0.7, 0.3)."""
    th = (np.arange(h) + 0.5) / h * np.pi
    ph = (np.arange(w) + 0.5) / w * 2 * np.pi - np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    sun = np.array([0.5, 0.7, 0.3])
    sun /= np.linalg.norm(sun)
    lobe = np.maximum(d @ sun, 0) ** 8
    base = 0.3 + 0.7 * np.maximum(d[..., 1], 0)
    rgb = np.stack([base * 0.6 + 4 * lobe, base * 0.8 + 3.5 * lobe, base * 1.0 + 2.5 * lobe], -1)
    return rgb.astype(np.float32)


def vcm_scene(seed: int = 3, n_meshes: int = 64, nu: int = 128, nv: int = 64, textures: bool = True,
              res_log2: int = 4):
    """VCM workload scene: n_meshes displaced-grid quad meshes of nu x nv quads (2 nu nv triangles
    each; 64 x 16,384 = 1,048,576 triangles) placed by a seeded RNG as panels of an enclosure: each
    mesh is a unit-by-half grid with value-noise displacement, scaled by 4..8, centred at a random
    direction and distance 3.5..5 from the centre of the 10^3 box and turned so that its front
    face (the side its geometric normal points to) faces the centre, with a random twist. Per mesh,
    Disney parameters: roughness uniform in [0.05, 1], metallic 0 or 1 (p 0.5), specTrans 1 with
    probability 0.2 (else 0), and (with `textures`) a per-face 16 x 16 RGB8 value-noise texture
    tinted by a base colour uniform in [0.2, 0.9]^3 (without textures that colour is the material's
    base colour). Lights inside the enclosure: one 1.5 x 1.5 quad light (radiance 20) facing down
    and one point light (intensity 15). The camera is inside too, so every surface it sees faces it
    (Selas's surfaces are one-sided for opaque materials: no normal flip in CalculateSurfaceParams).
    Merge radius 0.3 % of the scene diagonal."""
    rng = np.random.default_rng(seed)
    geoms, mats, tints = [], [], []
    centre = np.array([5.0, 5.0, 5.0])
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="xy")  # quad (i, j) = face j * nu + i
    pid = lambda i, j: i * (nv + 1) + j  # vertex (i, j), i along x, j along z
    quads = np.stack([pid(ii, jj), pid(ii, jj + 1), pid(ii + 1, jj + 1), pid(ii + 1, jj)], -1).reshape(-1, 4)
    adj = quad_adjacency(quads) if textures else None
    for m in range(n_meshes):
        xs = np.linspace(0, 1, nu + 1)
        zs = np.linspace(0, 0.5, nv + 1)
        X, Z = np.meshgrid(xs, zs, indexing="ij")
        H = 0.12 * value_noise_2d(rng, 8, X, Z * 2)
        local = np.stack([X, H, Z], -1).reshape(-1, 3) - np.array([0.5, 0.0, 0.25])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        R = _frame_from_normal(-d, rng.uniform(0, 2 * np.pi))
        s = rng.uniform(4.0, 8.0)
        t = centre + d * rng.uniform(3.5, 5.0)
        world = ((local @ R.T) * s + t).astype(np.float32)
        metallic = 1.0 if rng.random() < 0.5 else 0.0
        spec_trans = 1.0 if rng.random() < 0.2 else 0.0
        base = rng.uniform(0.2, 0.9, 3)
        flags = MAT_USES_PTEX if textures else 0
        mats.append(material(1, base, metallic, rng.uniform(0.05, 1.0), 1.5, spec_trans=spec_trans, flags=flags))
        g = {"positions": world, "quads": quads.astype(np.uint32)}
        if textures:
            g["ptex"] = {"res_log2": res_log2, "texels": None, "adjfaces": adj[0], "adjedges": adj[1]}
            tints.append(base)
        geoms.append(g)
    if textures:  # each mesh's texture has its own stream (seed, m): synthesised in parallel
        with ThreadPoolExecutor(min(8, n_meshes)) as ex:
            texels = ex.map(lambda m: grid_ptex(np.random.default_rng((seed, m)), nu, nv, res_log2, tints[m]),
                            range(n_meshes))
            for g, t in zip(geoms, texels):
                g["ptex"]["texels"] = t
    lights = np.zeros(2, LIGHT_DTYPE)
    lights[0] = (0, (5.0, 7.5, 5.0), (0.0, -1.0, 0.0), (1.5, 0.0, 0.0), (0.0, 0.0, 1.5), (20.0, 20.0, 20.0))
    lights[1] = (1, (4.0, 4.0, 6.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (15.0, 15.0, 15.0))
    allp = np.concatenate([g["positions"] for g in geoms]).reshape(-1, 3)
    diag = float(np.linalg.norm(allp.max(0) - allp.min(0)))
    return {
        "geoms": geoms,
        "materials": np.stack(mats),
        "geom_material": np.arange(n_meshes, dtype=np.uint32),
        "lights": lights,
        "background": np.zeros(3, np.float32),
        "camera": {"position": (5.0, 5.0, 3.0), "look_at": (5.0, 4.6, 8.0), "up": (0.0, 1.0, 0.0), "fov": 70.0},
        "merge_radius": 0.003 * diag,
        "radius_alpha": 0.75,
        "max_path_length": 10,
    }


def sized_vcm_scene(seed: int, n_tris: float, textures: bool = True):
    """vcm_scene with 64 meshes of nu x nv quads, nu = 2 nv, sized so that 128 nu nv is about
    n_tris (9,216 triangles at 10K, 2,120,448 at 2M)."""
    nv = max(2, int(round(np.sqrt(n_tris / 256.0))))
    return vcm_scene(seed=seed, n_meshes=64, nu=2 * nv, nv=nv, textures=textures)


def job_mix(seed: int, n_jobs: int, n_scenes: int = 0):
    """Uneven VCM render jobs: square images with side log-uniform in [64, 1024] (pixel count
    log-uniform in 64^2..1024^2), iterations uniform in 1..64, path length 10, per-job seeds. With
    n_scenes == 0 every job has a scene of its own (seed, triangle count log-uniform in 10K..2M);
    otherwise each job picks one of n_scenes shared scenes uniformly (scene ids 0..n_scenes-1, whose
    sizes scene_sizes gives)."""
    rng = np.random.default_rng(seed)
    side = np.rint(np.exp(rng.uniform(np.log(64), np.log(1024), n_jobs))).astype(int)
    its = rng.integers(1, 65, n_jobs)
    seeds = rng.integers(0, 1 << 40, n_jobs)
    if n_scenes:
        scene = rng.integers(0, n_scenes, n_jobs)
        tris = np.zeros(n_jobs)
    else:
        scene = np.arange(n_jobs)
        tris = np.exp(rng.uniform(np.log(1e4), np.log(2e6), n_jobs))
    return [{"width": int(side[k]), "height": int(side[k]), "iterations": int(its[k]), "max_path_length": 10,
             "seed": int(seeds[k]), "scene": int(scene[k]), "scene_tris": float(tris[k])} for k in range(n_jobs)]


def scene_sizes(seed: int, n_scenes: int):
    """Triangle counts (log-uniform in 10K..2M) of n_scenes shared scenes; scene k is
    sized_vcm_scene(seed * 1000 + k, size[k])."""
    rng = np.random.default_rng((seed, n_scenes))
    return np.exp(rng.uniform(np.log(1e4), np.log(2e6), n_scenes))


def anyhit_tfar(t_closest: np.ndarray, prim: np.ndarray, seed: int = 5):
    """tfar for the any-hit pass: 0.5 * t_closest on a seeded half of the hit rays (expected
    unoccluded), +inf elsewhere (expected occluded iff the ray has a closest hit)."""
    rng = np.random.default_rng(seed)
    half = rng.random(t_closest.shape[0]) < 0.5
    hit = prim != N.INVALID_ID
    tfar = np.full(t_closest.shape[0], np.inf, np.float32)
    sel = half & hit
    tfar[sel] = (0.5 * t_closest[sel]).astype(np.float32)
    expected = np.where(sel, 0, hit.astype(np.uint32)).astype(np.uint32)
    return tfar, expected
