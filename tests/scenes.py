"""Small scenes with known answers for the integrator tests.

plane_scene: one Lambert floor lit by a point light and a downward quad light, nothing else, so
every path of non-zero contribution is camera -> floor -> light and the image has a closed form:
    L_o(p) = (rho / pi) * (E_point(p) + E_quad(p))
    E_point = I cos(theta) / d^2
    E_quad  = Le * (1/2) * sum_i gamma_i n . unit(R_i x R_i+1)   (Lambert's polygon formula)
with rho = base_colour^2.2 (the material transform of CalculateSurfaceParams). Every estimator
that covers this path (NEE, light tracing, vertex merging, their MIS combination) must reproduce
it; a dropped or doubled MIS term shows up as a scale error.

sandwich_scene: a Lambert floor and ceiling facing each other (no corners within reach of the
merge radius), lit by a point light and a downward quad light between them; light bounces between
the planes, so every technique (NEE, light tracing, vertex connection, vertex merging) carries
multi-bounce paths. With Lambert BSDFs (pdfs consistent between sampling and evaluation) VCM and
the path tracer estimate the same integral.

cube_scene: a closed Lambert cube on a closed slab (every face wound outward, so every hit is on
a front face -- Selas's Lambert sampling does not check the side of the outgoing direction, which
would let light through the back of an open panel) under a point light, a quad light and a smooth
sky with a broad bright lobe (sky_map). With Lambert BSDFs VCM and the path tracer estimate the same
integral, and the image-based light adds paths of every technique: emission from the sky's disc,
NEE toward it, and camera paths that leave the scene.

dim_room: the closed room with albedos low enough (<= 0.1 after the 2.2 power) that paths longer
than 8 segments carry < 1e-7 of the energy, so the path-length limits of the techniques (which
differ between the path tracer and VCM) do not change the image.
"""
import numpy as np

from selas_amd import workloads as W

import oracle as O


def plane_scene(albedo=0.6):
    half = 40.0
    p0, p1, p2, p3 = (np.array(v, np.float64) for v in
                      ((-half, 0, -half), (-half, 0, half), (half, 0, half), (half, 0, -half)))
    tris = np.array([[p0, p1, p2], [p0, p2, p3]], np.float32).reshape(-1, 9)
    ng = np.cross(tris[0, 3:6] - tris[0, 0:3], tris[0, 6:9] - tris[0, 0:3])
    assert ng[1] > 0
    lights = np.zeros(2, W.LIGHT_DTYPE)
    lights[0] = (1, (0.6, 1.5, 0.4), (0, 0, 0), (0, 0, 0), (0, 0, 0), (3.0, 2.0, 1.0))
    lights[1] = (0, (-0.5, 1.2, 0.8), (0.0, -1.0, 0.0), (0.8, 0.0, 0.0), (0.0, 0.0, 0.6), (2.0, 3.0, 4.0))
    return {
        "geoms": [tris],
        "materials": np.stack([W.material(2, (albedo, albedo, albedo))]),
        "geom_material": np.zeros(1, np.uint32),
        "lights": lights,
        "background": np.zeros(3, np.float32),
        "camera": {"position": (0.0, 2.2, -2.5), "look_at": (0.0, 0.0, 0.6), "up": (0.0, 1.0, 0.0), "fov": 60.0},
    }


def _quad_irradiance(p, n, centre, ex, ez):
    """Irradiance factor at points p (k x 3) with normal n of a unit-radiance Lambertian quad."""
    c = np.asarray(centre, np.float64)
    ex, ez = np.asarray(ex, np.float64), np.asarray(ez, np.float64)
    verts = [c - 0.5 * ex - 0.5 * ez, c + 0.5 * ex - 0.5 * ez, c + 0.5 * ex + 0.5 * ez, c - 0.5 * ex + 0.5 * ez]
    total = np.zeros(p.shape[0])
    for i in range(4):
        r0 = verts[i][None, :] - p
        r1 = verts[(i + 1) % 4][None, :] - p
        r0 /= np.linalg.norm(r0, axis=1, keepdims=True)
        r1 /= np.linalg.norm(r1, axis=1, keepdims=True)
        cr = np.cross(r0, r1)
        gamma = np.arccos(np.clip(np.sum(r0 * r1, axis=1), -1, 1))
        total += gamma * (cr @ n) / np.linalg.norm(cr, axis=1)
    return 0.5 * np.abs(total)


def plane_image(sc, width, height, sub=8):
    """Pixel-area average of the closed-form radiance of plane_scene (height x width x 3)."""
    cam = sc["camera"]
    cam10 = list(cam["position"]) + list(cam["look_at"]) + list(cam["up"]) + [cam["fov"]]
    _, basis = O.camera_rays(cam10, width, height, np.array([[0, 0, 0, 1, 1, 0]], np.int32))
    bx, by, bz = (basis[3 * k:3 * k + 3].astype(np.float64) for k in range(3))
    pos = np.asarray(cam["position"], np.float64)
    off = (np.arange(sub) + 0.5) / sub
    xs = (np.arange(width)[:, None] + off[None, :]).reshape(-1)
    ys = (np.arange(height)[:, None] + off[None, :]).reshape(-1)
    X, Y = np.meshgrid(xs, ys)
    cx = X / width * 2.0 - 1.0
    cy = 1.0 - 2.0 * Y / height
    d = bz[None, None, :] - cx[..., None] * bx + cy[..., None] * by
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    t = -pos[1] / d[..., 1]
    p = (pos + t[..., None] * d).reshape(-1, 3)
    n = np.array([0.0, 1.0, 0.0])
    rho = float(sc["materials"][0][0]) ** 2.2
    out = np.zeros((p.shape[0], 3))
    for L in sc["lights"]:
        rad = np.asarray(L["radiance"], np.float64)
        if L["type"] == 1:
            v = np.asarray(L["position"], np.float64)[None, :] - p
            d2 = np.sum(v * v, axis=1)
            cos = np.maximum(v @ n, 0) / np.sqrt(d2)
            out += (cos / d2)[:, None] * rad[None, :]
        else:
            out += _quad_irradiance(p, n, L["position"], L["x"], L["z"])[:, None] * rad[None, :]
    out *= rho / np.pi
    out = out.reshape(height, sub, width, sub, 3).mean(axis=(1, 3))
    miss = (t <= 0).reshape(height, sub, width, sub).any(axis=(1, 3))
    assert not miss.any()
    return out


def sandwich_scene(albedo=0.7, height=1.6):
    sc = plane_scene(albedo)
    half = 40.0
    q = [(-half, height, -half), (half, height, -half), (half, height, half), (-half, height, half)]
    p0, p1, p2, p3 = (np.array(v, np.float64) for v in q)
    top = np.array([[p0, p1, p2], [p0, p2, p3]], np.float32).reshape(-1, 9)
    ng = np.cross(top[0, 3:6] - top[0, 0:3], top[0, 6:9] - top[0, 0:3])
    assert ng[1] < 0
    sc["geoms"] = sc["geoms"] + [top]
    sc["materials"] = np.stack([W.material(2, (albedo, 0.8 * albedo, 0.6 * albedo))] * 2)
    sc["geom_material"] = np.zeros(2, np.uint32)
    sc["camera"] = {"position": (0.0, 1.0, -2.0), "look_at": (0.0, 0.2, 0.6), "up": (0.0, 1.0, 0.0), "fov": 70.0}
    return sc


def dim_room(seed=1):
    sc = W.room_scene(seed)
    mats = sc["materials"].copy()
    mats[:, 0:3] = np.minimum(mats[:, 0:3], 0.35)
    sc["materials"] = mats
    return sc


def sky_map(h, w):
    """An equirectangular sky (h x w x 3): brighter toward +y, with a broad bright lobe."""
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


def closed_box(lo, hi):
    """The 12 triangles (9 floats each) of an axis-aligned box, wound so every normal points out."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    c = [np.array([hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]])
         for i in range(8)]
    centre = 0.5 * (lo + hi)
    tris = []
    for q in ((0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)):
        a, b, cc, d = (c[k] for k in q)
        for t in ((a, b, cc), (a, cc, d)):
            if np.dot(np.cross(t[1] - t[0], t[2] - t[0]), t[0] - centre) < 0:
                t = (t[0], t[2], t[1])
            tris.append(np.concatenate(t))
    return np.array(tris, np.float32)


def cube_scene(albedo=0.7, lights=True, ibl=True, sky=(64, 128)):
    sc = plane_scene(albedo)
    sc["geoms"] = [closed_box((-40, -0.2, -40), (40, 0.0, 40)), closed_box((-0.5, 0.0, 0.3), (0.5, 0.9, 1.3))]
    sc["materials"] = np.stack([W.material(2, (albedo, 0.8 * albedo, 0.6 * albedo)), W.material(2, (0.5, 0.7, 0.6))])
    sc["geom_material"] = np.array([0, 1], np.uint32)
    lt = sc["lights"].copy()
    lt[1]["position"] = (-0.3, 1.6, 0.2)
    sc["lights"] = lt if lights else lt[:0]
    if ibl:
        sc["ibl"] = {"rgb": sky_map(*sky), "rotation": 0.7, "exposure": 0.8}
    return sc


def uniform_sky_plane(albedo=0.6, h=64, w=128, radiance=1.0, rotation=0.3):
    """plane_scene lit only by a uniform sky of the given radiance: every camera pixel sees the floor,
    whose radiance is rho x radiance (a Lambert reflector under a uniform hemisphere)."""
    sc = plane_scene(albedo)
    sc["lights"] = sc["lights"][:0]
    sc["ibl"] = {"rgb": np.full((h, w, 3), radiance, np.float32), "rotation": rotation, "exposure": 1.0}
    return sc


def small_sky_plane(half=1.0, radiance=1.0):
    """A 2 half x 2 half Lambert floor under a uniform sky, seen from close by (every pixel on the
    floor, radiance rho x radiance). Its bounding sphere is small, so light tracing and merging from
    the sky's emission disc carry a large share of the image and every MIS term of sky emission
    (the first segment without d^2, dVC and dVM of an intersectable light) moves it by 9-11 %."""
    sc = plane_scene()
    p = [np.array(v, np.float64) for v in ((-half, 0, -half), (-half, 0, half), (half, 0, half), (half, 0, -half))]
    sc["geoms"] = [np.array([[p[0], p[1], p[2]], [p[0], p[2], p[3]]], np.float32).reshape(-1, 9)]
    sc["lights"] = sc["lights"][:0]
    sc["ibl"] = {"rgb": np.full((64, 128, 3), radiance, np.float32), "rotation": 0.3, "exposure": 1.0}
    sc["camera"] = {"position": (0.0, 1.2, -0.3), "look_at": (0.0, 0.0, 0.1), "up": (0.0, 1.0, 0.0), "fov": 50.0}
    return sc


def _m4(ry=0.0, scale=1.0, t=(0.0, 0.0, 0.0), rx=0.0):
    """4 x 4 local-to-world: rotation about x, then about y, uniform scale, translation (acts on
    column vectors)."""
    c, s_ = np.cos(ry), np.sin(ry)
    cx, sx = np.cos(rx), np.sin(rx)
    r = np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]]) @ np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    m = np.eye(4)
    m[:3, :3] = r * scale
    m[:3, 3] = t
    return m


def instanced_scene(ibl=False):
    """Three placed scenes -- a closed slab (Lambert), a closed crate (Disney) and a textured
    displaced panel (Disney with a per-face texture) -- arranged with two instance levels: a
    subscene (crate + panel above it) placed twice, the slab and a third crate placed directly.
    Lit by a quad and a point light (and, with `ibl`, a sky)."""
    floor = {"geoms": [closed_box((-6, -0.2, -6), (6, 0.0, 6))],
             "materials": np.stack([W.material(2, (0.6, 0.6, 0.55))]), "geom_material": np.zeros(1, np.uint32)}
    crate = {"geoms": [closed_box((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5))],
             "materials": np.stack([W.material(1, (0.7, 0.4, 0.2), 0.0, 0.5, 1.5)]),
             "geom_material": np.zeros(1, np.uint32)}
    # the workload's textured panel kind, flat and facing +y in its own space
    v = W.vcm_scene(seed=11, n_meshes=1, nu=8, nv=4)
    g = dict(v["geoms"][0])
    X, Z = np.meshgrid(np.linspace(0, 1, 9), np.linspace(0, 0.5, 5), indexing="ij")
    g["positions"] = np.stack([X - 0.5, np.zeros_like(X), Z - 0.25], -1).reshape(-1, 3).astype(np.float32)
    panel = {"geoms": [g], "materials": v["materials"][:1].copy(), "geom_material": np.zeros(1, np.uint32)}
    lights = np.zeros(2, W.LIGHT_DTYPE)
    lights[0] = (0, (0.0, 4.0, 0.0), (0.0, -1.0, 0.0), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0), (6.0, 6.0, 6.0))
    lights[1] = (1, (2.0, 3.0, -1.0), (0, 0, 0), (0, 0, 0), (0, 0, 0), (10.0, 9.0, 8.0))
    sc = {
        "models": [floor, crate, panel],
        "subscenes": [[(1, _m4(0.3)), (2, _m4(0.2, 1.2, (0.0, 1.5, 0.0), rx=-1.0))]],
        "instances": [("model", 0, _m4()), ("subscene", 0, _m4(0.0, 1.0, (-1.5, 0.0, 0.5))),
                      ("subscene", 0, _m4(1.0, 0.9, (1.2, 0.0, -0.3))), ("model", 1, _m4(0.7, 0.7, (0.0, 0.0, 2.0)))],
        "lights": lights,
        "background": np.zeros(3, np.float32),
        "camera": {"position": (0.0, 3.0, -5.0), "look_at": (0.0, 0.6, 0.5), "up": (0.0, 1.0, 0.0), "fov": 60.0},
    }
    if ibl:
        sc["ibl"] = {"rgb": sky_map(32, 64), "rotation": 0.4, "exposure": 0.5}
    return sc


def flatten_instanced(sc):
    """The scene of geometries an instanced scene dict describes: every composed instance's
    geometries moved to world space (in double, rounded to float), with their materials."""
    geoms, mats, gm = [], [], []
    for midx, l2w, _, _ in O.compose_instances(sc):
        m = np.array(l2w, np.float64).reshape(3, 4)
        model = sc["models"][midx]
        base = sum(x.shape[0] for x in mats)
        mats.append(np.asarray(model["materials"], np.float32))
        for gi, g in enumerate(model["geoms"]):
            gm.append(base + int(model["geom_material"][gi]))
            if isinstance(g, dict):
                p = np.asarray(g["positions"], np.float64)
                geoms.append(dict(g, positions=(p @ m[:, :3].T + m[:, 3]).astype(np.float32)))
            else:
                p = np.asarray(g, np.float64).reshape(-1, 3)
                geoms.append((p @ m[:, :3].T + m[:, 3]).astype(np.float32).reshape(-1, 9))
    out = {k: v for k, v in sc.items() if k not in ("models", "subscenes", "instances")}
    out.update(geoms=geoms, materials=np.concatenate(mats), geom_material=np.array(gm, np.uint32))
    return out
