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
