"""The oracle's round B-spline curves (oracle/orc_curve.c), pinned by closed forms.

Selas builds hair and grass as RTC_GEOMETRY_TYPE_ROUND_BSPLINE_CURVE (ModelResource.cpp:300-330).
Embree's library is not in the reference, so the pins are mathematical: the surface is the sphere
swept along the uniform cubic B-spline. Four equal control points collapse the segment to one ball
(ray-sphere closed form); equally spaced collinear points with equal radii make it the capsule from
the second to the third point (the basis reproduces linear functions: P(u) = p1 + u (p2 - p1));
the dense restatement (the definition, 8193 parameters and golden-section refinement) and the
piecewise one (the device's algorithm) agree on random curved segments away from grazing rays."""
import numpy as np
import pytest

import oracle as O


def _sphere_entry(c, r, o, d):
    oc = o - c
    a = d @ d
    hb = d @ oc
    disc = hb * hb - a * (oc @ oc - r * r)
    return None if disc < 0 else (-hb - np.sqrt(disc)) / a


def _capsule_entry(p, q, r, o, d):
    """Entry of o + t d into the capsule of radius r around segment pq (a union of convex parts:
    the first entry is the least entry into the cylinder body or an end ball)."""
    best = [t for t in (_sphere_entry(p, r, o, d), _sphere_entry(q, r, o, d)) if t is not None]
    ax = (q - p) / np.linalg.norm(q - p)
    L = np.linalg.norm(q - p)
    dp = d - (d @ ax) * ax
    op = (o - p) - ((o - p) @ ax) * ax
    a = dp @ dp
    if a > 0:
        hb = dp @ op
        disc = hb * hb - a * (op @ op - r * r)
        if disc >= 0:
            t = (-hb - np.sqrt(disc)) / a
            s = (o + t * d - p) @ ax
            if 0 <= s <= L:
                best.append(t)
    return min(best) if best else None


def _rays_at(rng, centre, spread, n, dist=8.0):
    """n rays from a sphere of radius `dist` around centre towards points within `spread` of it."""
    o = rng.normal(size=(n, 3))
    o = centre + dist * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = centre + spread * (rng.random((n, 3)) * 2 - 1)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= rng.uniform(0.5, 2.0, (n, 1))  # unnormalised directions: t is the ray's own parameter
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("dense", [True, False])
def test_collapsed_segment_is_a_sphere(dense):
    rng = np.random.default_rng(1)
    c, r = np.array([0.3, -0.2, 0.5]), 0.7
    cp = np.array([[*c, r]] * 4, np.float32)
    o, d = _rays_at(rng, c, 1.0, 400)
    hits = 0
    for k in range(len(o)):
        ref = _sphere_entry(c.astype(np.float32).astype(np.float64), np.float64(np.float32(r)), o[k].astype(np.float64),
                            d[k].astype(np.float64))
        got = O.curve_hit(cp, o[k], d[k], dense=dense)
        if ref is None:
            assert got is None
            continue
        hits += 1
        t, u, ng = got
        assert abs(t - ref) <= 1e-9 * max(1.0, abs(ref)), (k, t, ref)
        x = o[k] + t * d[k].astype(np.float64)
        assert np.allclose(ng, x - c.astype(np.float32), atol=1e-6)
    assert hits > 100


@pytest.mark.parametrize("dense", [True, False])
def test_collinear_segment_is_a_capsule(dense):
    rng = np.random.default_rng(2)
    a, e, r = np.array([-1.0, 0.2, 0.1]), np.array([0.8, 0.3, -0.1]), 0.25
    cp = np.array([[*(a + k * e), r] for k in range(4)], np.float32)
    p, q = cp[1, :3].astype(np.float64), cp[2, :3].astype(np.float64)
    o, d = _rays_at(rng, 0.5 * (p + q), 0.5, 400)
    hits = 0
    for k in range(len(o)):
        ref = _capsule_entry(p, q, np.float64(np.float32(r)), o[k].astype(np.float64), d[k].astype(np.float64))
        got = O.curve_hit(cp, o[k], d[k], dense=dense)
        if ref is None:
            assert got is None or abs(O.curve_clearance(cp, o[k], d[k])) < 1e-6
            continue
        hits += 1
        assert got is not None
        assert abs(got[0] - ref) <= 1e-6 * max(1.0, abs(ref)), (k, got[0], ref)
    assert hits > 100


def test_parameter_is_the_b_spline_parameter():
    # points (k, 0, 0), radius 0.1: P(u) = (1 + u, 0, 0), so a ray down onto x = 1.3 hits at u = 0.3
    cp = np.array([[k, 0, 0, 0.1] for k in range(4)], np.float32)
    for dense in (True, False):
        t, u, ng = O.curve_hit(cp, [1.3, 5.0, 0.0], [0.0, -1.0, 0.0], dense=dense)
        assert abs(t - 4.9) < 1e-6 and abs(u - 0.3) < 1e-6
        assert np.allclose(ng / np.linalg.norm(ng), [0, 1, 0], atol=1e-6)
    # the end balls are part of the surface: a ray along the axis enters the ball at p1
    t, u, _ = O.curve_hit(cp, [-3.0, 0.0, 0.0], [1.0, 0.0, 0.0])
    assert abs(t - 3.9) < 1e-6 and u < 1e-6


def _bspline(cp, u):
    u = np.asarray(u, np.float64)[:, None]
    s = 1 - u
    return (s ** 3 * cp[0] + (3 * u ** 3 - 6 * u ** 2 + 4) * cp[1] + (-3 * u ** 3 + 3 * u ** 2 + 3 * u + 1) * cp[2]
            + u ** 3 * cp[3]) / 6


def _rays_near_curve(rng, cp, n, dist=8.0):
    """n rays from `dist` away towards points within about two radii of the segment's curve."""
    tgt = _bspline(cp.astype(np.float64), rng.random(n))
    tgt = tgt[:, :3] + rng.normal(scale=1.5 * tgt[:, 3:4], size=(n, 3))
    o = rng.normal(size=(n, 3))
    o = tgt + dist * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = (tgt - o) * rng.uniform(0.05, 0.3, (n, 1))
    return o.astype(np.float32), d.astype(np.float32)


def _random_segment(rng, radius):
    base = rng.normal(size=3)
    pts = np.cumsum(rng.normal(scale=0.5, size=(4, 3)), axis=0) + base
    rad = radius * rng.uniform(0.5, 1.5, 4)
    return np.concatenate([pts, rad[:, None]], 1).astype(np.float32)


@pytest.mark.parametrize("radius", [0.05, 0.3])
def test_piecewise_agrees_with_the_definition(radius):
    """The device's algorithm (16 capsule-bounded pieces, golden-section search) against the dense
    restatement of the definition, rays that do not graze the surface."""
    rng = np.random.default_rng(3)
    compared = hits = 0
    for s in range(40):
        cp = _random_segment(rng, radius)
        o, d = _rays_near_curve(rng, cp, 40)
        for k in range(len(o)):
            if abs(O.curve_clearance(cp, o[k], d[k])) < 1e-3:
                continue
            compared += 1
            a = O.curve_hit(cp, o[k], d[k], dense=True)
            b = O.curve_hit(cp, o[k], d[k])
            assert (a is None) == (b is None), (s, k)
            if a is not None:
                hits += 1
                assert abs(a[0] - b[0]) <= 1e-7 * max(1.0, a[0]), (s, k, a[0], b[0])
                na, nb = a[2] / np.linalg.norm(a[2]), b[2] / np.linalg.norm(b[2])
                assert na @ nb > 1 - 1e-6
                assert O.curve_hit(cp, o[k], d[k], any_hit=True) is not None
    assert compared > 1000 and hits > 200


def test_batch_forms_and_interval():
    rng = np.random.default_rng(4)
    cps = np.stack([_random_segment(rng, 0.1) for _ in range(30)])
    o, d = _rays_at(rng, cps[:, :, :3].reshape(-1, 3).mean(0), 2.0, 500, dist=12.0)
    rays = {"org_x": o[:, 0], "org_y": o[:, 1], "org_z": o[:, 2], "dir_x": d[:, 0], "dir_y": d[:, 1], "dir_z": d[:, 2],
            "tnear": np.zeros(len(o), np.float32), "tfar": np.full(len(o), 1e30, np.float32)}
    h = O.curves_closest(cps, rays)
    hd = O.curves_closest(cps, rays, dense=True)
    occ = O.curves_occluded(cps, rays)
    assert (occ == (h["prim"] != 0xFFFFFFFF)).all()
    hit = h["prim"] != 0xFFFFFFFF
    assert hit.sum() > 50
    # per ray, the closest segment's t agrees with the dense closest (prim may differ only on ties)
    both = hit & (hd["prim"] != 0xFFFFFFFF)
    assert np.allclose(h["t"][both], hd["t"][both], rtol=1e-6)
    # an interval ending before the hit sees nothing; one starting after it sees a farther hit or none
    r2 = dict(rays, tfar=np.where(hit, h["t"] * 0.999, rays["tfar"]).astype(np.float32))
    assert (O.curves_occluded(cps, r2)[hit] == 0).all()
    r3 = dict(rays, tnear=np.where(hit, h["t"] * 1.001, 0).astype(np.float32))
    h3 = O.curves_closest(cps, r3)
    assert (h3["t"][hit & (h3["prim"] != 0xFFFFFFFF)] > h["t"][hit & (h3["prim"] != 0xFFFFFFFF)]).all()
