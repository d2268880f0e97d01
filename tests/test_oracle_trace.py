"""Pins the ray-cast oracle (oracle/orc_trace.c) before it is trusted as the checker.

The reference's ray casts go through Embree 3.2, whose library is not available; these tests pin
the oracle's contract (closest hit in (tnear, tfar], u/v = weights of v1/v2, ties to the smaller
index, occluded iff any hit) with closed forms, a double-precision brute force, and a planted
mistake that the checks must catch.
"""
import numpy as np
import pytest

import oracle as O
from selas_amd import native as N


def tri(v0, v1, v2):
    return np.array([list(v0) + list(v1) + list(v2)], np.float32)


def one_ray(o, d, tnear=0.0, tfar=np.inf):
    r = {"org_x": [o[0]], "org_y": [o[1]], "org_z": [o[2]], "dir_x": [d[0]], "dir_y": [d[1]], "dir_z": [d[2]],
         "tnear": [tnear], "tfar": [tfar]}
    return {k: np.asarray(v, np.float32) for k, v in r.items()}


def test_closed_form_hit_t_u_v():
    # triangle in the z=2 plane; the ray from the origin along +z hits at t=2, (u, v) = (0.25, 0.5)
    t9 = tri((0, 0, 2), (4, 0, 2), (0, 4, 2))
    hit, t, u, v = O.tri_test(t9[0, 0:3], t9[0, 3:6] - t9[0, 0:3], t9[0, 6:9] - t9[0, 0:3], [1, 2, 0], [0, 0, 1], 0.0)
    assert hit and t == 2.0 and u == 0.25 and v == 0.5
    # back side (det < 0) is hit too: Selas uses two-sided triangles
    hit, t, u, v = O.tri_test(t9[0, 0:3], t9[0, 3:6] - t9[0, 0:3], t9[0, 6:9] - t9[0, 0:3], [1, 2, 4], [0, 0, -1], 0.0)
    assert hit and t == 2.0 and u == 0.25 and v == 0.5


def test_interval_is_open_at_tnear_closed_at_tfar():
    t9 = tri((0, 0, 2), (4, 0, 2), (0, 4, 2))
    h = O.closest_brute_f32(t9, one_ray([1, 1, 0], [0, 0, 1], 0.0, 2.0))
    assert h["prim"][0] == 0 and h["t"][0] == 2.0
    h = O.closest_brute_f32(t9, one_ray([1, 1, 0], [0, 0, 1], 2.0, 5.0))
    assert h["prim"][0] == N.INVALID_ID and h["t"][0] == 5.0  # t == tnear is rejected; tfar is kept
    h = O.closest_brute_f32(t9, one_ray([1, 1, 0], [0, 0, 1], 0.0, 1.999))
    assert h["prim"][0] == N.INVALID_ID
    assert O.occluded_brute_f32(t9, one_ray([1, 1, 0], [0, 0, 1], 0.0, 2.0))[0] == 1
    assert O.occluded_brute_f32(t9, one_ray([1, 1, 0], [0, 0, 1], 2.0, 5.0))[0] == 0


def test_parallel_and_degenerate_miss():
    t9 = tri((0, 0, 2), (4, 0, 2), (0, 4, 2))
    assert O.closest_brute_f32(t9, one_ray([1, 1, 0], [1, 0, 0]))["prim"][0] == N.INVALID_ID
    deg = tri((0, 0, 2), (1, 1, 2), (2, 2, 2))  # zero area: det == 0 for every ray
    assert O.closest_brute_f32(deg, one_ray([1, 1, 0], [0, 0, 1]))["prim"][0] == N.INVALID_ID


def test_exact_tie_goes_to_smaller_index():
    a = tri((0, 0, 2), (4, 0, 2), (0, 4, 2))
    t9 = np.concatenate([a, a, a])
    rays = one_ray([1, 1, 0], [0, 0, 1])
    assert O.closest_brute_f32(t9, rays)["prim"][0] == 0
    assert O.CpuBvh(t9).closest(rays)["prim"][0] == 0


def dense_soup(seed, n_clusters=64, per=64):
    # larger triangles than the benchmark soup so most rays hit something
    return N.gen_cluster_soup(seed, n_clusters, per, 0.08, 2e-2, 2e-1)


@pytest.mark.parametrize("seed", [1, 2])
def test_f32_arithmetic_agrees_with_f64_brute_force(seed):
    t9 = dense_soup(seed)
    cam = N.gen_camera_rays((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), (0, 1, 0), 50.0, 48, 48, 1)
    bnc = N.gen_bounce_rays(seed, t9, 2304)
    for rays in (cam, bnc):
        f32 = O.closest_brute_f32(t9, rays)
        f64 = O.closest_brute_f64(t9, rays)
        hit = f64["prim"] != N.INVALID_ID
        assert hit.mean() > 0.3
        differ = f32["prim"] != f64["prim"]
        # a disagreement is only allowed for a near tie in t or a graze of a triangle edge
        with np.errstate(invalid="ignore"):
            near_tie = (f64["t2"] - f64["t"]) <= 1e-5 * np.maximum(np.abs(f64["t"]), 1e-3)
        graze = f64["margin"] < 1e-5
        assert not (differ & ~near_tie & ~graze).any()
        both = hit & ~differ
        # FP32 rounding of unit-scale coordinates: a few ulp of the hit distance or of 1
        err = np.abs(f32["t"][both] - f64["t"][both])
        assert (err <= 4e-6 * np.maximum(1.0, f64["t"][both])).all()


def test_planted_uv_swap_is_caught():
    # The same test with e1 and e2 exchanged (a transposed operand) swaps u and v and flips the
    # sign of det; the closed form catches it.
    t9 = tri((0, 0, 2), (4, 0, 2), (0, 4, 2))
    v0, e1, e2 = t9[0, 0:3], t9[0, 3:6] - t9[0, 0:3], t9[0, 6:9] - t9[0, 0:3]
    good = O.tri_test(v0, e1, e2, [1, 2, 0], [0, 0, 1], 0.0)
    bad = O.tri_test(v0, e2, e1, [1, 2, 0], [0, 0, 1], 0.0)
    assert good[2:] == (0.25, 0.5)
    assert bad[2:] != good[2:]


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_bvh_equals_brute_force_bit_exact(seed):
    t9 = dense_soup(seed, 64, 128)
    # add exact duplicates and degenerate triangles: ties and det == 0 paths
    t9 = np.concatenate([t9, t9[:64], np.repeat(t9[:8, 0:3], 3, axis=1)])
    rays = N.gen_bounce_rays(seed, t9, 4096)
    cam = N.gen_camera_rays((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), (0, 1, 0), 50.0, 64, 32, 2)
    bvh = O.CpuBvh(t9)
    for r in (rays, cam):
        a = O.closest_brute_f32(t9, r)
        b = bvh.closest(r)
        for k in ("prim", "t", "u", "v"):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
        assert np.array_equal(O.occluded_brute_f32(t9, r), bvh.occluded(r))
