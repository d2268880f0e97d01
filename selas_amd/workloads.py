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
