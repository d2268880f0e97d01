"""Instanced scenes for the traversal tests: models with their (geometry, primitive) tables and the
composition of sgpu_scene_add_instance's two levels as the library performs it (the oracle's
rows_from_colmajor / affine_mul / affine_inverse), so the oracle's composed instances carry the
device's world-to-local transforms bit for bit."""
import numpy as np

from selas_amd import native as N

import oracle as O


rows_from_colmajor = O.rows_from_colmajor
affine_mul = O.affine_mul
affine_inverse = O.affine_inverse


def random_transform(rng, scale=(0.5, 2.0), spread=4.0):
    """A column-major 4 x 4 float transform: random rotation, anisotropic scale, translation."""
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    s = rng.uniform(*scale, 3)
    m = np.eye(4)
    m[:3, :3] = q @ np.diag(s)
    m[:3, 3] = rng.uniform(-spread, spread, 3)
    return np.ascontiguousarray(m.T.reshape(-1), np.float32)


class Model:
    """A scene of geometries: soups and quad meshes, with the (geom, prim) of each triangle."""

    def __init__(self, parts):
        # parts: list of ("soup", tri9) or ("quads", positions, quads)
        self.parts = parts
        tris, geom, prim = [], [], []
        for g, p in enumerate(parts):
            if p[0] == "soup":
                t9 = np.asarray(p[1], np.float32)
                tris.append(t9)
                geom += [g] * t9.shape[0]
                prim += list(range(t9.shape[0]))
            else:
                pos, quads = np.asarray(p[1], np.float32), np.asarray(p[2], np.uint32)
                t = []
                for qi, (a, b, c, d) in enumerate(quads):
                    t.append(np.concatenate([pos[a], pos[b], pos[d]]))
                    t.append(np.concatenate([pos[c], pos[d], pos[b]]))
                    prim += [qi, qi]
                    geom += [g, g]
                tris.append(np.array(t, np.float32))
        self.tri9 = np.ascontiguousarray(np.concatenate(tris), np.float32)
        self.geom = np.array(geom, np.uint32)
        self.prim = np.array(prim, np.uint32)
        self.quad = np.array([parts[g][0] == "quads" for g in geom], bool)
        self.bvh = O.CpuBvh(self.tri9)

    def device(self, ctx):
        s = N.Scene(ctx)
        for p in self.parts:
            if p[0] == "soup":
                s.add_soup(p[1])
            else:
                s.add_quads(p[1], p[2])
        s.commit()
        return s


def compose(top):
    """top: list of (child, colmajor) where child is a Model or a list of (Model, colmajor).
    Returns the composed instances in composition order: (w2l float32 (12,), Model, id0, id1)."""
    out = []
    for i, (child, m) in enumerate(top):
        l2w = rows_from_colmajor(m)
        if isinstance(child, Model):
            out.append((np.array(affine_inverse(l2w), np.float32), child, i, N.INVALID_ID))
        else:
            for j, (model, m2) in enumerate(child):
                c = affine_mul(l2w, rows_from_colmajor(m2))
                out.append((np.array(affine_inverse(c), np.float32), model, i, j))
    return out


def oracle_hits(composed, rays):
    """The oracle's closest hits and occlusion of an instanced scene, mapped to Embree's fields."""
    sc = O.InstancedScene([(w, m.bvh) for w, m, _, _ in composed])
    r = sc.closest(rays)
    n = r["t"].shape[0]
    out = {"t": r["t"], "u": r["u"].copy(), "v": r["v"].copy(), "prim_id": np.full(n, N.INVALID_ID, np.uint32),
           "geom_id": np.full(n, N.INVALID_ID, np.uint32), "inst_id0": np.full(n, N.INVALID_ID, np.uint32),
           "inst_id1": np.full(n, N.INVALID_ID, np.uint32), "ng": np.zeros((n, 3), np.float32)}
    for i in np.nonzero(r["prim"] != N.INVALID_ID)[0]:
        w, model, id0, id1 = composed[r["inst"][i]]
        k = r["prim"][i]
        out["geom_id"][i] = model.geom[k]
        out["prim_id"][i] = model.prim[k]
        out["inst_id0"][i] = id0
        out["inst_id1"][i] = id1
        if model.quad[k] and (k - np.searchsorted(model.geom, model.geom[k])) % 2 == 1:
            out["u"][i] = np.float32(1.0) - r["u"][i]
            out["v"][i] = np.float32(1.0) - r["v"][i]
        t = model.tri9[k]
        e1 = t[3:6] - t[0:3]
        e2 = t[6:9] - t[0:3]
        out["ng"][i] = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
    out["occluded"] = sc.occluded(rays)
    out["inst"] = r["inst"]
    return out
