"""Minimal writer of Ptex files (.ptx) for test fixtures: quad meshes, 8-bit RGB, one resolution
for all faces, base level only, every face zip-encoded.

The layout is the one the Ptex reader parses (Middleware/ptex/Include/PtexIO.h:44-99 and
PtexReader.cpp:155-200, 595-700): header, extended header, zipped face-info block, zipped
per-face constant values, level-info block, level data (zipped face-data headers, then one zipped
block per face holding the texels plane by plane), then an empty meta-data block and the 8-byte
compatibility barrier. Written from that format description; the reference ships no writer.

Also builds the adjacency of the test meshes (Ptex edge ids: 0 bottom v=0, 1 right u=1, 2 top v=1,
3 left u=0; an edge's adjacent edge is the neighbour's edge id of the shared edge).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

MAGIC = ord("P") | (ord("t") << 8) | (ord("e") << 16) | (ord("x") << 24)
MT_QUAD = 1
DT_UINT8 = 0
ENC_ZIPPED = 1


def write_ptx(path, res_log2, texels, adjfaces, adjedges):
    """texels: (nfaces, res, res, 3) uint8, row index = v texel, column = u texel.
    adjfaces: (nfaces, 4) int (-1 = border); adjedges: (nfaces, 4) int in 0..3."""
    texels = np.ascontiguousarray(texels, np.uint8)
    nfaces, res = texels.shape[0], 1 << res_log2
    assert texels.shape == (nfaces, res, res, 3)
    # face info: res (ulog2, vlog2), adjedges (2 bits per edge), flags, adjfaces[4]
    fi = b"".join(
        struct.pack("<bbBB4i", res_log2, res_log2,
                    sum((int(adjedges[f][e]) & 3) << (2 * e) for e in range(4)), 0,
                    *[int(a) for a in adjfaces[f]])
        for f in range(nfaces))
    fi_z = zlib.compress(fi)
    const = np.round(texels.reshape(nfaces, -1, 3).mean(axis=1)).astype(np.uint8)
    const_z = zlib.compress(const.tobytes())
    blocks = [zlib.compress(np.ascontiguousarray(texels[f].transpose(2, 0, 1)).tobytes()) for f in range(nfaces)]
    fdh = struct.pack("<%dI" % nfaces, *[(len(b) & 0x3FFFFFFF) | (ENC_ZIPPED << 30) for b in blocks])
    fdh_z = zlib.compress(fdh)
    level = fdh_z + b"".join(blocks)
    levelinfo = struct.pack("<QII", len(level), len(fdh_z), nfaces)
    ext = struct.pack("<HHHHIIQQQ", 0, 0, 0, 0, 0, 0, 0, 0, 0)  # clamp borders, no edge filter, no edits
    header = struct.pack("<IIIIiHHIIIIIIQII", MAGIC, 1, MT_QUAD, DT_UINT8, -1, 3, 1, nfaces, len(ext), len(fi_z),
                         len(const_z), len(levelinfo), 4, len(level), 0, 0)
    assert len(header) == 64 and len(ext) == 40
    with open(path, "wb") as f:
        f.write(header + ext + fi_z + const_z + levelinfo + level + struct.pack("<Q", 0))


MT_TRIANGLE = 0
ENC_CONSTANT, ENC_DIFFZIPPED, ENC_TILED = 0, 2, 3
FLAG_CONSTANT = 1
_NP = {0: np.uint8, 1: np.uint16, 2: np.float16, 3: np.float32}


def _face_block(px, enc, dt):
    """One face's (or tile's) block: (encoding, bytes). px: (v, u, nchan) of dtype dt."""
    v, u, nc = px.shape
    if enc == "const":
        return ENC_CONSTANT, px[0, 0].astype(dt).tobytes()
    planes = np.ascontiguousarray(px.transpose(2, 0, 1)).astype(dt)
    if enc == "diff":
        a = planes.reshape(-1)
        d = np.empty_like(a)
        d[0] = a[0]
        d[1:] = a[1:] - a[:-1]  # encodeDifference (wrapping for unsigned types)
        return ENC_DIFFZIPPED, zlib.compress(d.tobytes())
    return ENC_ZIPPED, zlib.compress(planes.tobytes())


def write_ptx_general(path, faces, res_log2, adjfaces, adjedges, mesh_type=MT_QUAD, data_type=DT_UINT8,
                      encodings=None, flags=None, alpha=-1):
    """A Ptex file with per-face resolutions, encodings and flags, any channel count and data type.
    faces: list of (v res, u res, nchan) arrays (constant-flagged faces: any shape, first pixel
    used); res_log2: (n, 2) (ulog2, vlog2); encodings: per face "zip", "diff", "const" or
    ("tiled", tile ulog2, tile vlog2, [per-tile encodings]); flags: per-face FaceInfo flags."""
    n = len(faces)
    dt = _NP[data_type]
    nchan = faces[0].shape[2]
    encodings = encodings or ["zip"] * n
    flags = np.zeros(n, np.uint8) if flags is None else np.asarray(flags, np.uint8)
    fi = b"".join(
        struct.pack("<bbBB4i", int(res_log2[f][0]), int(res_log2[f][1]),
                    sum((int(adjedges[f][e]) & 3) << (2 * e) for e in range(4)), int(flags[f]),
                    *[int(a) for a in adjfaces[f]])
        for f in range(n))
    fi_z = zlib.compress(fi)
    const = np.stack([np.asarray(a, np.float64).reshape(-1, nchan).mean(0) for a in faces])
    const = (np.round(const) if np.dtype(dt).kind == "u" else const).astype(dt)
    for f in range(n):
        if flags[f] & FLAG_CONSTANT:
            const[f] = faces[f][0, 0]
    const_z = zlib.compress(const.tobytes())
    blocks, hdrs = [], []
    for f in range(n):
        if flags[f] & FLAG_CONSTANT or (res_log2[f][0] == 0 and res_log2[f][1] == 0):
            enc, b = _face_block(faces[f][:1, :1], "const", dt)
        elif isinstance(encodings[f], tuple):
            _, tul, tvl, tile_encs = encodings[f]
            tu, tv = 1 << tul, 1 << tvl
            px = faces[f]
            ntu, ntv = px.shape[1] // tu, px.shape[0] // tv
            tblocks, thdrs = [], []
            for i in range(ntv):
                for j in range(ntu):
                    te, tb = _face_block(px[i * tv:(i + 1) * tv, j * tu:(j + 1) * tu], tile_encs[i * ntu + j], dt)
                    tblocks.append(tb)
                    thdrs.append((len(tb) & 0x3FFFFFFF) | (te << 30))
            th_z = zlib.compress(struct.pack("<%dI" % len(thdrs), *thdrs))
            b = struct.pack("<bbI", tul, tvl, len(th_z)) + th_z + b"".join(tblocks)
            enc = ENC_TILED
        else:
            enc, b = _face_block(faces[f], encodings[f], dt)
        blocks.append(b)
        hdrs.append((len(b) & 0x3FFFFFFF) | (enc << 30))
    fdh_z = zlib.compress(struct.pack("<%dI" % n, *hdrs))
    level = fdh_z + b"".join(blocks)
    levelinfo = struct.pack("<QII", len(level), len(fdh_z), n)
    ext = struct.pack("<HHHHIIQQQ", 0, 0, 0, 0, 0, 0, 0, 0, 0)
    header = struct.pack("<IIIIiHHIIIIIIQII", MAGIC, 1, mesh_type, data_type, alpha, nchan, 1, n, len(ext),
                         len(fi_z), len(const_z), len(levelinfo), 4, len(level), 0, 0)
    with open(path, "wb") as fh:
        fh.write(header + ext + fi_z + const_z + levelinfo + level + struct.pack("<Q", 0))


def grid_adjacency(nu, nv):
    """Quad grid, face (i, j) = j * nu + i with u along i and v along j (the face order and uv
    orientation of quad_grid_mesh in selas_amd/workloads.py)."""
    n = nu * nv
    af = -np.ones((n, 4), np.int32)
    ae = np.zeros((n, 4), np.int32)
    for j in range(nv):
        for i in range(nu):
            f = j * nu + i
            if j > 0:
                af[f, 0], ae[f, 0] = f - nu, 2
            if i < nu - 1:
                af[f, 1], ae[f, 1] = f + 1, 3
            if j < nv - 1:
                af[f, 2], ae[f, 2] = f + nu, 0
            if i > 0:
                af[f, 3], ae[f, 3] = f - 1, 1
    return af, ae


def quad_mesh_adjacency(quads):
    """Adjacency of a quad mesh given as (n, 4) vertex indices, counter-clockwise (v0 v1 v2 v3 with
    u along v0->v1 and v along v0->v3): edge e runs from corner e to corner e+1."""
    quads = np.asarray(quads)
    n = quads.shape[0]
    edges = {}
    for f in range(n):
        for e in range(4):
            a, b = int(quads[f, e]), int(quads[f, (e + 1) % 4])
            edges[(a, b)] = (f, e)
    af = -np.ones((n, 4), np.int32)
    ae = np.zeros((n, 4), np.int32)
    for f in range(n):
        for e in range(4):
            a, b = int(quads[f, e]), int(quads[f, (e + 1) % 4])
            if (b, a) in edges:
                af[f, e], ae[f, e] = edges[(b, a)]
    return af, ae


def tri_mesh_adjacency(tris):
    """Adjacency of a triangle mesh given as (n, 3) vertex indices, counter-clockwise: Ptex's
    triangle edge e runs from corner e to corner e+1 (edge 0: v = 0, edge 1: w = 0, edge 2: u = 0);
    the fourth slot is unused (-1, 0)."""
    tris = np.asarray(tris)
    n = tris.shape[0]
    edges = {}
    for f in range(n):
        for e in range(3):
            edges[(int(tris[f, e]), int(tris[f, (e + 1) % 3]))] = (f, e)
    af = -np.ones((n, 4), np.int32)
    ae = np.zeros((n, 4), np.int32)
    for f in range(n):
        for e in range(3):
            a, b = int(tris[f, e]), int(tris[f, (e + 1) % 3])
            if (b, a) in edges:
                af[f, e], ae[f, e] = edges[(b, a)]
    return af, ae


def tri_grid(nu, nv):
    """A quad grid split into triangles (a, b, c), (a, c, d) per quad."""
    tris = []
    for j in range(nv):
        for i in range(nu):
            a = j * (nu + 1) + i
            b, c, d = a + 1, a + nu + 2, a + nu + 1
            tris += [[a, b, c], [a, c, d]]
    return np.array(tris)


def tetra_tris():
    return np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])


def icosa_tris():
    return np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                     [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                     [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])


def fan_tris(k=7):
    """k triangles around vertex 0 (an open fan: a border on both sides)."""
    return np.array([[0, i, i + 1] for i in range(1, k + 1)])


def cube_quads():
    """A cube as 6 quads with outward orientation and assorted rotations between neighbours (every
    corner has valence 3)."""
    quads = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [1, 2, 6, 5], [0, 4, 7, 3]])
    return quads


def star_quads(k=5):
    """k quads around a centre vertex (valence k) with a border: centre 0, ring vertices 1..2k."""
    quads = []
    for i in range(k):
        a = 1 + 2 * i
        b = 1 + 2 * i + 1
        c = 1 + (2 * i + 2) % (2 * k)
        quads.append([0, a, b, c])
    return np.array(quads)
