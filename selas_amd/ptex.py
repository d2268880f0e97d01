"""Ptex files (.ptx) read into the per-face textures sgpu_scene_set_ptex takes.

Selas opens its base-colour textures with PtexTexture::open (TextureCache::LoadTexturePtex,
Source/Core/TextureLib/TextureCache.cpp) and filters them with PtexFilter (SurfaceParameters.cpp:
196-203). This reader restates the parts of the Ptex 2.x reader that produce a face's full-resolution
texels (Middleware/ptex/Include/PtexReader.cpp:140-200 open, 290-330 face info / level info /
constant data, 560-720 zip blocks, levels, face data and tiles; PtexIO.h:44-101 the header records;
PtexUtils.cpp:162-195 interleave, 255-271 decodeDifference):

* header (64 bytes), extended header (its size from the header), then the zipped face-info block
  (20 bytes per face: ulog2, vlog2, packed adjacent edges, flags, 4 adjacent faces), the zipped
  constant-value block (one pixel per face), the level-info records and the level data;
* level 0 opens with the zipped face-data headers (30-bit size, 2-bit encoding per face), followed
  by each face's block: constant (one pixel), zipped or difference-zipped planes (channel planes of
  u-rows; difference coding is a running sum over the whole unzipped buffer), or tiled (tile
  resolution, zipped tile headers, then each tile's block, tiles row by row);
* a face flagged constant, or of resolution 1 x 1, takes its value from the constant block.

Reduction levels, meta data and edits are not read (edit blocks are refused). read_ptx returns every
face's texels; device_texture turns a file the device renders -- 8-bit data, at least 3 channels,
one resolution for all non-constant faces, quad or triangle meshes -- into the (res_log2, texels,
adjfaces, adjedges) that sgpu_scene_set_ptex takes, constant faces expanded to that resolution.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

MAGIC = 0x78657450  # 'P' 't' 'e' 'x'
MESH_TRIANGLE, MESH_QUAD = 0, 1
DT_UINT8, DT_UINT16, DT_HALF, DT_FLOAT = 0, 1, 2, 3
_DT = {DT_UINT8: np.uint8, DT_UINT16: np.uint16, DT_HALF: np.float16, DT_FLOAT: np.float32}
ENC_CONSTANT, ENC_ZIPPED, ENC_DIFFZIPPED, ENC_TILED = 0, 1, 2, 3
FLAG_CONSTANT, FLAG_HASEDITS, FLAG_NBCONSTANT, FLAG_SUBFACE = 1, 2, 4, 8
HEADER = struct.Struct("<IIIIiHHIIIIIIQII")  # PtexIO.h:44-63
EXT_HEADER = struct.Struct("<HHHHIIQQQ")  # PtexIO.h:64-74
LEVEL_INFO = struct.Struct("<QII")  # PtexIO.h:75-80
FACE_INFO = struct.Struct("<bbBB4i")  # Ptexture.h FaceInfo: Res, adjedges, flags, adjfaces[4]


class PtexError(ValueError):
    """A Ptex file this reader cannot read, or the device cannot render."""


def _unzip(buf, pos, zipsize, size, what):
    if pos + zipsize > len(buf):
        raise PtexError(f"{what}: block runs past the file's end")
    try:
        out = zlib.decompress(buf[pos:pos + zipsize])
    except zlib.error as e:
        raise PtexError(f"{what}: unzip failed ({e})") from None
    if len(out) != size:
        raise PtexError(f"{what}: {len(out)} bytes unzipped, {size} expected")
    return out


def _planes_to_pixels(raw, dt, nchan, uw, vw):
    """interleave: nchan planes of vw rows of uw values -> (vw, uw, nchan)."""
    a = np.frombuffer(raw, dt).reshape(nchan, vw, uw)
    return np.ascontiguousarray(a.transpose(1, 2, 0))


def read_ptx(src) -> dict:
    """A Ptex file's header fields, per-face info and level-0 texels: {"mesh_type", "data_type",
    "nchannels", "alpha_channel", "ulog2", "vlog2", "flags", "adjfaces" (n x 4), "adjedges" (n x 4),
    "constants" (n x nchannels), "faces" (list of (v res, u res, nchannels) arrays)}."""
    if isinstance(src, (bytes, bytearray, memoryview)):
        buf = bytes(src)
    else:
        with open(src, "rb") as f:
            buf = f.read()
    if len(buf) < HEADER.size:
        raise PtexError("not a Ptex file (too short)")
    (magic, version, meshtype, datatype, alphachan, nchan, nlevels, nfaces, extsize, fisize, cdsize, lisize,
     _minor, ldsize, mdzip, _mdmem) = HEADER.unpack_from(buf, 0)
    if magic != MAGIC:
        raise PtexError("not a Ptex file")
    if version != 1:
        raise PtexError(f"unsupported Ptex file version {version}")
    if datatype not in _DT or meshtype not in (MESH_TRIANGLE, MESH_QUAD):
        raise PtexError(f"unknown data type {datatype} or mesh type {meshtype}")
    dt = np.dtype(_DT[datatype]).newbyteorder("<")
    psize = dt.itemsize * nchan
    ext = bytes(buf[HEADER.size:HEADER.size + min(EXT_HEADER.size, extsize)]).ljust(EXT_HEADER.size, b"\0")
    _ub, _pad, _vb, _ef, _lmdzip, _lmdmem, _lmdsize, editsize, _editpos = EXT_HEADER.unpack(ext)
    if editsize:
        raise PtexError("Ptex edit blocks are not supported")
    pos = HEADER.size + extsize
    fi = _unzip(buf, pos, fisize, FACE_INFO.size * nfaces, "face info")
    pos += fisize
    const = np.frombuffer(_unzip(buf, pos, cdsize, psize * nfaces, "constant data"), dt).reshape(nfaces, nchan)
    pos += cdsize
    if lisize < LEVEL_INFO.size * nlevels or nlevels < 1:
        raise PtexError("level info block too small")
    level0_size, level0_hdr, level0_faces = LEVEL_INFO.unpack_from(buf, pos)
    pos += lisize
    if level0_faces != nfaces:
        raise PtexError(f"level 0 holds {level0_faces} faces of {nfaces}")
    ulog2 = np.zeros(nfaces, np.int32)
    vlog2 = np.zeros(nfaces, np.int32)
    flags = np.zeros(nfaces, np.uint8)
    adjf = np.zeros((nfaces, 4), np.int32)
    adje = np.zeros((nfaces, 4), np.int32)
    for f in range(nfaces):
        ul, vl, ae, fl, *af = FACE_INFO.unpack_from(fi, FACE_INFO.size * f)
        ulog2[f], vlog2[f], flags[f] = ul, vl, fl
        adjf[f] = af
        adje[f] = [(ae >> (2 * e)) & 3 for e in range(4)]
    fdh = np.frombuffer(_unzip(buf, pos, level0_hdr, 4 * nfaces, "level 0 face headers"), "<u4")
    data_pos = pos + level0_hdr
    offsets = data_pos + np.concatenate([[0], np.cumsum(fdh & 0x3FFFFFFF)[:-1]]).astype(np.int64)

    def face_block(p, header, uw, vw, what):
        enc, size = int(header) >> 30, int(header) & 0x3FFFFFFF
        if enc == ENC_CONSTANT:
            px = np.frombuffer(buf[p:p + psize], dt)
            return np.broadcast_to(px, (vw, uw, nchan)).copy()
        if enc in (ENC_ZIPPED, ENC_DIFFZIPPED):
            raw = _unzip(buf, p, size, psize * uw * vw, what)
            if enc == ENC_DIFFZIPPED and dt.kind == "u":
                raw = np.cumsum(np.frombuffer(raw, dt), dtype=dt).tobytes()  # decodeDifference, wrapping
            return _planes_to_pixels(raw, dt, nchan, uw, vw)
        # tiled: tile resolution, tile header size, zipped tile headers, tiles row by row
        tul, tvl = struct.unpack_from("<bb", buf, p)
        (ths,) = struct.unpack_from("<I", buf, p + 2)
        tu, tv = 1 << tul, 1 << tvl
        ntu, ntv = uw // tu, vw // tv
        if ntu < 1 or ntv < 1:
            raise PtexError(f"{what}: tile {tu} x {tv} larger than the face {uw} x {vw}")
        th = np.frombuffer(_unzip(buf, p + 6, ths, 4 * ntu * ntv, what + " tile headers"), "<u4")
        tp = p + 6 + ths + np.concatenate([[0], np.cumsum(th & 0x3FFFFFFF)[:-1]]).astype(np.int64)
        out = np.zeros((vw, uw, nchan), dt)
        for i in range(ntv):
            for j in range(ntu):
                t = i * ntu + j
                if int(th[t]) >> 30 == ENC_TILED:
                    raise PtexError(f"{what}: a tile is itself tiled")
                out[i * tv:(i + 1) * tv, j * tu:(j + 1) * tu] = face_block(int(tp[t]), th[t], tu, tv, what)
        return out

    faces = []
    for f in range(nfaces):
        uw, vw = 1 << int(ulog2[f]), 1 << int(vlog2[f])
        if flags[f] & FLAG_CONSTANT or (ulog2[f] == 0 and vlog2[f] == 0):
            faces.append(np.broadcast_to(const[f], (vw, uw, nchan)).copy())
        else:
            faces.append(face_block(int(offsets[f]), fdh[f], uw, vw, f"face {f}"))
    return {"mesh_type": meshtype, "data_type": datatype, "nchannels": nchan, "alpha_channel": alphachan,
            "ulog2": ulog2, "vlog2": vlog2, "flags": flags, "adjfaces": adjf, "adjedges": adje, "constants": const,
            "faces": faces, "level0_bytes": level0_size, "meta_zip_bytes": mdzip}


def device_texture(ptx: dict) -> dict:
    """The texture dict sgpu_scene_set_ptex takes ({"res_log2", "texels" (n, res, res, 3) uint8,
    "adjfaces", "adjedges"}) from read_ptx's result. Refused: other data types than 8-bit, fewer
    than 3 channels, subfaces, non-square faces, several resolutions among non-constant faces, or
    resolutions above 256 x 256. Constant faces are expanded to the common resolution and tiled faces
    assembled: Ptex weighs a constant face's value by the kernel's summed weight and applies the
    kernel to a tiled face tile by tile, where the device sums a face's texels in one pass, so
    lookups that touch such faces agree to float rounding (a few ulp); neighbourhood-constant faces,
    which Ptex returns unfiltered, likewise."""
    if ptx["data_type"] != DT_UINT8:
        raise PtexError("the device's per-face textures are 8-bit")
    if ptx["nchannels"] < 3:
        raise PtexError("the device's per-face textures have 3 channels (the renderer reads channels 0..2)")
    if np.any(ptx["flags"] & FLAG_SUBFACE):
        raise PtexError("subfaces (faces of a split non-quad) are not supported")
    if np.any(ptx["ulog2"] != ptx["vlog2"]):
        raise PtexError("non-square faces are not supported")
    n = len(ptx["faces"])
    varying = (ptx["flags"] & FLAG_CONSTANT) == 0
    varying &= ~((ptx["ulog2"] == 0) & (ptx["vlog2"] == 0))
    res = sorted(set(ptx["ulog2"][varying].tolist()))
    if len(res) > 1:
        raise PtexError(f"faces at several resolutions (log2 {res}); the device holds one per texture")
    rl = res[0] if res else 0
    if rl > 8:
        raise PtexError(f"face resolution 2^{rl} above 256")
    size = 1 << rl
    texels = np.zeros((n, size, size, 3), np.uint8)
    for f, a in enumerate(ptx["faces"]):
        texels[f] = a[:, :, :3] if varying[f] else np.broadcast_to(ptx["constants"][f][:3], (size, size, 3))
    return {"res_log2": rl, "texels": texels, "adjfaces": ptx["adjfaces"].copy(), "adjedges": ptx["adjedges"].copy(),
            "mesh_type": ptx["mesh_type"]}


def load_device_texture(path) -> dict:
    """read_ptx + device_texture: a ptex_loader for selas_amd.resources.load_scene."""
    return device_texture(read_ptx(path))
