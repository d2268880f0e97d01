"""The Ptex file reader (selas_amd/ptex.py) against the reference's PtexReader: files written by
tests/golden/ptx.py with every encoding (zipped, difference-zipped, tiled with zipped,
difference-zipped and constant tiles, constant faces, 1 x 1 faces), per-face resolutions, quad and
triangle meshes and 8-bit, 16-bit and half data, read back by PtexTexture::open + getFaceInfo +
getData (tests/golden/ptexio_*.npz, made by tests/golden/make_golden.py --only ptexio). Every face's
texels and face info must equal the reference's; device_texture's refusals and its expansion of
constant faces."""
import os

import numpy as np
import pytest

from selas_amd import ptex as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["mixed8", "tri16", "half1", "uniform8"]


def _gold(name):
    return np.load(os.path.join(GOLD, f"ptexio_{name}.npz")), os.path.join(GOLD, f"ptexio_{name}.ptx")


@pytest.mark.parametrize("name", NAMES)
def test_faces_and_face_info_match_reference_reader(name):
    g, path = _gold(name)
    r = P.read_ptx(path)
    info = g["info"]
    assert (r["mesh_type"], r["data_type"], r["nchannels"], len(r["faces"]), r["alpha_channel"]) == tuple(info[0][:5])
    offs = np.concatenate([[0], np.cumsum(g["sizes"])])
    for f, a in enumerate(r["faces"]):
        assert a.tobytes() == g["data"][offs[f]:offs[f + 1]].tobytes(), f"face {f}"
        i = info[f]
        assert (r["ulog2"][f], r["vlog2"][f], r["flags"][f]) == (i[5], i[6], i[8])
        assert sum(int(e) << (2 * k) for k, e in enumerate(r["adjedges"][f])) == i[7]
        assert list(r["adjfaces"][f]) == list(i[9:13])


def test_encodings_are_all_exercised():
    """The mixed file holds a difference-zipped face, a 1 x 1 face, a constant face and a tiled
    face whose tiles use three encodings (the coverage the comparison above relies on)."""
    import struct
    import zlib

    _, path = _gold("mixed8")
    buf = open(path, "rb").read()
    h = P.HEADER.unpack_from(buf, 0)
    extsize, fisize, cdsize, lisize = h[8], h[9], h[10], h[11]
    pos = P.HEADER.size + extsize + fisize + cdsize
    _, hdr, n = P.LEVEL_INFO.unpack_from(buf, pos)
    fdh = struct.unpack("<%dI" % n, zlib.decompress(buf[pos + lisize:pos + lisize + hdr]))
    encs = [x >> 30 for x in fdh]
    assert {P.ENC_ZIPPED, P.ENC_DIFFZIPPED, P.ENC_CONSTANT, P.ENC_TILED} <= set(encs)
    r = P.read_ptx(path)
    assert r["flags"][3] & P.FLAG_CONSTANT and (r["ulog2"][2], r["vlog2"][2]) == (0, 0)


def test_device_texture_of_a_uniform_file():
    g, path = _gold("uniform8")
    r = P.read_ptx(path)
    t = P.device_texture(r)
    assert t["res_log2"] == 3 and t["texels"].shape == (12, 8, 8, 3) and t["texels"].dtype == np.uint8
    for f in range(12):
        np.testing.assert_array_equal(t["texels"][f], r["faces"][f][:, :, :3])
    np.testing.assert_array_equal(t["adjfaces"], r["adjfaces"])
    # the constant faces: every texel the constant value
    for f in (2, 7):
        assert r["flags"][f] & P.FLAG_CONSTANT and np.all(t["texels"][f] == r["constants"][f][:3])


@pytest.mark.parametrize("name,match", [("mixed8", "non-square"), ("tri16", "8-bit"),
                                        ("half1", "8-bit")])
def test_device_texture_refusals(name, match):
    _, path = _gold(name)
    with pytest.raises(P.PtexError, match=match):
        P.device_texture(P.read_ptx(path))


def test_device_texture_refuses_fewer_channels_and_subfaces():
    _, path = _gold("uniform8")
    r = P.read_ptx(path)
    r1 = dict(r, nchannels=2)
    with pytest.raises(P.PtexError, match="3 channels"):
        P.device_texture(r1)
    assert P.device_texture(dict(r, mesh_type=P.MESH_TRIANGLE))["mesh_type"] == P.MESH_TRIANGLE
    with pytest.raises(P.PtexError, match="subfaces"):
        P.device_texture(dict(r, flags=r["flags"] | P.FLAG_SUBFACE))


def test_malformed_files_are_refused():
    _, path = _gold("uniform8")
    buf = open(path, "rb").read()
    with pytest.raises(P.PtexError, match="not a Ptex"):
        P.read_ptx(b"XXXX" + buf[4:])
    with pytest.raises(P.PtexError):
        P.read_ptx(buf[:200])
    bad = bytearray(buf)
    bad[4] = 2  # version
    with pytest.raises(P.PtexError, match="version"):
        P.read_ptx(bytes(bad))


def test_fixture_scene_texture_loads():
    """The resource fixture's floor texture through load_device_texture (the ptex_loader
    load_scene takes) equals the texture the fixture was written from."""
    import resource_fixtures as F
    from selas_amd import resources as R

    ls = R.load_scene(F.FIXTURE_ROOT, "rooms", ptex_loader=P.load_device_texture)
    t = ls.description["models"][2]["geoms"][0]["ptex"]
    want = F.floor_ptex()
    assert t["res_log2"] == want["res_log2"]
    np.testing.assert_array_equal(t["texels"], want["texels"])
    np.testing.assert_array_equal(t["adjfaces"], want["adjfaces"])
    np.testing.assert_array_equal(t["adjedges"], want["adjedges"])
