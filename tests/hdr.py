"""A reader of Radiance .hdr pictures for the tests (the format as its documentation states it:
header lines up to a blank line, the resolution line "-Y h +X w", then scanlines, each either run-
length coded per byte plane -- the marker 2 2 w_hi w_lo, then per plane: n > 128 a run of n - 128
copies of the next byte, else n literal bytes -- or flat RGBE quadruples)."""
import numpy as np


def read_rgbe(path_or_bytes):
    """(header lines, RGBE bytes as h x w x 4 uint8)."""
    data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    end = data.index(b"\n\n")
    header = data[:end].decode().split("\n")
    pos = end + 2
    nl = data.index(b"\n", pos)
    res = data[pos:nl].decode().split()
    assert res[0] == "-Y" and res[2] == "+X", res
    h, w = int(res[1]), int(res[3])
    pos = nl + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and not data[pos + 2] & 0x80:
            assert (data[pos + 2] << 8 | data[pos + 3]) == w
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = data[pos]
                    pos += 1
                    if n > 128:
                        n -= 128
                        out[y, x:x + n, c] = data[pos]
                        pos += 1
                    else:
                        assert n > 0
                        out[y, x:x + n, c] = np.frombuffer(data[pos:pos + n], np.uint8)
                        pos += n
                    x += n
                assert x == w
        else:
            out[y] = np.frombuffer(data[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    assert pos == len(data)
    return header, out


def rgbe_to_float(rgbe):
    """Components byte x 2^(e - 136) (0 where e == 0)."""
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), 0.0).astype(np.float32)
    return rgbe[..., :3].astype(np.float32) * scale[..., None]
