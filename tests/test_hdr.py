"""The .hdr reader the GPU tests use, pinned against the reference's own HDR writer and reader:
stbi_write_hdr's files (FrameBuffer_Save's writer, recorded in tests/golden/hdr_*.npz by
tests/golden/make_golden.py) read back to the pixels stbi_loadf reads from them, bit for bit, for
a run-length coded image (21 x 6) and a flat one (5 x 3)."""
import os

import numpy as np
import pytest

import hdr

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("case", [0, 1])
def test_reader_matches_stb(case):
    g = np.load(os.path.join(GOLD, f"hdr_{case}.npz"))
    header, rgbe = hdr.read_rgbe(bytes(g["bytes"]))
    assert header[0] == "#?RADIANCE" and "FORMAT=32-bit_rle_rgbe" in header
    assert rgbe.shape[:2] == g["image"].shape[:2]
    np.testing.assert_array_equal(hdr.rgbe_to_float(rgbe), g["decoded"])
    # the quantization is Ward's: within one part in 256 of the largest component
    img = g["image"]
    m = img.max(axis=2, keepdims=True)
    err = np.abs(hdr.rgbe_to_float(rgbe) - img)
    assert (err <= m / 128 + 1e-30).all()
