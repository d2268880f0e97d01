"""GPU: sgpu_framebuffer_save_hdr (the replacement of FrameBuffer_Save) against the reference's own
HDR writer: the RGBE bytes of every pixel equal those stbi_write_hdr wrote for the same image
(tests/golden/hdr_*.npz), so stbi_loadf reads the same floats; run-length coded and flat
scanlines; one file per layer."""
import ctypes as C
import os

import numpy as np
import pytest

import hdr
from selas_amd import native as N

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "libselas_ref.so")


@pytest.mark.parametrize("case", [0, 1])
def test_rgbe_bytes_match_stb(gpu_ctx, tmp_path, case):
    g = np.load(os.path.join(GOLD, f"hdr_{case}.npz"))
    img = g["image"]
    h, w = img.shape[:2]
    prefix = str(tmp_path / f"img{case}")
    print(f"hdr case {case} {w}x{h}", flush=True)
    N.save_hdr(gpu_ctx, prefix, img, w, h, 1)
    header, mine = hdr.read_rgbe(prefix + "_0.hdr")
    _, theirs = hdr.read_rgbe(bytes(g["bytes"]))
    np.testing.assert_array_equal(mine, theirs)
    np.testing.assert_array_equal(hdr.rgbe_to_float(mine), g["decoded"])
    if os.path.exists(REF):  # stb's own reader on this file
        lib = C.CDLL(REF)
        dec = np.zeros((h, w, 3), np.float32)
        n = lib.ref_read_hdr((prefix + "_0.hdr").encode(), dec.ctypes.data_as(C.POINTER(C.c_float)), dec.size)
        assert n == w * h
        np.testing.assert_array_equal(dec, g["decoded"])


def test_layers_and_long_rows(gpu_ctx, tmp_path):
    rng = np.random.default_rng(3)
    w, h = 600, 7
    fb = rng.lognormal(0.0, 1.5, (2, h, w, 3)).astype(np.float32)
    fb[0, :, 100:400] = 0.5  # long runs: split into runs of at most 127
    N.save_hdr(gpu_ctx, str(tmp_path / "pt"), fb, w, h, 2)
    for layer in range(2):
        _, rgbe = hdr.read_rgbe(str(tmp_path / f"pt_{layer}.hdr"))
        dec = hdr.rgbe_to_float(rgbe)
        m = fb[layer].max(axis=2, keepdims=True)
        assert (np.abs(dec - fb[layer]) <= m / 128).all()
    with pytest.raises(N.SgpuError):
        N.save_hdr(gpu_ctx, str(tmp_path / "no_such_dir" / "x"), fb[0], w, h, 1)
