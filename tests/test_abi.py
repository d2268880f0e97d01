"""CPU-side checks of the C ABI: the library loads, exports every declared entry point, and its
host-only generators are deterministic. No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np

from selas_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "selas_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sgpu_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in selas_gpu.h but not exported"


def test_binding_covers_header():
    assert sorted(N.declared_symbols()) == header_symbols()


def test_cluster_soup_deterministic():
    a = N.gen_cluster_soup(11, 4, 16)
    b = N.gen_cluster_soup(11, 4, 16)
    c = N.gen_cluster_soup(12, 4, 16)
    assert a.shape == (64, 9)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    e1 = a[:, 3:6] - a[:, 0:3]
    e2 = a[:, 6:9] - a[:, 0:3]
    l1 = np.linalg.norm(e1, axis=1)
    l2 = np.linalg.norm(e2, axis=1)
    assert (l1 > 0.9e-4).all() and (l1 < 1.1e-2).all()
    assert (l2 > 0.9e-4).all() and (l2 < 1.1e-2).all()


def test_camera_rays_shape_and_norm():
    r = N.gen_camera_rays((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), (0, 1, 0), 50.0, 8, 4, 3)
    assert r["org_x"].shape == (96,)
    d = np.stack([r["dir_x"], r["dir_y"], r["dir_z"]], 1)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert (r["dir_z"] > 0).all()
    assert np.isinf(r["tfar"]).all() and (r["tnear"] == 0).all()


def test_invalid_arguments_are_errors_not_crashes():
    out = np.empty((4, 9), np.float32)
    rc = N.lib.sgpu_gen_cluster_soup(1, 0, 4, 0.02, 1e-4, 1e-2, N.fptr(out))
    assert rc == N.SGPU_E_INVALID
    assert b"bad argument" in N.lib.sgpu_last_error()
    rc = N.lib.sgpu_gen_camera_rays(None, None, None, 50.0, 4, 4, 1, *([None] * 8))
    assert rc == N.SGPU_E_INVALID


def test_c_caller_compiles_against_the_header(tmp_path):
    """tests/c_caller/capi_threads.c is plain C11 against include/selas_gpu.h (no HIP headers), warnings
    as errors, linked with libselas_gpu; it is run on the GPU by tests/test_c_caller_gpu.py."""
    import subprocess
    src = os.path.join(ROOT, "tests", "c_caller", "capi_threads.c")
    out = tmp_path / "capi_threads"
    r = subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-I" + os.path.join(ROOT, "include"), "-o", str(out), src,
                        "-L" + os.path.dirname(N.LIB_PATH), "-lselas_gpu", "-pthread"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.exists()
