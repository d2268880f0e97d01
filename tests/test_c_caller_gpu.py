"""A C program built against include/selas_gpu.h queries one committed scene from 8 host threads at
once, each on its own stream, as Selas's path tracer calls RayPick / OcclusionRay from its worker
threads (Source/Applications/Selas/Source/PathTracer.cpp:325-344). Every thread's closest hits
(t, u, v, primID, geomID) and occlusion flags must equal the serial pass bit for bit, 4 rounds each.
The program is built by tests/c_caller/Makefile (from build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "c_caller", "capi_threads")


def test_eight_threads_on_their_own_streams_match_the_serial_pass():
    if not os.path.exists(EXE):
        pytest.fail("tests/c_caller/capi_threads is not built (make -C tests/c_caller)")
    print("capi_threads: 8 threads x 4 rounds, 65,536 triangles, 49,152 rays", flush=True)
    r = subprocess.run([EXE, "8", "4", "1024"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    lines = r.stdout.strip().splitlines()
    assert lines[-1].startswith("ok: 8 threads, 0 mismatches")
    assert sum(1 for ln in lines if ln.startswith("thread ")) == 8
    # the scene is hit by a fair share of the rays (the comparison is not of empty results)
    serial = lines[0]
    hits = int(serial.split("serial pass: ")[1].split(" hits")[0])
    assert hits > 10000, serial
