"""The C++ drop-in HostTriangulation (include/mcleap/Host_Triangulation.h)
compiled with g++ against libmcleap_hip.so and run on the CPU: two batches
of insertions + one single insertion (vertex ids stay stable), point
location, movePoint + untangle + delonize, moveFlipflop / moveFlipflopDelaunay on
the host.  The final mesh must
be valid, exactly Delaunay and equal to the oracle's triangulation of the
final positions."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "host_test")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-D__HIP_PLATFORM_AMD__",
           os.path.join(ROOT, "tests", "cpp", "host_test.cpp"), "-o", exe,
           "-L" + os.path.join(ROOT, "mcleap_amd"), "-lmcleap_hip", "-Wl,-rpath," + os.path.join(ROOT, "mcleap_amd")]
    subprocess.check_call(cmd)
    return exe


def test_cpp_host_dropin(oracle, host_exe):
    O = oracle
    out = subprocess.run([host_exe, "600"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.returncode
    lines = out.stdout.split("\n")
    npos, nhe, nt = map(int, lines[0].split())
    pos = np.array([list(map(float, l.split())) for l in lines[1:1 + npos]], np.float32)
    he = np.array([list(map(int, l.split())) for l in lines[1 + npos:1 + npos + nhe]], np.int32)
    t = np.array([int(l) for l in lines[1 + npos + nhe:1 + npos + nhe + nt]], np.int32)
    v = np.stack([np.arange(npos), np.zeros(npos)], 1).astype(np.int32)
    for h in range(nhe):  # any outgoing half-edge per vertex
        v[he[h, 1], 1] = h
    rc, msg = O.check_mesh(pos, v, he, t, periodic=False, require_ccw=True)
    assert rc == 0, msg
    assert O.count_non_delaunay(pos, he) == 0
    ref = O.delaunay(pos[4:])
    _, _, h2, t2 = ref.export()
    assert np.array_equal(O.canonical_triangles(he, t), O.canonical_triangles(h2, t2))
