"""The C++ drop-in headers on the GPU: tests/cpp/dropin_test.cpp (built by
__graft_entry__.build() with hipcc) runs the reference's call sequence --
HostTriangulation construction, DeviceTriangulation movePoints / untangle2 /
delonize2 / oneRing / closestNeighbors / getFRNN / transferToHost and
GridCount2d update + apply_f_frnn with a device functor -- and checks each
result on the host against brute force and the library's exact predicates."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_test")


def test_cpp_dropin_sequence():
    assert os.path.exists(EXE), "tests/cpp/dropin_test is not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=110)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "DROPIN OK" in r.stdout
