"""Builds libmcleap_hip.so in-tree with hipcc for gfx950.

    python -m mcleap_amd.build        (or mcleap_amd.build.build())

Every .hip file under csrc/ is compiled to an object in _build/ and linked
into mcleap_amd/libmcleap_hip.so.  Flags: -O3, no floating-point
contraction (the predicates' error bounds assume separately rounded
operations), no packed-FP32 instructions.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmcleap_hip.so")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCL_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    return obj, warn


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            for line in w:
                print(line)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, OUT)
    return OUT


def build_cpp_example(verbose=False):
    """the drop-in C++ test program (tests/cpp/dropin_test.cpp) against the
    headers in include/mcleap and the built library"""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", src, "-o", exe, "-L" + HERE, "-lmcleap_hip",
                             "-Wl,-rpath,$ORIGIN/../../mcleap_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_example())
