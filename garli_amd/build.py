"""Build libgle.so (HIP kernels + C++ host layers) in-tree for gfx950.

    python -m garli_amd.build

Objects go to garli_amd/_build/, the library to garli_amd/libgle.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GLE_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), p.stdout, p.stderr))
    return p


def sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = sources()
    jobs = []
    for f in hip:
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f),
                                          "-o", os.path.join(BUILD, f + ".o")])
    for f in cpp:
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, f),
                                          "-o", os.path.join(BUILD, f + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for cmd, p in zip(jobs, ex.map(_run, jobs)):
            if verbose:
                print(" ".join(cmd[:1] + cmd[-3:]))
    objs = [os.path.join(BUILD, f + ".o") for f in hip + cpp]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
