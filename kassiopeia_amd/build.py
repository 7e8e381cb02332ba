"""Build the HIP library ``kassiopeia_amd/libkemhip.so`` for gfx950 (MI355X) with hipcc.

Every call compiles from source (objects in ``kassiopeia_amd/_build``), so the library always
matches the tree.  Usage: ``python -m kassiopeia_amd.build``.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libkemhip.so")
BUILD = os.path.join(HERE, "_build")

ARCH = os.environ.get("KEM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# packed FP32 math is disabled for every HIP translation unit (the backend emits no
# v_pk_*_f32 then); the host pass only warns that it does not know the feature.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", f"--offload-arch={ARCH}", "-munsafe-fp-atomics"]

SOURCES = [
    ("kem_kernels.hip", True),
    ("kem_sort.hip", True),
    ("kem_batch.hip", True),
    ("kem_adjoint.hip", True),
    ("kem_zh.hip", True),
    ("kem_capi.cpp", True),  # holds the no-allocation scratch (a __device__ array)
    ("kem_prep.cpp", False),
    ("kem_cache.cpp", False),
    ("kem_zh_gen.cpp", False),
]


def _compile(src, device):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if device:
        cmd[1:1] = ["-x", "hip"] + DEVICE_FLAGS
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp], check=True)
    os.replace(tmp, OUT)
    # the C++ facade example (include/kemhip.hpp), linked against the library just built
    facade = os.path.join(BUILD, "facade_test")
    subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall",
                    "-I", os.path.join(os.path.dirname(HERE), "include"),
                    os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_test.cpp"),
                    "-L", HERE, "-lkemhip", "-Wl,-rpath,$ORIGIN/..", "-o", facade], check=True)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
