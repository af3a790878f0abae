"""Build the native library in-tree: ``sylver_amd/libsylver_amd.so``.

One shared object holds the C ABI (include/sylver/sylver.h, include/sylver_amd.h),
the CPU symbolic analysis and the gfx950 HIP kernels. Built with hipcc for
gfx950 only; run ``python -m sylver_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsylver_amd.so")

SOURCES = ["numeric.hip", "indef.hip", "costmodel.hip", "lu.hip", "analyse.cpp", "scaling.cpp", "rb_io.cpp",
           "sylver_api.cpp"]
# the reference driver drivers/spldlt_test.F90 on the C interface (gcc)
DRIVER = os.path.join(HERE, "drivers", "spldlt_test.c")
DRIVER_OUT = os.path.join(HERE, "spldlt_test")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compiles every source (in parallel, one hipcc per source) and links."""
    from concurrent.futures import ThreadPoolExecutor

    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmds.append([hipcc(), *HIP_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", path, "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        list(ex.map(run, cmds))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    cmd = ["gcc", "-O2", "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-Wall", "-Werror",
           "-I", os.path.join(HERE, "..", "include"), DRIVER, "-L", HERE, "-lsylver_amd",
           "-Wl,-rpath,$ORIGIN", "-lm", "-o", DRIVER_OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER_OUT + ".tmp", DRIVER_OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
