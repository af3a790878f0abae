"""Build the CPU oracle: ``oracle/liboracle.so`` (g++, OpenMP).

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg load it.

The reference itself cannot be compiled here (its API layer is Fortran, no
gfortran exists, and its SPRAL submodule, StarPU, METIS and hwloc are
absent), so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v3",
         "-fno-fast-math"]


def build(verbose: bool = False) -> str:
    """oracle.cpp keeps the compiler's default FMA contraction (it is the
    timed CPU baseline); oracle_indef.cpp is built with -ffp-contract=off
    because its roundings are written out explicitly, and so is
    oracle_lu.cpp."""
    objs = []
    for src, extra in (("oracle.cpp", []), ("oracle_indef.cpp", ["-ffp-contract=off"]),
                       ("oracle_lu.cpp", ["-ffp-contract=off"])):
        obj = os.path.join(HERE, "_build", src.replace(".cpp", ".o"))
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = ["g++", *FLAGS, *extra, "-c", os.path.join(HERE, src), "-o", obj]
        cmd.remove("-shared")
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = ["g++", "-shared", "-fopenmp", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
