"""A compiled C caller of the public API (tests/c/spldlt_example.c), built by
gcc against include/ (the reference's include path "sylver/sylver.h") and
linked to the in-tree libsylver_amd.so.

CPU: it compiles and links, and without a HIP device it reports the
factorization's no-device flag (-51: no CPU fallback) through the C ABI.
GPU: it runs to the known answer x = [1.5, 2, 1.5]
(examples/C/spldlt_simple_example_c.c in the reference).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "spldlt_example.c")
LIBDIR = os.path.join(ROOT, "sylver_amd")


def _build(tmp_path):
    exe = str(tmp_path / "spldlt_example")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
                    "-L", LIBDIR, "-lsylver_amd", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe], check=True)
    return exe


def test_c_caller_builds_and_refuses_without_device(S, tmp_path):
    exe = _build(tmp_path)
    if S.device_count() > 0:
        pytest.skip("a HIP device is visible: the GPU test runs the program")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "spldlt_factorize flag -51" in r.stdout


@pytest.mark.gpu
def test_c_caller_known_answer(S, gpu, tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    x = [float(v) for v in r.stdout.split("=")[1].split()]
    assert max(abs(a - b) for a, b in zip(x, [1.5, 2.0, 1.5])) <= 1e-14, r.stdout
