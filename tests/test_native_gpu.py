"""The C++ drop-in program (tests/native/dropin_test.cpp) run on the GPU:
the reference's own call patterns through pcg_hip:: (plateau checksum,
output layer, poly-mul patterns, OLE relation, the two parties in two
threads)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "pcg-square_amd", "pcgsq")


def test_dropin_program_passes(tmp_path):
    exe = tmp_path / "dropin_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "dropin_test.cpp"), "-L" + LIBDIR,
                    "-lpcgsq", "-pthread", "-Wl,-rpath," + LIBDIR, "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RESULT: PASS" in r.stdout
    assert "checksum=16e395febfd6d23f" in r.stdout
