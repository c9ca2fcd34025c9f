"""Reverse-mode primitives of the discrete adjoint (zfv_adjoint_ops.hpp)
against central differences of the forward primitives (zfv_physics.hpp).

The device headers compile as plain C++ with the qualifier stand-ins of
tests/hostshim; tests/adjoint_ops_check.cpp draws random physical states and
checks <rev(ybar), d> = <ybar, (f(x + eps d) - f(x - eps d)) / 2 eps> for the
conversions, the physical flux, HLLC (all wave-speed branches), the face
extrapolation and the minmod / MC limiter products.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_reverse_primitives_match_central_differences(tmp_path):
    exe = tmp_path / "adjoint_ops_check"
    subprocess.run(["g++", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "tests", "hostshim"),
                    "-I", os.path.join(ROOT, "zephyr_amd", "csrc"),
                    os.path.join(ROOT, "tests", "adjoint_ops_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("worst rel err") == 8
