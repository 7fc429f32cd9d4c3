"""The HLLC shear factor's reduced-range power (csrc/device/powu.hpp), built
for the host and checked against long-double powl."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_powu_matches_libm(tmp_path):
    exe = tmp_path / "powu_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(HERE, "powu_check.cpp")], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    rel, ab = float(out[0]), float(out[1])
    edges = [float(v) for v in out[2:]]
    assert rel < 32.0, rel         # < 3.6e-15 relative where the power is not tiny
    assert ab <= 2.3e-16, ab       # 1 - x^y, the quantity HLLC uses, to one ulp of 1
    assert edges == [1.0, 1.0, 0.0, 0.5, 0.0]
