"""The Riemann solvers' reciprocal, quotient and square root (rcp_n / div_n /
sqrt_n, device/physics.hpp) against the IEEE operations on the GPU: the
check program evaluates 2^26 log-uniform operand pairs of both signs over
[2^-40, 2^40) and must find every result equal to the IEEE one, sqrt_n(0) = 0
and sqrt_n(-1) NaN."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kfvm-kokkos_amd", "build", "div_sqrt_check")


@pytest.mark.gpu
def test_division_and_square_root_match_ieee():
    assert os.path.exists(EXE), f"{EXE} not built; run kfvm-kokkos_amd/build.py"
    print("[div_sqrt] running", EXE, flush=True)
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=60, check=True).stdout
    print(out, flush=True)
    vals = {}
    for line in out.splitlines():
        tok = line.split()
        if tok[0] in ("rcp_n", "div_n", "sqrt_n") and tok[1] == "mismatches":
            vals[tok[0]] = (int(tok[2]), int(tok[4]))
    assert set(vals) == {"rcp_n", "div_n", "sqrt_n"}
    for name, (mismatches, max_ulp) in vals.items():
        assert mismatches == 0 and max_ulp == 0, f"{name}: {mismatches} results differ, up to {max_ulp} ulp"
    assert "sqrt_n(0) zero 1" in out and "sqrt_n(-1) nan 1" in out
