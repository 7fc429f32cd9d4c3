"""The C-ABI library: loads, exports every declared symbol, parses init files
with the reference's float semantics, and fails loudly (no CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

import kfvm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kfvm.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(kfvm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(kfvm_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(kfvm_amd.EXPORTS) == names


def test_parse_init_uses_float_semantics_and_defaults():
    text = """# comment
mesh:resolution 64 32 16
mesh:extent 0.0 1.0 -0.5 0.5 0.0 2.0
stencil:lfac 5
time:final 0.2
time:cfl 1.25
time:atol 1.e-3
eos:gamma 1.4
eos:wenothresh 1.0
bogus:key 7
problem:bc periodic outflow reflecting noslip outflow outflow
problem:riemann hll
"""
    p = kfvm_amd.parse_init(text)
    assert p["n"] == [64, 32, 16]
    assert p["gamma"] == float(np.float32(1.4)) and p["gamma"] != 1.4
    assert p["final"] == float(np.float32(0.2))
    assert p["atol"] == float(np.float32(1e-3))
    assert p["lo"][1] == -0.5 and p["hi"][2] == 2.0
    assert p["lfac"] == 5.0
    assert p["bc"] == ["periodic", "outflow", "reflecting", "noslip", "outflow", "outflow"]
    assert p["rs"] == 1
    assert p["warnings"] == 1  # the unknown category is reported and skipped
    # defaults of ProblemSetup / EosParameters
    q = kfvm_amd.parse_init("")
    assert q["maxsteps"] == 1000 and q["rejectionlimit"] == 4
    assert q["wenothresh"] == 10.0 and q["lfac"] == 20.0
    assert q["atol"] == 1e-4 and q["rejectionthresh"] == 0.8


def test_weight_builder_rejects_bad_shapes():
    lib = kfvm_amd.load_library()
    f = np.zeros(10)
    rc = lib.kfvm_build_weights(3, 2, 5.0, 0.1, 0.1, 0.1, kfvm_amd._dptr(f), f.size,
                                kfvm_amd._dptr(f), f.size)
    assert rc != 0
    assert b"wrong size" in lib.kfvm_last_error()
    assert lib.kfvm_weight_sizes(4, 2, None, None) != 0


def test_solver_fails_loudly_without_a_gpu(have_gpu):
    if have_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(kfvm_amd.KfvmError):
        kfvm_amd.Solver(init_text="mesh:resolution 8 8 8\nproblem:name riemann3d_planar\n")


def test_solver_rejects_invalid_configuration_before_device_work():
    # configuration errors are reported before any HIP call
    for text, msg in [("problem:name riemann3d_planar\nproblem:dim 4\n", "dim"),
                      ("mesh:resolution 8 8 8\nproblem:equation linadv\nproblem:name linadv_sine\n", "exact"),
                      ("mesh:resolution 8 8 8\nproblem:name nope\n", "unknown problem"),
                      ("mesh:resolution 8 8 8\nproblem:name riemann3d_planar\n"
                       "problem:bc periodic outflow outflow outflow outflow outflow\n", "periodic"),
                      # a reference init file as it is: its problem is compiled in there
                      ("mesh:resolution 8 8 8\ntime:final 0.1\n", "names no problem")]:
        with pytest.raises(kfvm_amd.KfvmError, match=msg):
            kfvm_amd.Solver(init_text=text)


def test_flux_brick_shape_is_checked_before_any_device_work():
    """tune:fluxbrick selects among the 2D flux-kernel brick shapes; 3D has one
    shape, and the solver refuses another before it touches the device."""
    import kfvm_amd
    from helpers import config, init_text
    text = init_text(config(n=(8, 8, 8))) + "tune:fluxbrick 1\n"
    with pytest.raises(RuntimeError, match="tune:fluxbrick"):
        kfvm_amd.Solver(init_text=text)
