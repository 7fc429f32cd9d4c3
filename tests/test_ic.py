"""Initial conditions and init files (no GPU).

IntegrateIC_K (Numeric_K.H:31-107) averages a point IC over the NQ^D Gauss
points of each cell.  The library's registry (kfvm_initial_condition, the
same code kfvm_create runs) is compared with
  * the oracle's independent restatement (orc_set_ic) for every problem the
    parity tests and benchmark configurations use;
  * closed forms: the exact cell average of sin(2 pi (a x + b y + c z)) is
    the centre value times sinc(pi a dx) sinc(pi b dy) sinc(pi c dz), which
    the 3-point Gauss rule meets to ~1e-14 at these resolutions (catches a
    wrong node order, weight, cell centre or axis); constant states of the
    Riemann problems in cells the discontinuity does not cut.
Init files of the reference's examples (tests/golden/*.init, copied
unchanged from exec/Hydro/Riem_3d_1c/riem.init and
exec/Hydro/Riem_2d_3c/riem.init) parse to their values, with the scheme the
example compiles in taken from the registry entry the problem:name line
selects.
"""
import os

import numpy as np
import pytest

import kfvm_amd
from helpers import config, init_text, oracle_for

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_ic(c):
    text = init_text(c)
    p = kfvm_amd.parse_init(text)
    n = p["n"]
    dims = [(p["hi"][a] - p["lo"][a]) / n[a] for a in range(3)]
    fw, dw = kfvm_amd.build_weights(p["dim"], p["nq"], p["lfac"], *dims)
    o = oracle_for(c, p, fw, dw, threads=2)
    o.set_ic(c["problem"], p["user"], p["seed"])
    return o.state


IC_CASES = {
    "riemann3d_planar": config(n=(12, 10, 8), user=(1.5,)),
    "riemann3d_planar_nq3": config(n=(9, 9, 9), nq=3, user=(1.5,)),
    "linadv_sine_2d": config(dim=2, n=(16, 12, 1), nq=3, equation="linadv", riemann="exact",
                             problem="linadv_sine", user=(1.0, 1.0, 0.0), bc="periodic"),
    "shear2d": config(dim=2, n=(16, 16, 1), problem="shear2d", user=(), seed=3, bc="periodic"),
    "quadrant2d": config(dim=2, n=(12, 12, 1), problem="quadrant2d", user=(), seed=11),
    "mhd_vortex3d": config(n=(8, 8, 8), equation="mhd_glm", riemann="llf", problem="mhd_vortex3d",
                           user=(), bc="periodic", gamma=5.0 / 3.0),
    "orszag_tang": config(dim=2, n=(12, 12, 1), equation="mhd_glm", riemann="hll", problem="orszag_tang",
                          user=(), bc="periodic", gamma=5.0 / 3.0),
    "mhd_blast": config(n=(8, 8, 8), equation="mhd_glm", riemann="llf", problem="mhd_blast",
                        user=(0.3, 0.5, 0.5), gamma=5.0 / 3.0),
    "riem_3d_1c": config(n=(10, 10, 10), problem="riem_3d_1c", user=(), lo=(-1.0, -1.0, -1.0)),
    "riem_2d_3c": config(dim=2, n=(15, 15, 1), problem="riem_2d_3c", user=()),
    "disprel": config(n=(8, 4, 4), equation="linadv", riemann="exact", problem="disprel",
                      user=(0.3, 2.0), bc="periodic"),
}


@pytest.mark.parametrize("name", sorted(IC_CASES))
def test_library_ic_matches_oracle(name):
    """Two independent restatements of the same point ICs and Gauss rule:
    bitwise equal (same operation order), except where the point functions
    are written with different but equivalent arithmetic."""
    c = IC_CASES[name]
    U = kfvm_amd.initial_condition(init_text(c))
    V = _oracle_ic(c)
    assert U.shape == V.shape
    scale = np.abs(V).max()
    assert np.abs(U - V).max() <= 1e-15 * scale


def test_linadv_sine_exact_cell_average():
    """The configuration of the 2D linear-advection case (128^2, NQ = 3) and a
    skewed wave (a, b) = (1, 2) on a non-square grid: the 3-point Gauss rule's
    error is (k h)^6 / 2016000 relative, ~1e-14 and ~1e-11 here."""
    for (nx, ny), (a, b), tol in [((128, 128), (1.0, 1.0), 1e-13), ((96, 80), (1.0, 2.0), 2e-11)]:
        c = config(dim=2, n=(nx, ny, 1), nq=3, equation="linadv", riemann="exact", problem="linadv_sine",
                   user=(a, b, 0.0), bc="periodic")
        U = kfvm_amd.initial_condition(init_text(c))[0, 0]
        dx, dy = 1.0 / nx, 1.0 / ny
        xc = (np.arange(nx) + 0.5) * dx
        yc = (np.arange(ny) + 0.5) * dy
        # np.sinc(t) = sin(pi t) / (pi t)
        exact = np.sin(2 * np.pi * (a * xc[None, :] + b * yc[:, None])) * np.sinc(a * dx) * np.sinc(b * dy)
        assert np.abs(U - exact).max() <= tol
        # the 2-point rule is exact only to its own (lower) order: the closed form tells it apart
        U2 = kfvm_amd.initial_condition(init_text(dict(c, nq=2)))[0, 0]
        assert np.abs(U2 - exact).max() > 100 * tol


def test_riemann_ics_uncut_cells_hold_the_constant_states():
    g = float(np.float32(1.4))  # eos:gamma goes through std::stof
    # planar problem: cells entirely on one side of x + y + z = 1.5
    n = 12
    c = config(n=(n, n, n), user=(1.5,))
    U = kfvm_amd.initial_condition(init_text(c))
    h = 1.0 / n
    idx = np.arange(n)
    s = idx[None, None, :] + idx[None, :, None] + idx[:, None, None]  # i + j + k
    left = (s + 3) * h < 1.5   # the cell's far corner is left of the plane
    right = s * h > 1.5        # its near corner is right of it
    assert left.sum() > 100 and right.sum() > 100
    np.testing.assert_array_equal(U[0][left], 1.0)
    np.testing.assert_array_equal(U[4][left], 1.0 / (g - 1.0))
    np.testing.assert_array_equal(U[0][right], 0.125)
    np.testing.assert_array_equal(U[4][right], 0.1 / (g - 1.0))
    # cut cells: the 2-point Gauss rule (nodes +-1/(2 sqrt 3) of the cell,
    # weight 1/8 each) sees m of its 8 nodes left of the plane
    g1 = 0.5 / np.sqrt(3.0)
    cut = ~(left | right)
    for k, j, i in zip(*np.nonzero(cut)):
        m = sum((i + 0.5 + a) * h + (j + 0.5 + b) * h + (k + 0.5 + e) * h < 1.5
                for a in (-g1, g1) for b in (-g1, g1) for e in (-g1, g1))
        assert abs(U[0, k, j, i] - (0.125 + 0.875 * m / 8.0)) <= 1e-15
    # octant problem on [-1,1]^3 (exec/Hydro/Riem_3d_1c/InitialCondition.H): with an
    # even cell count no cell is cut; the top-north-east octant is (0.53125, 0, 0, 0, 0.4)
    c = config(n=(8, 8, 8), problem="riem_3d_1c", user=(), lo=(-1.0, -1.0, -1.0))
    U = kfvm_amd.initial_condition(init_text(c))
    # (sums of 8 equal Gauss terms round at the last bit)
    close = lambda a, b: np.testing.assert_allclose(a, b, rtol=1e-15, atol=0)  # noqa: E731
    close(U[0, 4:, 4:, 4:], 0.53125)
    close(U[4, 4:, 4:, 4:], 0.4 / (g - 1.0))
    close(U[0, :4, :4, :4], 1.016216216)
    close(U[1, :4, :4, :4], 1.016216216 * -0.401442839)
    # top south west octant (z >= 0, y < 0, x < 0): rho 0.8, w = -0.727606875
    close(U[3, 4:, :4, :4], 0.8 * -0.727606875)
    assert np.all(U[1:4, 4:, 4:, 4:] == 0.0)


def _golden(name):
    path = os.path.join(GOLDEN, name)
    with open(path) as f:
        text = f.read()
    ref = "/root/reference/exec/Hydro/" + {"riem_3d_1c.init": "Riem_3d_1c/riem.init",
                                           "riem_2d_3c.init": "Riem_2d_3c/riem.init"}[name]
    if os.path.exists(ref):
        with open(ref) as f:
            assert f.read() == text, "the fixture must stay a verbatim copy of the reference's file"
    return text


def test_reference_init_file_needs_a_problem_name():
    text = _golden("riem_3d_1c.init")
    p = kfvm_amd.parse_init(text)
    assert p["problem_set"] is False
    with pytest.raises(kfvm_amd.KfvmError, match="names no problem"):
        kfvm_amd.Solver(init_text=text)


def test_reference_init_file_parses_with_its_compiled_in_scheme():
    text = _golden("riem_3d_1c.init") + "problem:name riem_3d_1c\n"
    p = kfvm_amd.parse_init(text)
    assert p["problem_set"] is True and p["problem"] == "riem_3d_1c"
    assert p["n"] == [128, 128, 128]
    assert p["lo"] == [-1.0, -1.0, -1.0] and p["hi"] == [1.0, 1.0, 1.0]
    # reals go through std::stof (ProblemSetup.H:236-305)
    assert p["final"] == float(np.float32(0.53)) and p["cfl"] == 1.25
    assert p["atol"] == float(np.float32(1e-3)) and p["rtol"] == float(np.float32(1e-3))
    assert p["maxsteps"] == 2500 and p["lfac"] == 5.0 and p["gamma"] == float(np.float32(1.4))
    assert p["wenothresh"] == 1.0
    # Riem_3d_1c/Definitions.H and ProblemSetup.cpp: 3D, NQ 2, hydro, HLLC, SSP(4,3,2), outflow
    assert (p["dim"], p["nq"], p["eq"], p["rs"], p["rk"]) == (3, 2, 0, 2, 2)
    assert p["bc"] == ["outflow"] * 6
    assert p["warnings"] == 0  # io:pdiconf and mesh:numblocks 1 1 1 are understood
    # explicit problem: keys still override the preset
    p2 = kfvm_amd.parse_init(text + "problem:nquad 3\nproblem:riemann hll\n")
    assert p2["nq"] == 3 and p2["rs"] == 1
    p3 = kfvm_amd.parse_init(_golden("riem_2d_3c.init") + "problem:name riem_2d_3c\n")
    assert (p3["dim"], p3["nq"], p3["n"][:2]) == (2, 2, [400, 400])


def test_named_user_parameters_of_a_reference_problem():
    """DispRel's UserParameters.H: ang = pi * stof(value), freq = stof(value)."""
    base = "mesh:resolution 8 4 4\nproblem:name disprel\n"
    p = kfvm_amd.parse_init(base + "user:ang 0.25\nuser:freq 3.0\n")
    assert p["user"][0] == np.pi * float(np.float32(0.25)) and p["user"][1] == 3.0
    assert (p["dim"], p["eq"], p["rs"]) == (3, 1, 3) and p["bc"] == ["periodic"] * 6
    assert p["warnings"] == 0
    p = kfvm_amd.parse_init(base)
    assert p["user"][1] == 1.0  # freq's default
    p = kfvm_amd.parse_init(base + "user:nope 1\n")
    assert p["warnings"] == 1
