"""Parity of the HIP stage library with the CPU oracle on the same inputs.

Gate (per the path's contract): per-cell relative difference <= 1e-12 after one
fixed-dt step, same weight table on both sides, each side starting from the
initial condition it computes itself (the two are first checked to agree to
1e-15).  The relative difference of a cell is max_v |a_v - b_v| / max_v |b_v|.
Cells whose discrete decisions differ (WENO flag) are counted separately and
must be zero here.
"""
import os
import sys

import numpy as np
import pytest

from helpers import cell_rel_diff, config, init_text, oracle_for

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12


def _pair(c):
    import kfvm_amd
    s = kfvm_amd.Solver(init_text=init_text(c))
    parsed = kfvm_amd.parse_init(init_text(c))
    f, d = s.weights
    o = oracle_for(c, parsed, f, d)
    o.parsed = parsed
    return s, o


def _start_oracle(s, o, c):
    """The oracle computes its own initial condition (orc_set_ic), which must
    equal the library's; then it FSAL-primes like the solver's constructor."""
    p = o.parsed
    o.set_ic(c["problem"], p["user"], p["seed"])
    U0, V0 = s.state, o.state
    d = np.abs(U0 - V0).max() / max(np.abs(V0).max(), 1e-300)
    assert d <= 1e-15, f"initial conditions differ: {d:.3e}"
    fl = np.zeros((2,) + s.shape[1:])
    fl[1] = 1.0
    o.flags = fl
    o.prime()
    return U0


def _report(name, r, extra=""):
    print(f"[parity] {name}: max cell rel diff {r.max():.3e} {extra}", flush=True)


CASES = {
    "riemann3d_sparse": config(n=(24, 24, 24), wenothresh=1.0, fixeddt=0.004),
    "riemann3d_full": config(n=(24, 24, 24), wenothresh=-1.0, fixeddt=0.004),
    "riemann3d_nq3_full": config(n=(16, 16, 16), nq=3, wenothresh=-1.0, fixeddt=0.004),
    "riemann3d_hll_reflect": config(n=(16, 20, 12), riemann="hll", bc="reflecting", fixeddt=0.004),
    "riemann3d_llf_noslip_threestar": config(n=(16, 16, 16), riemann="llf", bc="noslip",
                                              rk="threestarp", fixeddt=0.004),
    "quadrant2d_sparse": config(dim=2, n=(48, 40, 1), problem="quadrant2d", seed=7, user=(),
                                fixeddt=0.002),
    "quadrant2d_full_nq3": config(dim=2, n=(32, 32, 1), nq=3, problem="quadrant2d", seed=3,
                                  user=(), wenothresh=-1.0, fixeddt=0.002),
    "shear2d_periodic": config(dim=2, n=(48, 48, 1), problem="shear2d", seed=1, user=(),
                               bc="periodic", fixeddt=0.002),
    # WENO cells on every periodic edge (the discontinuity plane wraps), NQ=3
    # and a grid that is no multiple of the bricks
    "riemann3d_periodic_nq3_sparse": config(n=(18, 13, 10), nq=3, bc="periodic", wenothresh=1.0,
                                            fixeddt=0.002),
    "linadv2d_full_nq3": config(dim=2, n=(32, 32, 1), nq=3, equation="linadv", riemann="exact",
                                problem="linadv_sine", user=(1.0, 1.0, 0.0), bc="periodic",
                                wenothresh=-1.0, advvel=(1.0, 1.0, 0.0), fixeddt=0.01),
    # the reference's LinAdv velocity (cos gamma, sin gamma) from eos:gamma
    # (LinAdv_Riemann.H:30,76), no z flux, with the DispRel problem
    "disprel_default_velocity": config(n=(16, 12, 8), equation="linadv", riemann="exact", problem="disprel",
                                       user=(0.25, 2.0), bc="periodic", gamma=0.3, wenothresh=-1.0,
                                       fixeddt=0.01),
    # the two remaining RK tables (Numeric.H:163-282)
    "riemann3d_fourstarp": config(n=(16, 16, 16), rk="fourstarp", fixeddt=0.002),
    "quadrant2d_ssp10": config(dim=2, n=(32, 32, 1), problem="quadrant2d", seed=5, user=(), rk="ssp10_4_3",
                               fixeddt=0.004),
    "linadv3d_sparse": config(n=(16, 16, 16), equation="linadv", riemann="exact",
                              problem="linadv_sine", user=(1.0, 1.0, 1.0), bc="periodic",
                              advvel=(1.0, 0.5, -0.25), fixeddt=0.01),
    # GLM-MHD: interior source points, cleaning speed, KEPES / HLL / LLF
    "mhd_vortex3d_llf_sparse": config(n=(16, 16, 16), equation="mhd_glm", riemann="llf",
                                      problem="mhd_vortex3d", user=(), bc="periodic",
                                      gamma=5.0 / 3.0, fixeddt=0.002),
    "mhd_vortex3d_kepes_full": config(n=(16, 16, 16), equation="mhd_glm", riemann="kepes",
                                      problem="mhd_vortex3d", user=(), bc="periodic",
                                      gamma=5.0 / 3.0, wenothresh=-1.0, fixeddt=0.002),
    "mhd_vortex3d_nq3_hll": config(n=(12, 12, 12), nq=3, equation="mhd_glm", riemann="hll",
                                   problem="mhd_vortex3d", user=(), bc="periodic",
                                   gamma=5.0 / 3.0, fixeddt=0.002),
    "orszag_tang2d_hll": config(dim=2, n=(48, 48, 1), equation="mhd_glm", riemann="hll",
                                problem="orszag_tang", user=(), bc="periodic", gamma=5.0 / 3.0,
                                fixeddt=0.002),
    "orszag_tang2d_kepes_nq3": config(dim=2, n=(32, 32, 1), nq=3, equation="mhd_glm",
                                      riemann="kepes", problem="orszag_tang", user=(),
                                      bc="periodic", gamma=5.0 / 3.0, fixeddt=0.002),
    "mhd_blast2d_reflect": config(dim=2, n=(40, 40, 1), equation="mhd_glm", riemann="llf",
                                  problem="mhd_blast", user=(0.1, 0.70710678118654752, 0.70710678118654752),
                                  bc="reflecting", gamma=5.0 / 3.0, fixeddt=0.0005),
    "mhd_blast3d_kepes_outflow": config(n=(16, 16, 16), equation="mhd_glm", riemann="kepes",
                                        problem="mhd_blast", user=(0.2, 0.5, 0.0), bc="outflow",
                                        gamma=5.0 / 3.0, fixeddt=0.0005),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_first_rhs_and_one_step(name):
    c = CASES[name]
    s, o = _pair(c)
    _start_oracle(s, o, c)
    # first RHS (the FSAL prime) and the flags it leaves
    K_gpu, K_cpu = s.rhs, o.rhs
    scale = max(np.abs(K_cpu).max(), 1e-300)
    rk = np.abs(K_gpu - K_cpu).max() / scale
    print(f"[parity] {name}: first RHS max diff / max|RHS| = {rk:.3e}", flush=True)
    assert rk <= 1e-12
    fg, fc = s.flags, o.flags
    assert np.array_equal(fg[0], fc[0]), f"WENO flag flips: {int((fg[0] != fc[0]).sum())}"
    assert np.abs(fg[1] - fc[1]).max() <= 1e-12
    # one fixed-dt step
    st_g = s.take_step()
    st_c = o.take_step()
    assert st_g == st_c == 0
    r = cell_rel_diff(s.state, o.state)
    flips = int((s.flags[0] != o.flags[0]).sum())
    _report(name, r, f"flag flips {flips}")
    assert flips == 0
    assert r.max() <= TOL_STEP
    tg, dtg, ng = s.time
    tc, dtc, nc = o.time
    assert tg == tc and ng == nc


def test_three_adaptive_steps_match():
    """Adaptive dt: the controller sees the same error norm up to summation order."""
    c = config(n=(16, 16, 16), wenothresh=1.0)
    s, o = _pair(c)
    _start_oracle(s, o, c)
    t0g, dt0g, _ = s.time
    o.set_time(0.0, dt0g)
    for _ in range(3):
        assert s.take_step() == o.take_step()
    tg, dtg, _ = s.time
    tc, dtc, _ = o.time
    assert abs(tg - tc) <= 1e-12 * tc
    assert abs(dtg - dtc) <= 1e-10 * dtc
    r = cell_rel_diff(s.state, o.state)
    _report("adaptive 3 steps", r)
    assert r.max() <= 1e-10


def test_unphysical_rejection_forces_full_weno():
    """cfl far too large: the step is rejected as unphysical, retried with
    full WENO, exactly like the reference's state machine."""
    c = config(n=(16, 16, 16), fixeddt=None, cfl=50.0)
    s, o = _pair(c)
    _start_oracle(s, o, c)
    _, dt, _ = s.time
    s.set_time(0.0, 0.2)
    o.set_time(0.0, 0.2)
    sg = s.take_step()
    sc = o.take_step()
    assert sg == sc
    cnt = s.counters
    assert cnt["rejects_unphys"] >= 1
    r = cell_rel_diff(s.state, o.state)
    _report("unphysical retry", r)
    assert r.max() <= 1e-10


FUSED_CASES = ["riemann3d_sparse", "riemann3d_periodic_nq3_sparse", "quadrant2d_sparse", "shear2d_periodic",
               "riemann3d_hll_reflect", "riemann3d_llf_noslip_threestar"]


@pytest.mark.parametrize("name", FUSED_CASES)
def test_fused_path_equals_unfused_path(name, monkeypatch):
    """The fused path (linear pass storing only brick-boundary states, rows for
    the WENO and limited cells, reconstruction recomputed inside the flux
    kernel) does the same arithmetic as the unfused one (every face state
    through HBM): three fixed-dt steps agree bit for bit, flags included."""
    import kfvm_amd
    c = CASES[name]
    text = init_text(c)
    monkeypatch.setenv("KFVM_FUSED", "0")
    ref = kfvm_amd.Solver(init_text=text)
    monkeypatch.setenv("KFVM_FUSED", "1")
    fus = kfvm_amd.Solver(init_text=text)
    for _ in range(3):
        assert ref.take_step() == fus.take_step() == 0
    d = np.abs(fus.state - ref.state).max()
    nw_ref, nw_fus = ref.counters["last_nweno"], fus.counters["last_nweno"]
    print(f"[fused] {name}: max |fused - unfused| {d:.3e}, WENO cells {nw_fus} / {nw_ref}", flush=True)
    assert d == 0.0
    assert np.array_equal(fus.flags, ref.flags)


@pytest.mark.parametrize("name", ["riemann3d_sparse", "riemann3d_full", "riemann3d_hll_reflect"])
def test_matrix_core_weno_matches_scalar_weno(name, monkeypatch):
    """3D NQ=2 runs KFVM-WENO on the FP64 matrix cores (k_wenorec5_mfma); the
    scalar kernel (KFVM_WENO_SCALAR=1) sums the same products in the
    reference's order.  Three fixed-dt steps of each, every launch of one run
    on the matrix cores and of the other on the scalar kernel, agree to
    rounding (per-cell relative difference <= 1e-13) with the same WENO
    flags; the limiter's theta, computed from the WENO states, to 1e-12."""
    import kfvm_amd
    c = CASES[name]
    text = init_text(c)
    monkeypatch.setenv("KFVM_WENO_SCALAR", "1")
    ref = kfvm_amd.Solver(init_text=text)
    monkeypatch.setenv("KFVM_WENO_SCALAR", "0")
    mat = kfvm_amd.Solver(init_text=text)
    for _ in range(3):
        monkeypatch.setenv("KFVM_WENO_SCALAR", "1")
        assert ref.take_step() == 0
        monkeypatch.setenv("KFVM_WENO_SCALAR", "0")
        assert mat.take_step() == 0
    r = cell_rel_diff(mat.state, ref.state)
    nw = mat.counters["last_nweno"]
    fm, fr = mat.flags, ref.flags
    dth = np.abs(fm[1] - fr[1]).max()
    print(f"[weno] {name}: matrix-core vs scalar WENO max cell rel diff {r.max():.3e}, max |d theta| "
          f"{dth:.3e}, WENO cells {nw}", flush=True)
    assert nw > 0
    assert r.max() <= 1e-13
    assert np.array_equal(fm[0], fr[0])
    assert dth <= 1e-12
