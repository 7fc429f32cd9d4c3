"""Pins of the oracle functions that convergence and symmetry tests leave open
(no GPU).  Each test names the plausible mistake it catches.

  * KXRCF selector (Numeric_K.H:292-398, threshold Solver.cpp:425): a
    brute-force numpy evaluation of the indicator from the weight table and
    an explicit outflow mirror, with the threshold placed between indicator
    values so that the dmin^1.5 scaling decides flags;
  * positivity limiter (Common_K.H:299-643): the invariants every limited
    state must meet, and a cell whose theta is known in closed form;
  * WENO-AO weights (Stencil_K.H:506-580): equal smoothness on linear data
    reduces the scheme to the full-stencil linear reconstruction, and a
    one-sided jump hands all weight to the one smooth biased substencil;
  * error norm (Numeric_K.H:262-290): hand-made states with the sum and the
    positivity written out.
"""
import math

import numpy as np
import pytest

import kfvm_amd
from helpers import config, init_text, oracle_for
from test_weights import stencil3d

GAMMA = 1.4


def _oracle(c):
    text = init_text(c)
    p = kfvm_amd.parse_init(text)
    n = p["n"]
    dims = [(p["hi"][a] - p["lo"][a]) / n[a] for a in range(3)]
    fw, dw = kfvm_amd.build_weights(p["dim"], p["nq"], p["lfac"], *dims)
    return oracle_for(c, p, fw, dw, threads=2), p, fw, dims


def _mirror(i, n):
    """Outflow ghost cell -> the interior cell it copies (BoundaryConditions_K.H:65-69)."""
    return -i - 1 if i < 0 else (2 * n - 1 - i if i >= n else i)


def _prim2cons(rho, u, v, w, p):
    return np.array([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


def _linear_faces(U, fw, nq):
    """Brute-force linear face states [F, Q, NV, nz, ny, nx] of a 3D state with
    outflow walls: each cell's stencil gathered through the mirror map and
    contracted with the full-stencil weights faceW[0][f][q][:]."""
    off, _ = stencil3d()
    nv, nz, ny, nx = U.shape
    F, Q, C = 6, nq * nq, 33
    W = np.asarray(fw).reshape(8, F, Q, C)[0]
    out = np.zeros((F, Q, nv, nz, ny, nx))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                S = np.array([U[:, _mirror(k + o[2], nz), _mirror(j + o[1], ny), _mirror(i + o[0], nx)]
                              for o in off])  # [C, NV]
                out[:, :, :, k, j, i] = np.einsum("fqc,cv->fqv", W, S)
    return out


def _pres(s):
    return (GAMMA - 1.0) * (s[4] - 0.5 * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2) / s[0])


def _eint(s):
    return s[4] - 0.5 * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2) / s[0]


def _kxrcf(U, fw, nq):
    """max over face points of |w_out - w_in| / w_cell, w = p / rho^gamma;
    across an outflow wall the outer state is the inner one (jump 0)."""
    fs = np.moveaxis(_linear_faces(U, fw, nq), 2, 0)  # [NV, F, Q, nz, ny, nx]
    w = _pres(fs) / fs[0] ** GAMMA  # [F, Q, nz, ny, nx]
    nv, nz, ny, nx = U.shape
    ind = np.zeros((nz, ny, nx))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                m = 0.0
                for f in range(6):
                    ax, hi = f // 2, f % 2
                    d = [0, 0, 0]
                    d[ax] = 1 if hi else -1
                    ni, nj, nk = i + d[0], j + d[1], k + d[2]
                    inside = 0 <= ni < nx and 0 <= nj < ny and 0 <= nk < nz
                    for q in range(nq * nq):
                        win = w[f, q, k, j, i]
                        wout = w[f ^ 1, q, nk, nj, ni] if inside else win
                        m = max(m, abs(wout - win))
                ind[k, j, i] = m / (_pres(U[:, k, j, i]) / U[0, k, j, i] ** GAMMA)
    return ind


def _step_state(n, profile):
    """3D state varying along x only: profile(i) -> (rho, u, v, w, p)."""
    U = np.zeros((5, n, n, n))
    for i in range(n):
        U[:, :, :, i] = _prim2cons(*profile(i))[:, None, None]
    return U


def test_kxrcf_selector_brute_force():
    """Flags = (indicator >= wenothresh * dmin^1.5) or (previous theta < 1).
    The threshold sits between two indicator values, so a missing or wrong
    power of dmin, a missing division by the cell's own p/rho^gamma, a missing
    absolute value or a wrong neighbour face / point pairing flips flags."""
    n = 6
    c = config(n=(n, n, n), wenothresh=1.0, problem="uniform", user=(1, 0, 0, 0, 1))
    o, p, fw, dims = _oracle(c)
    # a smooth pressure bump plus a density jump, so indicator values spread out
    U = _step_state(n, lambda i: (1.0 if i < 3 else 0.4, 0.1 * i, 0.0, 0.0, 1.0 + 0.05 * i * i))
    U[0, :, 2, :] *= 1.1  # break the y symmetry
    ind = _kxrcf(U, fw, 2)
    vals = np.unique(np.round(ind[ind > 1e-12], 12))
    assert vals.size >= 4
    thr_eff = 0.5 * (vals[vals.size // 2 - 1] + vals[vals.size // 2])
    dmin = min(dims)
    wenothresh = thr_eff / dmin ** 1.5
    c2 = config(n=(n, n, n), wenothresh=wenothresh, problem="uniform", user=(1, 0, 0, 0, 1))
    o, p, fw, dims = _oracle(c2)
    # the library stores wenothresh with std::stof semantics: use what it parsed
    thr = p["wenothresh"] * min(dims) ** 1.5
    expect = ind >= thr
    assert 0 < expect.sum() < expect.size
    o.state = U
    fl = np.zeros((2, n, n, n))
    fl[1] = 1.0
    fl[1, 5, 5, 5] = 0.5  # limited by the previous RHS -> flagged whatever its indicator
    assert ind[5, 5, 5] < thr
    expect[5, 5, 5] = True
    o.flags = fl
    o.eval_rhs(0)
    got = o.flags[0] == 1.0
    assert np.array_equal(got, expect), f"{int((got != expect).sum())} flag differences"


def test_positivity_limiter_invariants_and_closed_form_theta():
    """A density step 1 -> 1e-3 at rest with uniform total energy: linear
    face states undershoot next to the step.  Catches a wrong widening
    (kappa), a wrong bound neighbourhood, theta from the wrong side, or
    mixing that does not go towards the cell average.
      * theta = 1 cells keep their states bitwise;
      * limited states stay within [minDens, maxDens] (widened by kappa = 0.2
        at rest) and keep internal energy above the widened minimum;
      * with E uniform and no momentum, e_int = E at every point, so
        theta = theta_rho = min over points below minDens of
        (rho_c - minDens) / (rho_c - rho_f), in closed form."""
    n = 8
    c = config(n=(n, n, n), wenothresh=1.0, problem="uniform", user=(1, 0, 0, 0, 1))
    o, p, fw, dims = _oracle(c)
    eps = 1e-3
    U = _step_state(n, lambda i: (1.0 if i < 4 else eps, 0.0, 0.0, 0.0, 1.0))
    o.state = U
    lin = o.face_states(weno=False)  # [F*Q, NV, nz, ny, nx]
    lim = o.face_states_limited()
    fl = np.zeros((2, n, n, n))
    fl[1] = 1.0
    o.flags = fl
    o.eval_rhs(0)
    # sparse mode: WENO overwrites flagged cells' states; the limiter's theta
    # of unflagged cells is that of the linear states
    theta = o.flags[1]
    wflag = o.flags[0]
    FQ = lin.shape[0]
    kappa = 0.2
    checked = closed = 0
    for k in range(n):
        for j in range(n):
            for i in range(n):
                rho_c = U[0, k, j, i]
                nb = U[0, max(k - 1, 0):k + 2, max(j - 1, 0):j + 2, max(i - 1, 0):i + 2]
                # outflow mirror: the ghost copies the edge cell, inside the same range
                mind, maxd = nb.min() * (1 - kappa), nb.max() * (1 + kappa)
                mine = (1.0 / (GAMMA - 1.0)) * (1 - kappa)
                L = lin[:, :, k, j, i]
                M = lim[:, :, k, j, i]
                th_rho = 1.0
                for fq in range(FQ):
                    r = L[fq, 0]
                    if not r > mind:
                        th_rho = min(th_rho, (rho_c - mind) / (rho_c - r))
                    if not r < maxd:
                        th_rho = min(th_rho, (maxd - rho_c) / (r - rho_c))
                if th_rho == 1.0:
                    assert np.array_equal(M, L)
                else:
                    checked += 1
                    assert np.all(M[:, 0] >= mind * (1 - 1e-12)) and np.all(M[:, 0] <= maxd * (1 + 1e-12))
                    assert np.all(_eint(M.T) > mine * (1 - 1e-12))
                    # the limited states are the convex mix towards the cell average
                    ref = U[:, k, j, i]
                    np.testing.assert_allclose(M, ref + th_rho * (L - ref), rtol=0, atol=1e-13)
                    if wflag[k, j, i] == 0.0:
                        closed += 1
                        assert abs(theta[k, j, i] - th_rho) <= 1e-12
    assert checked >= 2 * n * n and closed >= n * n


def test_weno_equal_smoothness_reduces_to_linear():
    """Density linear in x, y, z at rest: every substencil reproduces the
    linear field, so every beta_s is the same, omega = (1, 0, ..., 0) after
    the WENO-AO transform, and the WENO face states equal the full-stencil
    linear ones to round-off (cells whose stencils stay inside the box).
    Catches a wrong linear weight gamma, a beta not squared twice, a wrong
    normalisation or AO transform, or a bad linearised-primitive round trip."""
    n = 10
    c = config(n=(n, n, n), wenothresh=-1.0, problem="uniform", user=(1, 0, 0, 0, 1))
    o, p, fw, dims = _oracle(c)
    xs = (np.arange(n) + 0.5) / n
    rho = 1.0 + 0.3 * xs[None, None, :] + 0.2 * xs[None, :, None] - 0.1 * xs[:, None, None]
    U = np.zeros((5, n, n, n))
    U[0] = rho
    U[4] = 2.5
    o.state = U
    lin = o.face_states(weno=False)
    wen = o.face_states(weno=True)
    inner = (slice(None), slice(None), slice(2, n - 2), slice(2, n - 2), slice(2, n - 2))
    scale = np.abs(lin[inner]).max()
    assert np.abs(wen[inner] - lin[inner]).max() <= 1e-13 * scale
    for cell in [(2, 3, 4), (5, 5, 5), (7, 2, 6)]:
        om = o.weno_omegas(*cell)
        expect = np.zeros_like(om)
        expect[0] = 1.0
        np.testing.assert_allclose(om, expect, rtol=0, atol=1e-12)


def test_weno_one_sided_jump_picks_smooth_biased_substencil():
    """Density jump between x-cells 4 and 5 at rest, uniform pressure.  For
    cell 4 only the west-biased substencil lies on one side of the jump:
    its weight in the density variable is 1 and every other substencil's 0
    (its beta is exactly 0 on constant data); variables without a jump keep
    omega = (1, 0, ..., 0).  Catches a wrong substencil index map or a
    swapped west/east orientation."""
    n = 10
    c = config(n=(n, n, n), wenothresh=-1.0, problem="uniform", user=(1, 0, 0, 0, 1))
    o, p, fw, dims = _oracle(c)
    U = _step_state(n, lambda i: (1.0 if i <= 4 else 0.25, 0.0, 0.0, 0.0, 1.0))
    o.state = U
    om = o.weno_omegas(4, 5, 5)  # substencils: full, centre, W, E, S, N, B, T
    assert abs(om[2, 0] - 1.0) <= 1e-12
    assert np.abs(np.delete(om[:, 0], 2)).max() <= 1e-12
    expect = np.zeros((8, 4))
    expect[0] = 1.0
    np.testing.assert_allclose(om[:, 1:], expect, rtol=0, atol=1e-12)
    # the mirror image: cell 5 is smooth only towards the east
    om = o.weno_omegas(5, 5, 5)
    assert abs(om[3, 0] - 1.0) <= 1e-12
    assert np.abs(np.delete(om[:, 0], 3)).max() <= 1e-12


def test_error_norm_closed_form():
    """Two Euler cells with the sum written out by hand:
    err = sum ((u - uh) / (atol + rtol max(|u|, |uh|)))^2, positivity =
    min(rho, e_int), -1 for a NaN.  Catches max(|u|, |uh|) taken as |u|
    only, atol / rtol swapped, a missing square or a wrong positivity."""
    from orc import error_norm
    atol, rtol = 1e-3, 1e-2
    u = np.array([[1.0, 0.5, 0.0, 0.0, 3.0], [0.2, -0.1, 0.0, 0.0, 0.5]])
    uh = np.array([[1.1, 0.5, 0.0, 0.0, 2.9], [0.2, -0.3, 0.0, 0.0, 0.5]])
    terms = [
        (1.0 - 1.1) / (atol + rtol * 1.1), 0.0, 0.0, 0.0, (3.0 - 2.9) / (atol + rtol * 3.0),
        0.0, (-0.1 + 0.3) / (atol + rtol * 0.3), 0.0, 0.0, 0.0,
    ]
    e, pos = error_norm("hydro", u, uh, atol, rtol)
    assert abs(e - sum(t * t for t in terms)) <= 1e-13 * e
    eint = [3.0 - 0.5 * 0.25 / 1.0, 0.5 - 0.5 * 0.01 / 0.2]
    assert pos == min(1.0, eint[0], 0.2, eint[1])
    u[1, 4] = np.nan
    assert error_norm("hydro", u, uh, atol, rtol)[1] == -1.0
    # linear advection: positivity is 1, only the norm counts
    e1, p1 = error_norm("linadv", np.array([[2.0]]), np.array([[1.0]]), 0.5, 0.25)
    assert e1 == pytest.approx((1.0 / (0.5 + 0.25 * 2.0)) ** 2, rel=1e-15) and p1 == 1.0
