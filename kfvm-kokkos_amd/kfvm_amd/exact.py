"""Exact solution of the Riemann problem of the ideal-gas Euler equations
(Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, ch. 4),
vectorised over sample points: the error measure of the planar Riemann
problems (python/solErrorL1.py in the reference computes L1 errors the same
way, as the mean absolute difference over cells)."""
import math

import numpy as np


def _f(p, rk, pk, ck, g):
    if p > pk:
        A, B = 2.0 / ((g + 1.0) * rk), (g - 1.0) / (g + 1.0) * pk
        return (p - pk) * math.sqrt(A / (p + B))
    return 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)


def riemann(left, right, xi, t, gamma=1.4):
    """(rho, u, p) at signed distances xi from the initial jump at time t > 0;
    left / right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    g = gamma
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    lo, hi = 1e-12, 10.0 * max(pl, pr) + 1.0
    for _ in range(300):
        m = 0.5 * (lo + hi)
        if _f(m, rl, pl, cl, g) + _f(m, rr, pr, cr, g) + ur - ul > 0.0:
            hi = m
        else:
            lo = m
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (_f(ps, rr, pr, cr, g) - _f(ps, rl, pl, cl, g))
    gm = (g - 1.0) / (g + 1.0)
    s = np.asarray(xi, dtype=float) / t
    rho = np.empty_like(s)
    u = np.empty_like(s)
    p = np.empty_like(s)
    # left of the contact
    L = s < us
    if ps > pl:  # left shock
        rsl = rl * (ps / pl + gm) / (gm * ps / pl + 1.0)
        SL = ul - cl * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
        a = L & (s < SL)
        b = L & ~(s < SL)
        rho[a], u[a], p[a] = rl, ul, pl
        rho[b], u[b], p[b] = rsl, us, ps
    else:        # left rarefaction
        rsl = rl * (ps / pl) ** (1.0 / g)
        csl = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
        a = L & (s < ul - cl)
        fan = L & (s >= ul - cl) & (s < us - csl)
        b = L & (s >= us - csl)
        rho[a], u[a], p[a] = rl, ul, pl
        uf = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + s[fan])
        cf = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - s[fan]))
        rho[fan], u[fan], p[fan] = rl * (cf / cl) ** (2.0 / (g - 1.0)), uf, pl * (cf / cl) ** (2.0 * g / (g - 1.0))
        rho[b], u[b], p[b] = rsl, us, ps
    R = ~L
    if ps > pr:  # right shock
        rsr = rr * (ps / pr + gm) / (gm * ps / pr + 1.0)
        SR = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
        a = R & (s >= SR)
        b = R & ~(s >= SR)
        rho[a], u[a], p[a] = rr, ur, pr
        rho[b], u[b], p[b] = rsr, us, ps
    else:        # right rarefaction
        rsr = rr * (ps / pr) ** (1.0 / g)
        csr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
        a = R & (s >= ur + cr)
        fan = R & (s < ur + cr) & (s >= us + csr)
        b = R & (s < us + csr)
        rho[a], u[a], p[a] = rr, ur, pr
        uf = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + s[fan])
        cf = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - s[fan]))
        rho[fan], u[fan], p[fan] = rr * (cf / cr) ** (2.0 / (g - 1.0)), uf, pr * (cf / cr) ** (2.0 * g / (g - 1.0))
        rho[b], u[b], p[b] = rsr, us, ps
    return rho, u, p


def planar_riemann_l1(rho, t, offset=1.5, gamma=1.4):
    """L1 density error of the 3D planar Riemann problem across x+y+z = offset
    on [0,1]^3 ((1,0,0,0,1) / (0.125,0,0,0,0.1)); rho [nz, ny, nx] at cell
    centres; the exact solution is Sod's along the plane normal."""
    nz, ny, nx = rho.shape
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx,
                          indexing="ij")
    xi = (x + y + z - offset) / math.sqrt(3.0)
    ex, _, _ = riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), xi.ravel(), t, gamma)
    return float(np.abs(rho.ravel() - ex).mean())
