"""Pins of the flow-statistics restatement (helpers.flow_stats_ref, the
checker of kfvm_flow_stats) against closed forms, no GPU.

With point values at the cell centres of a periodic grid of spacing h:
  * u = (sin 2 pi y, 0, 0), rho = 1: kinetic energy 1/4 exactly (the mean of
    sin^2 over n >= 3 equispaced points is 1/2), enstrophy
    (sin(2 pi h) / h)^2 / 4 (the central difference of sin is
    cos(2 pi y) sin(2 pi h) / h) -- catches a dropped 1/2, a wrong spacing, a
    forward instead of central difference, or weights that are not dvol/volume;
  * B = (0, sin 2 pi x, cos 2 pi x), u = 0: helicity B . curl B =
    sin(2 pi h) / h in every cell -- catches a sign or index error in the curl;
  * the same fields rotated onto other axes give the same values (each
    vorticity / curl component is exercised);
  * reflecting walls with a uniform flow normal to them: the mirrored ghost
    momentum makes the normal velocity's tangential differences vanish, the
    statistics are those of the interior only.
"""
import math

import numpy as np
import pytest

from helpers import flow_stats_ref


def _grid(n, dim):
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    if dim == 3:
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    else:
        Y, X = np.meshgrid(x, x, indexing="ij")
        Z = np.zeros_like(X)
        X, Y, Z = X[None], Y[None], Z[None]
    return h, X, Y, Z


@pytest.mark.parametrize("dim,n", [(3, 8), (3, 13), (2, 16)])
def test_shear_flow_kinetic_energy_and_enstrophy(dim, n):
    h, X, Y, Z = _grid(n, dim)
    for axis_u, coord in ((1, Y), (2, X)):  # u_x(y) and u_y(x)
        U = np.zeros((5,) + X.shape)
        U[0] = 1.0
        U[axis_u] = np.sin(2 * math.pi * coord)
        U[4] = 3.0
        s = flow_stats_ref(U, dim, (h, h, h), 1.0, "periodic", "hydro")
        assert abs(s[0] - 0.25) < 1e-14
        assert abs(s[1] - 0.25 * (math.sin(2 * math.pi * h) / h) ** 2) < 1e-12
    if dim == 3:  # u_z(x): the omega_y component
        U = np.zeros((5,) + X.shape)
        U[0] = 1.0
        U[3] = np.sin(2 * math.pi * X)
        s = flow_stats_ref(U, dim, (h, h, h), 1.0, "periodic", "hydro")
        assert abs(s[1] - 0.25 * (math.sin(2 * math.pi * h) / h) ** 2) < 1e-12


@pytest.mark.parametrize("dim,n", [(3, 8), (3, 11)])
def test_force_free_field_helicity(dim, n):
    h, X, Y, Z = _grid(n, dim)
    U = np.zeros((9,) + X.shape)
    U[0] = 1.0
    U[4] = 2.0
    U[6] = np.sin(2 * math.pi * X)
    U[7] = np.cos(2 * math.pi * X)
    s = flow_stats_ref(U, dim, (h, h, h), 1.0, "periodic", "mhd_glm")
    assert abs(s[0]) < 1e-15 and abs(s[1]) < 1e-15
    assert abs(s[2] - math.sin(2 * math.pi * h) / h) < 1e-12
    # rotated: B = (cos 2 pi y, 0, sin 2 pi y)
    U = np.zeros((9,) + X.shape)
    U[0] = 1.0
    U[5] = np.cos(2 * math.pi * Y)
    U[7] = np.sin(2 * math.pi * Y)
    s = flow_stats_ref(U, dim, (h, h, h), 1.0, "periodic", "mhd_glm")
    assert abs(abs(s[2]) - math.sin(2 * math.pi * h) / h) < 1e-12


def test_reflecting_walls_mirror_the_normal_velocity():
    """u = (1, 0, 0) between reflecting x walls: the ghost momentum is -1, so
    d(u)/dx is nonzero at the walls but u_x has no y / z variation: the
    enstrophy is 0 and the kinetic energy 1/2."""
    h, X, Y, Z = _grid(6, 3)
    U = np.zeros((5,) + X.shape)
    U[0] = 1.0
    U[1] = 1.0
    bc = ["reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic"]
    s = flow_stats_ref(U, 3, (h, h, h), 1.0, bc, "hydro")
    assert abs(s[0] - 0.5) < 1e-15 and s[1] == 0.0
