"""Pinning of the learned indicator's CPU restatement (oracle/indicator_ref.py):
the backward pass against central finite differences of the forward pass, and
features / labels against cells computed by hand."""
import numpy as np
import pytest

import indicator_ref as ir
from test_weights import stencil2d, stencil3d


@pytest.mark.parametrize("dim", [2, 3])
def test_backward_matches_finite_differences(dim):
    rng = np.random.default_rng(dim)
    nin, npar = ir.net_sizes(dim)
    P = rng.normal(0.0, 0.3, npar)
    X = rng.uniform(-1, 1, (40, nin))
    Y = (rng.uniform(size=40) < 0.3).astype(float)
    loss, g = ir.loss_and_grad(P, X, Y, dim, pos_weight=3.0)
    idx = rng.choice(npar, 60, replace=False)
    idx = np.concatenate([idx, [npar - 1, npar - 2, nin * 32]])  # b3, w3, b1[0]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (ir.loss_and_grad(Pp, X, Y, dim, 3.0)[0] - ir.loss_and_grad(Pm, X, Y, dim, 3.0)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the loss is the class-weighted logistic loss of the forward logits
    z = ir.forward(P, X, dim)
    ref = np.mean(np.where(Y > 0.5, 3.0 * np.log1p(np.exp(-z)), np.log1p(np.exp(z))))
    assert abs(loss - ref) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_features_of_a_density_ramp(dim):
    """rho = 1 + 0.1 x_index, uniform velocity and pressure, periodic across:
    the density features are the stencil's x offsets / 2 (the ramp's largest
    step is two cells), the pressure features vanish, the relative range is
    0.2 / rho_o."""
    off = (stencil3d() if dim == 3 else stencil2d())[0]
    nx, ny, nz = 9, 5, (4 if dim == 3 else 1)
    g = 1.4
    xi = np.arange(nx, dtype=float)
    rho = np.broadcast_to(1.0 + 0.1 * xi, (nz, ny, nx))
    U = np.zeros((5, nz, ny, nx))
    U[0] = rho
    U[1] = 0.3 * rho
    U[4] = 1.0 / (g - 1) + 0.5 * 0.09 * rho
    bc = ["outflow", "outflow", "periodic", "periodic", "periodic", "periodic"]
    X = ir.features(U, off, g, bc).reshape(nz, ny, nx, -1)
    C = len(off)
    o = off.index((0, 0, 0))
    others = [a for n, a in enumerate(off) if n != o]
    cell = X[nz // 2, 2, 4]
    np.testing.assert_allclose(cell[:C - 1], [a[0] / 2.0 for a in others], rtol=0, atol=1e-6)
    assert np.all(np.abs(cell[C - 1:2 * C - 2]) < 1e-6)
    assert abs(cell[2 * C - 2] - 0.2 / 1.4) < 1e-6 and abs(cell[2 * C - 1]) < 1e-6


@pytest.mark.parametrize("dim", [2, 3])
def test_mhd_features_see_the_gas_pressure(dim):
    """GLM-MHD states whose field and cleaning scalar carry energy
    (E = eint + |m|^2/(2 rho) + |B|^2/2 + psi^2/2) give the features of the
    Euler state with the same density, momenta and internal energy."""
    off = (stencil3d() if dim == 3 else stencil2d())[0]
    nx, ny, nz = 8, 6, (5 if dim == 3 else 1)
    g = 5.0 / 3.0
    rng = np.random.default_rng(7 + dim)
    rho = rng.uniform(0.5, 2.0, (nz, ny, nx))
    m = rng.uniform(-1.0, 1.0, (3, nz, ny, nx))
    B = rng.uniform(-1.0, 1.0, (3, nz, ny, nx))
    psi = rng.uniform(-0.1, 0.1, (nz, ny, nx))
    p = rng.uniform(0.2, 3.0, (nz, ny, nx))
    ek = 0.5 * (m ** 2).sum(0) / rho
    H = np.concatenate([rho[None], m, (p / (g - 1) + ek)[None]])
    M = np.concatenate([rho[None], m, (p / (g - 1) + ek + 0.5 * (B ** 2).sum(0) + 0.5 * psi ** 2)[None],
                        B, psi[None]])
    bc = ["periodic"] * 6
    np.testing.assert_allclose(ir.features(M, off, g, bc), ir.features(H, off, g, bc), rtol=0, atol=2e-5)


def test_labels_of_hand_made_face_states():
    """2D, two cells differ: a WENO face state moved by 0.3 of the stencil range
    labels its cell at tau = 0.25 and not at tau = 0.35."""
    off = stencil2d()[0]
    nz, ny, nx = 1, 6, 6
    U = np.ones((5, nz, ny, nx))
    U[0, 0, 2, 3] = 2.0           # density range 1 in the stencils that see it
    FSL = np.zeros((8, 5, nz, ny, nx))
    FSW = FSL.copy()
    FSW[3, 0, 0, 2, 2] = 0.3      # cell (x=2, y=2): its stencil contains (3, 2)
    FSW[5, 4, 0, 4, 4] = 1e-3     # energy range 0 there: the floor 1e-3 |E| = 1e-3 is the scale
    bc = ["periodic"] * 6
    y25 = ir.labels(U, FSL, FSW, off, bc, 0.25).reshape(ny, nx)
    y35 = ir.labels(U, FSL, FSW, off, bc, 0.35).reshape(ny, nx)
    assert y25[2, 2] == 1 and y35[2, 2] == 0
    assert y25[4, 4] == 1 and y35[4, 4] == 1          # ratio 1
    assert y25.sum() == 2 and y35.sum() == 1
    FSW[5, 4, 0, 4, 4] = 2e-4                          # ratio 0.2: below both thresholds
    assert ir.labels(U, FSL, FSW, off, bc, 0.25).reshape(ny, nx)[4, 4] == 0
