"""CPU restatement of the learned troubled-cell indicator's arithmetic.

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in
kfvm-kokkos_amd/csrc/device/indicator.hpp (features, labels, network forward /
backward, NAD count).  The reference has no learned indicator; what this
restates is this library's own definition of it.  Pinning: the backward pass
is checked against central finite differences of the forward pass in float64;
features and labels against hand-computed cells.
"""
import numpy as np

H = 32


def net_sizes(dim):
    C = 33 if dim == 3 else 13
    nin = 2 * C
    return nin, H * nin + H + H * H + H + H + 1


def unpack(P, dim):
    nin, npar = net_sizes(dim)
    P = np.asarray(P, dtype=np.float64)
    assert P.size == npar
    o = 0
    W1 = P[o:o + H * nin].reshape(H, nin); o += H * nin
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].reshape(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    w3 = P[o:o + H]; o += H
    b3 = P[o]
    return W1, b1, W2, b2, w3, b3


def forward(P, X, dim):
    """Logits of rows X [n, nin] in float64."""
    W1, b1, W2, b2, w3, b3 = unpack(P, dim)
    a1 = X.astype(np.float64) @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ w3 + b3


def loss_and_grad(P, X, Y, dim, pos_weight=1.0):
    """Mean class-weighted logistic loss and its gradient (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(P, dim)
    X = X.astype(np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n = len(Y)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    w = np.where(Y > 0.5, pos_weight, 1.0)
    m = np.where(Y > 0.5, -z, z)
    loss = np.mean(w * (np.maximum(m, 0.0) + np.log1p(np.exp(-np.abs(m)))))
    dz = w * (1.0 / (1.0 + np.exp(-z)) - Y) / n
    d2 = np.outer(dz, w3) * (a2 > 0)
    d1 = (d2 @ W2) * (a1 > 0)
    g = np.concatenate([(d1.T @ X).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0),
                        h2.T @ dz, [dz.sum()]])
    return loss, g


def _ghosted(U, bc, r=2):
    """U [NV, nz, ny, nx] padded by r ghost cells: periodic wrap or mirror copy
    (outflow; reflecting negates the normal momentum).  bc: 6 names."""
    dim = 3 if U.shape[1] > 1 else 2
    V = U
    for ax in range(dim):
        axis = 3 - ax  # x is the last array axis
        lo, hi = bc[2 * ax], bc[2 * ax + 1]
        pad = [(0, 0)] * 4
        pad[axis] = (r, r)
        if lo == "periodic":
            V = np.pad(V, pad, mode="wrap")
            continue
        V = np.pad(V, pad, mode="symmetric")
        for side, name in ((0, lo), (1, hi)):
            if name in ("reflecting", "noslip"):
                sl = [slice(None)] * 4
                sl[axis] = slice(0, r) if side == 0 else slice(-r, None)
                vs = [1 + ax] if name == "reflecting" else [1, 2, 3]
                for v in vs:
                    idx = tuple([v] + sl[1:])
                    V[idx] = -V[idx]
    return V


def _stencil_values(V, off, shape, dim, r=2):
    nz, ny, nx = shape
    out = []
    for (a, b, c) in off:
        zs = slice(r + c, r + c + nz) if dim == 3 else slice(0, 1)
        out.append(V[:, zs, r + b:r + b + ny, r + a:r + a + nx])
    return out


def _eint(s):
    """Internal energy per volume of the Euler (5 variables, Hydro_EOS.H) or
    GLM-MHD (9 variables, MHD_GLM_EOS.H) state s, in the device's operation
    order."""
    ek = 0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]) / s[0]
    if len(s) == 9:
        pb = 0.5 * (s[5] * s[5] + s[6] * s[6] + s[7] * s[7])
        return s[4] - pb - ek - 0.5 * s[8] * s[8]
    return s[4] - ek


def features(U, off, gamma, bc):
    """Features [ncell, 2C] (float32) of every cell, cells in x-fastest order."""
    dim = 3 if U.shape[1] > 1 else 2
    V = _ghosted(U, bc)
    S = _stencil_values(V, off, U.shape[1:], dim)
    o = off.index((0, 0, 0))
    gm1 = gamma - 1.0
    rho = [s[0] for s in S]
    pr = [gm1 * _eint(s) for s in S]
    others = [n for n in range(len(off)) if n != o]
    sr = np.zeros_like(rho[o])
    sp = np.zeros_like(rho[o])
    for n in others:
        sr = np.fmax(sr, np.abs(rho[n] - rho[o]))
        sp = np.fmax(sp, np.abs(pr[n] - pr[o]))
    ir = 1.0 / (sr + 1e-300)
    ip = 1.0 / (sp + 1e-300)
    cols = [(rho[n] - rho[o]) * ir for n in others] + [(pr[n] - pr[o]) * ip for n in others]
    cols += [np.fmin(sr / np.abs(rho[o]), 4.0), np.fmin(sp / np.abs(pr[o]), 4.0)]
    return np.stack([c.ravel() for c in cols], axis=1).astype(np.float32)


def labels(U, FSL, FSW, off, bc, tau):
    """Labels [ncell] from linear / WENO face states [F*Q, NV, nz, ny, nx]."""
    dim = 3 if U.shape[1] > 1 else 2
    V = _ghosted(U, bc)
    S = _stencil_values(V, off, U.shape[1:], dim)
    d = np.zeros(U.shape[1:])
    for v in (0, 4):
        vals = np.stack([s[v] for s in S])
        sc = vals.max(0) - vals.min(0) + 1e-3 * np.abs(U[v]) + 1e-300
        d = np.fmax(d, (np.abs(FSW[:, v] - FSL[:, v]) / sc).max(0))
    return (d > tau).astype(np.float32).ravel()
