"""Pinning of the GP reconstruction weights (host side, no GPU).

The weights solve the saddle system of ReconVectors.cpp:26-86.  They are
pinned here by what the mathematics fixes, independently of the solver:
  * the constraint rows: polynomial reproduction of cell averages of every
    monomial up to the substencil's degree (the reference's own testPV
    identities, Stencil.cpp:249-279, generalised), for face points and for
    derivative functionals (with the Monomials.cpp:29 argument order);
  * the kernel rows: K w - k(q) lies in the range of P (the residual of the
    first block row can be absorbed by the multipliers);
  * an independent high-precision (mpmath, 40 digits) solve of the same
    system, agreeing to the conditioning of the matrix;
  * exact (bitwise) stencil symmetry between faces.
"""
import math

import numpy as np
import pytest

import kfvm_amd

# ---------------------------------------------------------------- own stencils


def stencil3d():
    rm, rad = 1, 2
    off = []
    for i in range(-rm, rm + 1):
        for j in range(-rm, rm + 1):
            for k in range(-rm, rm + 1):
                if i * i + j * j + k * k <= rm * rm:
                    off.append((i, j, k))
    for i in range(-rad, rad + 1):
        for j in range(-rad, rad + 1):
            for k in range(-rad, rad + 1):
                r = i * i + j * j + k * k
                if rm * rm < r <= rad * rad:
                    off.append((i, j, k))
    sub = [[] for _ in range(6)]
    for n, (l, m, q) in enumerate(off):
        if -l >= abs(m) and -l >= abs(q): sub[0].append(n)
        if l >= abs(m) and l >= abs(q): sub[1].append(n)
        if -m >= abs(l) and -m >= abs(q): sub[2].append(n)
        if m >= abs(l) and m >= abs(q): sub[3].append(n)
        if -q >= abs(l) and -q >= abs(m): sub[4].append(n)
        if q >= abs(l) and q >= abs(m): sub[5].append(n)
    return off, sub


def stencil2d():
    off = [(0, 0, 0)]
    for nl in (1, 2):
        off += [(-nl, 0, 0), (nl, 0, 0), (0, -nl, 0), (0, nl, 0)]
        if nl == 2:
            off += [(-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0)]
    sub = [[off.index((-h, 0, 0)) for h in range(3)], [off.index((h, 0, 0)) for h in range(3)],
           [off.index((0, -h, 0)) for h in range(3)], [off.index((0, h, 0)) for h in range(3)]]
    sub[0] += [off.index((-1, -1, 0)), off.index((-1, 1, 0))]
    sub[1] += [off.index((1, -1, 0)), off.index((1, 1, 0))]
    sub[2] += [off.index((-1, -1, 0)), off.index((1, -1, 0))]
    sub[3] += [off.index((-1, 1, 0)), off.index((1, 1, 0))]
    return off, sub


def test_stencil_counts_brute_force():
    off, sub = stencil3d()
    assert len(off) == 33 and len(set(off)) == 33
    assert off[:7] == [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
    assert all(len(s) == 11 for s in sub)
    # brute force: lattice points of the radius-2 ball
    ball = {(i, j, k) for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)
            if i * i + j * j + k * k <= 4}
    assert set(off) == ball
    off2, sub2 = stencil2d()
    assert len(off2) == 13 and set(off2) == {(i, j, 0) for i in range(-2, 3) for j in range(-2, 3)
                                             if i * i + j * j <= 4}
    assert all(len(s) == 5 for s in sub2)


# ------------------------------------------------------------ monomial algebra

def monos(dim, tdeg):
    out = []
    for t in range(tdeg + 1):
        if dim == 2:
            for i in range(t + 1):
                out.append((i, t - i, 0))
        else:
            for i in range(t + 1):
                for j in range(t - i + 1):
                    out.append((i, j, t - i - j))
    return out


def avg1(p, x):
    return ((x + 0.5) ** (p + 1) - (x - 0.5) ** (p + 1)) / (p + 1)


def mono_avg(m, c, dim):
    v = avg1(m[0], c[0]) * avg1(m[1], c[1])
    return v * avg1(m[2], c[2]) if dim == 3 else v


def quirk_der0(p, N):
    """d^N x^p at 0 as the reference evaluates it (Monomials.cpp:29 order)."""
    if N == 0:
        return 1.0 if p == 0 else 0.0
    if N > p:
        return 0.0
    return p * float(N - 1) ** (p - 1)


def true_der0(p, N):
    return math.factorial(p) if N == p else 0.0


FACE_PTS = {
    0: lambda a, b: (-0.5, a, b), 1: lambda a, b: (0.5, a, b), 2: lambda a, b: (a, -0.5, b),
    3: lambda a, b: (a, 0.5, b), 4: lambda a, b: (a, b, -0.5), 5: lambda a, b: (a, b, 0.5)}


def face_points(dim, nq):
    ab = {2: [-0.577350269189625764509148 / 2, 0.577350269189625764509148 / 2],
          3: [0.0, -0.7745966692414833770358530 / 2, 0.7745966692414833770358530 / 2]}[nq]
    pts = {}
    for f in range(2 * dim):
        if dim == 3:
            pts[f] = [FACE_PTS[f](ab[a], ab[b]) for a in range(nq) for b in range(nq)]
        else:
            pts[f] = [FACE_PTS[f](ab[a], 0.0)[:2] + (0.0,) for a in range(nq)]
    return pts


def sub_cells(dim, s):
    off, sub = stencil3d() if dim == 3 else stencil2d()
    C, Cc = (33, 7) if dim == 3 else (13, 5)
    if s == 0:
        return list(range(C)), off
    if s == 1:
        return list(range(Cc)), off
    return sub[s - 2], off


@pytest.mark.parametrize("dim,nq", [(2, 2), (2, 3), (3, 2), (3, 3)])
def test_face_weights_reproduce_polynomials(dim, nq):
    fw, dw = kfvm_amd.build_weights(dim, nq, 5.0, 1.0 / 64, 1.0 / 64, 1.0 / 64)
    pts = face_points(dim, nq)
    nsub = fw.shape[0]
    worst = 0.0
    for s in range(nsub):
        cells, off = sub_cells(dim, s)
        deg = 3 if s == 0 else 1
        for f in range(2 * dim):
            for q, x in enumerate(pts[f]):
                w = fw[s, f, q, : len(cells)]
                assert np.all(fw[s, f, q, len(cells):] == 0.0)
                for m in monos(dim, deg):
                    lhs = sum(w[n] * mono_avg(m, off[c], dim) for n, c in enumerate(cells))
                    rhs = x[0] ** m[0] * x[1] ** m[1] * (x[2] ** m[2] if dim == 3 else 1.0)
                    worst = max(worst, abs(lhs - rhs))
    assert worst < 1e-9, worst


@pytest.mark.parametrize("dim", [2, 3])
def test_derivative_weights_follow_reference_monomial_functional(dim):
    dx = 1.0 / 64
    fw, dw = kfvm_amd.build_weights(dim, 2, 5.0, dx, dx, dx)
    dset = monos(dim, 3 if dim == 3 else 2)[1:]
    quirk_differs = False
    for s in (0, 1, 2):
        cells, off = sub_cells(dim, s)
        deg = 3 if s == 0 else 1
        for i, Nd in enumerate(dset):
            w = dw[s, i, : len(cells)] / dx ** sum(Nd)
            for m in monos(dim, deg):
                lhs = sum(w[n] * mono_avg(m, off[c], dim) for n, c in enumerate(cells))
                rhs = quirk_der0(m[0], Nd[0]) * quirk_der0(m[1], Nd[1])
                tru = true_der0(m[0], Nd[0]) * true_der0(m[1], Nd[1])
                if dim == 3:
                    rhs *= quirk_der0(m[2], Nd[2])
                    tru *= true_der0(m[2], Nd[2])
                assert abs(lhs - rhs) < 1e-7 * max(1.0, abs(rhs)), (s, Nd, m, lhs, rhs)
                quirk_differs |= rhs != tru
    # the reference's argument order changes second/third-derivative targets
    assert quirk_differs


def test_bitwise_face_symmetry():
    fw, _ = kfvm_amd.build_weights(3, 2, 5.0, 1 / 32, 1 / 32, 1 / 32)
    off, _ = stencil3d()
    pts = face_points(3, 2)
    idx = {o: n for n, o in enumerate(off)}
    # east face = x-reflection of the west face, point by point
    for qe, pe in enumerate(pts[1]):
        qw = pts[0].index((-pe[0], pe[1], pe[2]))
        perm = [idx[(-o[0], o[1], o[2])] for o in off]
        assert np.array_equal(fw[0, 1, qe], fw[0, 0, qw][perm])


def test_kernel_rows_in_range_of_polynomial_block():
    """k(q) - K w must be P lambda for some lambda (first block row)."""
    from scipy.special import erf
    dim, nq, lfac = 3, 2, 5.0
    eps = 1.0 / (lfac * math.sqrt(2.0))
    fw, _ = kfvm_amd.build_weights(dim, nq, lfac, 1 / 32, 1 / 32, 1 / 32)
    off, _ = stencil3d()
    X = np.array(off, dtype=float)
    kav = lambda d: math.sqrt(math.pi) / (2 * eps) * (erf(eps * (d + 0.5)) - erf(eps * (d - 0.5)))
    K = np.array([[kav(a[0] - b[0]) * kav(a[1] - b[1]) * kav(a[2] - b[2]) for b in X] for a in X])
    P = np.array([[mono_avg(m, c, 3) for m in monos(3, 3)] for c in X])
    q = face_points(3, 2)[0][0]
    kq = np.array([math.exp(-eps * eps * sum((q[d] - c[d]) ** 2 for d in range(3))) for c in X])
    r = kq - K @ fw[0, 0, 0]
    lam, *_ = np.linalg.lstsq(P, r, rcond=None)
    assert np.abs(P @ lam - r).max() < 1e-8 * np.abs(kq).max()


def _mp_weights(dim, cells, monodeg, q, eps, digits=40):
    import mpmath as mp
    mp.mp.dps = digits
    e = mp.mpf(eps)

    def kav(d):
        return mp.sqrt(mp.pi) / (2 * e) * (mp.erf(e * (d + mp.mpf(1) / 2)) - mp.erf(e * (d - mp.mpf(1) / 2)))

    def av(p, x):
        x = mp.mpf(x)
        return ((x + mp.mpf(1) / 2) ** (p + 1) - (x - mp.mpf(1) / 2) ** (p + 1)) / (p + 1)

    ms = monos(dim, monodeg)
    n = len(cells)
    N = n + len(ms)
    A = mp.zeros(N, N)
    b = mp.zeros(N, 1)
    for i, ci in enumerate(cells):
        for j, cj in enumerate(cells):
            v = kav(ci[0] - cj[0]) * kav(ci[1] - cj[1])
            A[i, j] = v * kav(ci[2] - cj[2]) if dim == 3 else v
        for k, m in enumerate(ms):
            v = av(m[0], ci[0]) * av(m[1], ci[1])
            v = v * av(m[2], ci[2]) if dim == 3 else v
            A[i, n + k] = v
            A[n + k, i] = v
        d2 = sum((mp.mpf(q[a]) - ci[a]) ** 2 for a in range(dim))
        b[i] = mp.exp(-e * e * d2)
    for k, m in enumerate(ms):
        v = mp.mpf(q[0]) ** m[0] * mp.mpf(q[1]) ** m[1]
        b[n + k] = v * mp.mpf(q[2]) ** m[2] if dim == 3 else v
    x = mp.lu_solve(A, b)
    return np.array([float(x[i]) for i in range(n)])


@pytest.mark.parametrize("dim,s", [(2, 0), (3, 1), (3, 2)])
def test_against_high_precision_solve(dim, s):
    lfac = 5.0
    eps = 1.0 / (lfac * math.sqrt(2.0))
    fw, _ = kfvm_amd.build_weights(dim, 2, lfac, 1 / 32, 1 / 32, 1 / 32)
    cells, off = sub_cells(dim, s)
    xs = [off[c] for c in cells]
    q = face_points(dim, 2)[0][1]
    ref = _mp_weights(dim, xs, 3 if s == 0 else 1, q, eps)
    got = fw[s, 0, 1, : len(cells)]
    # the double-precision matrix entries alone move the solution by about
    # cond(A) * 1e-16; these substencils are well enough conditioned for 1e-9
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
