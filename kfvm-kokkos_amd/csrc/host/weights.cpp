// Reconstruction-weight generation (host set-up, not on the timed path).
//
// For each (sub)stencil the weights w solve the saddle-point system
//     [ K_avg  P_avg ] [ w ]   [ k(q) ]
//     [ P_avg^T  0   ] [ l ] = [ p(q) ]
// with K_avg the squared-exponential kernel averaged over one cell
// (SEKernel.H:19-25), P_avg cell-averaged monomials (Monomials.cpp:14-19) and
// (k, p) the kernel / monomials evaluated (or differentiated) at the query q
// (SEKernel.H:27-33, Monomials.cpp:21-31).  ReconVectors.cpp:26-86 factors the
// matrix with LAPACK dgetrf; here the factorisation and solves run in IEEE
// binary128 on the double-precision matrix entries and the result is rounded
// once to double, so the table does not depend on a BLAS/LAPACK build.
//
// Weights that the reference obtains by symmetry (StencilSymmetry.cpp) are
// produced the same way here: an exact permutation of an already computed
// vector, with the symmetry found by the same ordered search over the 13
// reflections / rotations, so the table has bitwise stencil symmetry.
#include "weights.hpp"

#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "stencil_geom.hpp"

namespace kfvm {
namespace {

using qreal = __float128;

struct Pt {
  double x, y, z;
};

// ----------------------------------------------------------------- 1D pieces
double se_avg(double eps, double x) {
  const double pf = std::sqrt(M_PI) / (2.0 * eps);
  return pf * (std::erf(eps * (x + 0.5)) - std::erf(eps * (x - 0.5)));
}

double se_der(double eps, long N, double x) {
  const double sgn = N % 2 == 0 ? 1.0 : -1.0;
  const double Hn = sgn * std::pow(eps, N) * std::hermite(static_cast<unsigned>(N), eps * x);
  return Hn * std::exp(-eps * eps * x * x);
}

double mono_avg(long p, double x) {
  const double ub = std::pow(x + 0.5, p + 1) / (p + 1);
  const double lb = std::pow(x - 0.5, p + 1) / (p + 1);
  return ub - lb;
}

// Derivative functional applied to x^p.  The recursive call passes its
// arguments in the order (p-1, x, N-1) into the (p, N, x) slots, exactly as
// Monomials.cpp:29 does; for the derivative queries at x = 0 this evaluates
// d^N x^p to p*(N-1)^(p-1).  The smoothness weights depend on it.
double mono_der(long p, long N, double x) {
  if (N == 0) return std::pow(x, p);
  if (N > p) return 0.0;
  return p * mono_der(p - 1, static_cast<long>(x), static_cast<double>(N - 1));
}

// ------------------------------------------------------------ saddle system
struct System {
  int dim, n, nMono, size;
  double eps;
  std::vector<Pt> xs;
  std::vector<std::array<int, 3>> monos;
  std::vector<qreal> lu;  // row-major size x size
  std::vector<int> piv;

  System(int dim_, double eps_, int monoDeg, const std::vector<Pt> &xs_)
      : dim(dim_), n(int(xs_.size())), eps(eps_), xs(xs_) {
    for (int t = 0; t <= monoDeg; t++) {
      if (dim == 2) {
        for (int i = 0; i <= t; i++) monos.push_back({i, t - i, 0});
      } else {
        for (int i = 0; i <= t; i++)
          for (int j = 0; j <= t - i; j++) monos.push_back({i, j, t - i - j});
      }
    }
    nMono = int(monos.size());
    size = n + nMono;
    std::vector<double> A(size_t(size) * size, 0.0);
    auto at = [&](int i, int j) -> double & { return A[size_t(i) * size + j]; };
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double dx = xs[i].x - xs[j].x, dy = xs[i].y - xs[j].y, dz = xs[i].z - xs[j].z;
        double k = se_avg(eps, dx) * se_avg(eps, dy);
        if (dim == 3) k = k * se_avg(eps, dz);
        at(i, j) = k;
      }
    for (int i = 0; i < n; i++)
      for (int m = 0; m < nMono; m++) {
        double v = mono_avg(monos[m][0], xs[i].x) * mono_avg(monos[m][1], xs[i].y);
        if (dim == 3) v = v * mono_avg(monos[m][2], xs[i].z);
        at(i, n + m) = v;
        at(n + m, i) = v;
      }
    factor(A);
  }

  void factor(const std::vector<double> &A) {
    lu.assign(A.begin(), A.end());
    piv.resize(size);
    for (int k = 0; k < size; k++) {
      int p = k;
      qreal best = lu[size_t(k) * size + k];
      if (best < 0) best = -best;
      for (int i = k + 1; i < size; i++) {
        qreal v = lu[size_t(i) * size + k];
        if (v < 0) v = -v;
        if (v > best) {
          best = v;
          p = i;
        }
      }
      piv[k] = p;
      if (best == 0) throw std::runtime_error("singular GP reconstruction system");
      if (p != k)
        for (int j = 0; j < size; j++) std::swap(lu[size_t(k) * size + j], lu[size_t(p) * size + j]);
      const qreal inv = qreal(1) / lu[size_t(k) * size + k];
      for (int i = k + 1; i < size; i++) {
        qreal &lik = lu[size_t(i) * size + k];
        lik = lik * inv;
        if (lik == 0) continue;
        for (int j = k + 1; j < size; j++) lu[size_t(i) * size + j] -= lik * lu[size_t(k) * size + j];
      }
    }
  }

  // Query functional: derivative orders Nd at point q.
  std::vector<double> solve(const int Nd[3], const Pt &q) const {
    std::vector<qreal> b(size);
    for (int i = 0; i < n; i++) {
      double k = se_der(eps, Nd[0], q.x - xs[i].x) * se_der(eps, Nd[1], q.y - xs[i].y);
      if (dim == 3) k = k * se_der(eps, Nd[2], q.z - xs[i].z);
      b[i] = k;
    }
    for (int m = 0; m < nMono; m++) {
      double v = mono_der(monos[m][0], Nd[0], q.x) * mono_der(monos[m][1], Nd[1], q.y);
      if (dim == 3) v = v * mono_der(monos[m][2], Nd[2], q.z);
      b[n + m] = v;
    }
    for (int k = 0; k < size; k++) std::swap(b[k], b[piv[k]]);
    for (int i = 0; i < size; i++) {
      qreal s = b[i];
      for (int j = 0; j < i; j++) s -= lu[size_t(i) * size + j] * b[j];
      b[i] = s;
    }
    for (int i = size - 1; i >= 0; i--) {
      qreal s = b[i];
      for (int j = i + 1; j < size; j++) s -= lu[size_t(i) * size + j] * b[j];
      b[i] = s / lu[size_t(i) * size + i];
    }
    std::vector<double> w(n);
    for (int i = 0; i < n; i++) w[i] = static_cast<double>(b[i]);
    return w;
  }
};

// --------------------------------------------------------------- symmetries
enum Sym { RefX, RefY, RefZ, RefXY, RefXZ, RefYZ, RefXYZ, RotXY, RotXZ, RotYZ, RotXYm, RotXZm, RotYZm, NSym };

bool deq(double a, double b) { return std::fabs(a - b) < 1.e-6; }

// Does symmetry s take reference point b onto destination point a?
bool test_sym(int dim, int s, const Pt &a, const Pt &b) {
  if (dim == 2) {
    switch (s) {
    case RefX: return deq(a.x, -b.x) && deq(a.y, b.y);
    case RefY: return deq(a.x, b.x) && deq(a.y, -b.y);
    case RefXY: return deq(a.x, -b.x) && deq(a.y, -b.y);
    case RotXY: return deq(a.x, -b.y) && deq(a.y, b.x);
    case RotXYm: return deq(a.x, b.y) && deq(a.y, -b.x);
    default: return false;
    }
  }
  switch (s) {
  case RefX: return deq(a.x, -b.x) && deq(a.y, b.y) && deq(a.z, b.z);
  case RefY: return deq(a.x, b.x) && deq(a.y, -b.y) && deq(a.z, b.z);
  case RefZ: return deq(a.x, b.x) && deq(a.y, b.y) && deq(a.z, -b.z);
  case RefXY: return deq(a.x, -b.x) && deq(a.y, -b.y) && deq(a.z, b.z);
  case RefXZ: return deq(a.x, -b.x) && deq(a.y, b.y) && deq(a.z, -b.z);
  case RefYZ: return deq(a.x, b.x) && deq(a.y, -b.y) && deq(a.z, -b.z);
  case RefXYZ: return deq(a.x, -b.x) && deq(a.y, -b.y) && deq(a.z, -b.z);
  case RotXY: return deq(a.x, -b.y) && deq(a.y, b.x) && deq(a.z, b.z);
  case RotXZ: return deq(a.x, -b.z) && deq(a.y, b.y) && deq(a.z, b.x);
  case RotYZ: return deq(a.x, b.x) && deq(a.y, -b.z) && deq(a.z, b.y);
  case RotXYm: return deq(a.x, b.y) && deq(a.y, -b.x) && deq(a.z, b.z);
  case RotXZm: return deq(a.x, b.z) && deq(a.y, b.y) && deq(a.z, -b.x);
  case RotYZm: return deq(a.x, b.x) && deq(a.y, b.z) && deq(a.z, -b.y);
  default: return false;
  }
}

// For every destination point: the reference point it is the image of, and the
// stencil-cell permutation that carries the reference weights over.
struct SymMap {
  std::vector<int> refPt;
  std::vector<std::vector<int>> idx;
};

SymMap make_symmap(int dim, bool isSelf, const std::vector<Pt> &offD, const std::vector<Pt> &offR,
                   const std::vector<Pt> &ptD, const std::vector<Pt> &ptR) {
  const int N = int(offD.size()), nPts = int(ptD.size());
  // symmetries under which every destination cell has a reference cell
  std::vector<int> stenSym;
  for (int s = 0; s < NSym; s++) {
    bool all = true;
    for (int nD = 0; nD < N && all; nD++) {
      bool found = false;
      for (int nR = 0; nR < N && !found; nR++) found = test_sym(dim, s, offD[nD], offR[nR]);
      all = found;
    }
    if (all) stenSym.push_back(s);
  }
  SymMap m;
  m.refPt.assign(nPts, -1);
  m.idx.assign(nPts, std::vector<int>(N, -1));
  std::vector<int> symOf(nPts, -1);
  for (int nD = 0; nD < nPts; nD++) {
    const int refLim = isSelf ? nD : int(ptR.size());
    for (int nR = 0; nR < refLim && m.refPt[nD] < 0; nR++)
      for (int s : stenSym)
        if (test_sym(dim, s, ptD[nD], ptR[nR])) {
          m.refPt[nD] = nR;
          symOf[nD] = s;
          break;
        }
  }
  for (int nP = 0; nP < nPts; nP++) {
    if (m.refPt[nP] < 0) continue;
    for (int nD = 0; nD < N; nD++)
      for (int nR = 0; nR < N; nR++)
        if (test_sym(dim, symOf[nP], offD[nD], offR[nR])) {
          m.idx[nP][nD] = nR;
          break;
        }
  }
  return m;
}

// Accessor over a block of weight rows: row(p)[c] for p in points, c < C.
struct Rows {
  double *base;
  int stride;  // distance between consecutive points
  double *row(int p) const { return base + size_t(p) * stride; }
};

}  // namespace

WeightTable build_weights(int dim, int nq, double lfac, double dx, double dy, double dz) {
  if (dim != 2 && dim != 3) throw std::runtime_error("dimension must be 2 or 3");
  const QuadRule qr = quad_rule(nq);
  if (qr.n == 0) throw std::runtime_error("unsupported number of quadrature points");

  WeightTable T;
  T.dim = dim;
  T.nq = nq;
  int B, Cc, Cb, nDeriv;
  std::vector<Pt> off;
  std::vector<std::vector<int>> sub;
  if (dim == 2) {
    constexpr auto st = make_stencil_tables<2>();
    T.nSub = SInfo<2>::nSub; T.F = 4; T.C = SInfo<2>::C; T.I = SInfo<2>::I;
    B = SInfo<2>::B; Cc = SInfo<2>::Cc; Cb = SInfo<2>::Cb; nDeriv = SInfo<2>::nDeriv;
    for (int c = 0; c < T.C; c++) off.push_back({double(st.off[c][0]), double(st.off[c][1]), 0.0});
    for (int b = 0; b < B; b++) sub.emplace_back(st.sub[b], st.sub[b] + Cb);
    T.Q = nq;
    T.nqCell = nq * nq;
  } else {
    constexpr auto st = make_stencil_tables<3>();
    T.nSub = SInfo<3>::nSub; T.F = 6; T.C = SInfo<3>::C; T.I = SInfo<3>::I;
    B = SInfo<3>::B; Cc = SInfo<3>::Cc; Cb = SInfo<3>::Cb; nDeriv = SInfo<3>::nDeriv;
    for (int c = 0; c < T.C; c++)
      off.push_back({double(st.off[c][0]), double(st.off[c][1]), double(st.off[c][2])});
    for (int b = 0; b < B; b++) sub.emplace_back(st.sub[b], st.sub[b] + Cb);
    T.Q = nq * nq;
    T.nqCell = nq * nq * nq;
  }
  T.face.assign(size_t(T.nSub) * T.F * T.Q * T.C, 0.0);
  T.deriv.assign(size_t(T.nSub) * T.I * T.C, 0.0);
  T.cell.assign(size_t(T.nSub) * T.nqCell * T.C, 0.0);

  const double eps = 1.0 / (lfac * std::sqrt(2.0));
  const int monoDeg[3] = {3, 1, 1};

  // derivative set: monomial exponents up to nDeriv, entry 0 (point value) skipped
  std::vector<std::array<int, 3>> dset;
  for (int t = 0; t <= nDeriv; t++) {
    if (dim == 2) {
      for (int i = 0; i <= t; i++) dset.push_back({i, t - i, 0});
    } else {
      for (int i = 0; i <= t; i++)
        for (int j = 0; j <= t - i; j++) dset.push_back({i, j, t - i - j});
    }
  }

  // face quadrature point sets, index q = nQ*NQ + nR in 3D (Stencil.cpp:303-319)
  std::vector<double> fq1(T.Q), fq2(T.Q, 0.0);
  if (dim == 2) {
    for (int q = 0; q < nq; q++) fq1[q] = qr.ab[q];
  } else {
    for (int a = 0; a < nq; a++)
      for (int b = 0; b < nq; b++) {
        fq1[a * nq + b] = qr.ab[a];
        fq2[a * nq + b] = qr.ab[b];
      }
  }
  std::vector<Pt> fpts[6];
  for (int q = 0; q < T.Q; q++) {
    fpts[West].push_back({-0.5, fq1[q], fq2[q]});
    fpts[East].push_back({0.5, fq1[q], fq2[q]});
    fpts[South].push_back({fq1[q], -0.5, fq2[q]});
    fpts[North].push_back({fq1[q], 0.5, fq2[q]});
    if (dim == 3) {
      fpts[Bottom].push_back({fq1[q], fq2[q], -0.5});
      fpts[Top].push_back({fq1[q], fq2[q], 0.5});
    }
  }
  if (dim == 2)
    for (int f = 0; f < 4; f++)
      for (auto &p : fpts[f]) p.z = 0.0;
  std::vector<Pt> cpts;
  for (int a = 0; a < nq; a++)
    for (int b = 0; b < nq; b++) {
      if (dim == 2) cpts.push_back({qr.ab[a], qr.ab[b], 0.0});
      else
        for (int c = 0; c < nq; c++) cpts.push_back({qr.ab[a], qr.ab[b], qr.ab[c]});
    }
  const std::vector<Pt> dpt = {{0.0, 0.0, 0.0}};
  const int pointSpec[3] = {0, 0, 0};

  auto sten = [&](int s) {
    std::vector<Pt> xs;
    const int size = s == 0 ? T.C : (s == 1 ? Cc : Cb);
    for (int n = 0; n < size; n++) xs.push_back(off[s < 2 ? n : sub[s - 2][n]]);
    return xs;
  };
  auto faceRows = [&](int s, int f) { return Rows{&T.f(s, f, 0, 0), T.C}; };
  auto cellRows = [&](int s) { return Rows{&T.cw(s, 0, 0), T.C}; };

  // Solve for all points of a set, reusing earlier points related by symmetry.
  auto recvecs = [&](const System &sys, const std::vector<Pt> &xs, const std::vector<Pt> &pts,
                     Rows dst) {
    SymMap m = make_symmap(dim, true, xs, xs, pts, pts);
    for (size_t p = 0; p < pts.size(); p++) {
      double *r = dst.row(int(p));
      if (m.refPt[p] < 0) {
        std::vector<double> w = sys.solve(pointSpec, pts[p]);
        for (size_t j = 0; j < xs.size(); j++) r[j] = w[j];
        T.nSolved++;
      } else {
        const double *src = dst.row(m.refPt[p]);
        for (size_t j = 0; j < xs.size(); j++) r[j] = src[m.idx[p][j]];
      }
    }
  };
  auto mapweights = [&](const std::vector<Pt> &xsD, const std::vector<Pt> &xsR,
                        const std::vector<Pt> &ptD, const std::vector<Pt> &ptR, Rows dst,
                        Rows src) {
    SymMap m = make_symmap(dim, false, xsD, xsR, ptD, ptR);
    for (size_t p = 0; p < ptD.size(); p++) {
      if (m.refPt[p] < 0) throw std::runtime_error("missing stencil symmetry");
      const double *s = src.row(m.refPt[p]);
      double *d = dst.row(int(p));
      for (size_t j = 0; j < xsD.size(); j++) d[j] = s[m.idx[p][j]];
    }
  };
  auto derivs = [&](const System &sys, int s) {
    const size_t n = sys.xs.size();
    for (int nI = 0; nI < T.I; nI++) {
      const int Nd[3] = {dset[nI + 1][0], dset[nI + 1][1], dset[nI + 1][2]};
      std::vector<double> w = sys.solve(Nd, dpt[0]);
      T.nSolved++;
      double scl = std::pow(dx, Nd[0]);
      scl *= std::pow(dy, Nd[1]);
      if (dim == 3) scl *= std::pow(dz, Nd[2]);
      for (size_t c = 0; c < n; c++) T.d(s, nI, int(c)) = w[c] * scl;
    }
  };

  // full and centre substencils (Stencil.cpp:350-420)
  for (int s = 0; s < 2; s++) {
    const std::vector<Pt> xs = sten(s);
    System sys(dim, eps, monoDeg[s], xs);
    recvecs(sys, xs, fpts[West], faceRows(s, West));
    mapweights(xs, xs, fpts[East], fpts[West], faceRows(s, East), faceRows(s, West));
    mapweights(xs, xs, fpts[South], fpts[West], faceRows(s, South), faceRows(s, West));
    mapweights(xs, xs, fpts[North], fpts[South], faceRows(s, North), faceRows(s, South));
    if (dim == 3) {
      mapweights(xs, xs, fpts[Bottom], fpts[West], faceRows(s, Bottom), faceRows(s, West));
      mapweights(xs, xs, fpts[Top], fpts[Bottom], faceRows(s, Top), faceRows(s, Bottom));
    }
    recvecs(sys, xs, cpts, cellRows(s));
    derivs(sys, s);
  }

  // west-biased substencil solved directly (Stencil.cpp:422-491)
  const std::vector<Pt> xw = sten(2);
  {
    System sys(dim, eps, monoDeg[2], xw);
    recvecs(sys, xw, fpts[West], faceRows(2, West));
    recvecs(sys, xw, fpts[East], faceRows(2, East));
    recvecs(sys, xw, fpts[South], faceRows(2, South));
    mapweights(xw, xw, fpts[North], fpts[South], faceRows(2, North), faceRows(2, South));
    if (dim == 3) {
      mapweights(xw, xw, fpts[Bottom], fpts[South], faceRows(2, Bottom), faceRows(2, South));
      mapweights(xw, xw, fpts[Top], fpts[Bottom], faceRows(2, Top), faceRows(2, Bottom));
    }
    recvecs(sys, xw, cpts, cellRows(2));
    derivs(sys, 2);
  }

  // remaining biased substencils as images of the west one (Stencil.cpp:493-824):
  // for substencil s, face f takes the west-stencil weights of face refFace[s][f]
  const int refFace[8][6] = {
      {0, 0, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0},
      {East, West, South, North, Bottom, Top},   // east stencil
      {North, South, West, East, Bottom, Top},   // south stencil
      {South, North, East, West, Bottom, Top},   // north stencil
      {Top, Bottom, South, North, West, East},   // bottom stencil
      {Bottom, Top, South, North, East, West}};  // top stencil
  for (int s = 3; s < T.nSub; s++) {
    const std::vector<Pt> xs = sten(s);
    for (int f = 0; f < T.F; f++)
      mapweights(xs, xw, fpts[f], fpts[refFace[s][f]], faceRows(s, f), faceRows(2, refFace[s][f]));
    mapweights(xs, xw, cpts, cpts, cellRows(s), cellRows(2));
    SymMap m = make_symmap(dim, false, xs, xw, dpt, dpt);
    if (m.refPt[0] < 0) throw std::runtime_error("missing stencil symmetry");
    for (int nI = 0; nI < T.I; nI++)
      for (int j = 0; j < Cb; j++) T.d(s, nI, j) = T.d(2, nI, m.idx[0][j]);
  }
  return T;
}

}  // namespace kfvm
