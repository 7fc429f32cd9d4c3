// Compile-time geometry of the radius-2 KFVM stencils, usable from host and
// device code (all tables are constexpr so fully unrolled device loops fold
// them into immediates).
//
//   counts       follow src/stencil/StencilInfo.H:21-82 (r = 2)
//   enumeration  follows Stencil::Core::findStencil / findSubStencils,
//                src/stencil/Stencil.cpp:47-229: centre substencil first, then
//                the remaining cells; biased substencils list their cells in
//                full-stencil order (3D) or axis-first (2D).
#pragma once

#ifndef KFVM_HD
#if defined(__HIPCC__)
#define KFVM_HD __host__ __device__
#else
#define KFVM_HD
#endif
#endif

namespace kfvm {

template <int D>
struct SInfo;

template <>
struct SInfo<2> {
  static constexpr int C = 13;   // cells in the full stencil
  static constexpr int Cc = 5;   // centre substencil
  static constexpr int Cb = 5;   // each biased substencil
  static constexpr int B = 4;    // biased substencils
  static constexpr int nSub = 6; // full, centre, W, E, S, N
  static constexpr int nDeriv = 2;
  static constexpr int I = 5;    // derivatives in the smoothness indicator
  static constexpr int F = 4;    // faces
  static constexpr int monoDeg[3] = {3, 1, 1};
};

template <>
struct SInfo<3> {
  static constexpr int C = 33;
  static constexpr int Cc = 7;
  static constexpr int Cb = 11;
  static constexpr int B = 6;
  static constexpr int nSub = 8;  // full, centre, W, E, S, N, B, T
  static constexpr int nDeriv = 3;
  static constexpr int I = 19;
  static constexpr int F = 6;
  static constexpr int monoDeg[3] = {3, 1, 1};
};

template <int D>
struct StencilTables {
  int off[SInfo<D>::C][3];             // (x, y, z) offsets of each stencil cell
  int sub[SInfo<D>::B][SInfo<D>::Cb];  // biased substencil -> full-stencil index
};

namespace detail {
constexpr int iabs(int a) { return a < 0 ? -a : a; }
constexpr int ceil_sqrt(int r2) {
  int r = 0;
  while (r * r < r2) r++;
  return r;
}
template <int D>
constexpr int find_idx(const StencilTables<D> &t, int i, int j, int k) {
  for (int n = 0; n < SInfo<D>::C; n++)
    if (t.off[n][0] == i && t.off[n][1] == j && t.off[n][2] == k) return n;
  return -1;
}
}  // namespace detail

template <int D>
constexpr StencilTables<D> make_stencil_tables() {
  StencilTables<D> t{};
  constexpr int rad = 2;
  int nC = 0;
  if (D == 2) {
    t.off[nC][0] = 0; t.off[nC][1] = 0; t.off[nC][2] = 0; nC++;
    for (int nL = 1; nL <= rad; nL++) {
      const int ax[4][2] = {{-nL, 0}, {nL, 0}, {0, -nL}, {0, nL}};
      for (int a = 0; a < 4; a++) {
        t.off[nC][0] = ax[a][0]; t.off[nC][1] = ax[a][1]; t.off[nC][2] = 0; nC++;
      }
      for (int nR = 1; nR <= nL; nR++) {
        const int rsq = 2 * nR * nR;
        if (rsq <= rad * rad && detail::ceil_sqrt(rsq) == nL) {
          const int dg[4][2] = {{-nR, -nR}, {-nR, nR}, {nR, -nR}, {nR, nR}};
          for (int a = 0; a < 4; a++) {
            t.off[nC][0] = dg[a][0]; t.off[nC][1] = dg[a][1]; t.off[nC][2] = 0; nC++;
          }
        }
      }
      for (int i = 2; i <= rad; i++) {
        for (int j = 1; j < i; j++) {
          const int rsq = i * i + j * j;
          if (rsq <= rad * rad && detail::ceil_sqrt(rsq) == nL) {
            const int kn[8][2] = {{-i, j}, {-i, -j}, {i, j}, {i, -j},
                                  {-j, i}, {-j, -i}, {j, i}, {j, -i}};
            for (int a = 0; a < 8; a++) {
              t.off[nC][0] = kn[a][0]; t.off[nC][1] = kn[a][1]; t.off[nC][2] = 0; nC++;
            }
          }
        }
      }
    }
    // biased substencils: on-axis cells first, then off-axis pairs
    int nW = 0, nE = 0, nS = 0, nN = 0;
    for (int h = 0; h <= rad; h++) {
      t.sub[0][nW++] = detail::find_idx<D>(t, -h, 0, 0);
      t.sub[1][nE++] = detail::find_idx<D>(t, h, 0, 0);
      t.sub[2][nS++] = detail::find_idx<D>(t, 0, -h, 0);
      t.sub[3][nN++] = detail::find_idx<D>(t, 0, h, 0);
    }
    for (int h = 0; h <= rad; h++) {
      for (int l = 1; l <= h; l++) {
        if (detail::find_idx<D>(t, -h, -l, 0) >= 0) {
          t.sub[0][nW++] = detail::find_idx<D>(t, -h, -l, 0);
          t.sub[0][nW++] = detail::find_idx<D>(t, -h, l, 0);
          t.sub[1][nE++] = detail::find_idx<D>(t, h, -l, 0);
          t.sub[1][nE++] = detail::find_idx<D>(t, h, l, 0);
          t.sub[2][nS++] = detail::find_idx<D>(t, -l, -h, 0);
          t.sub[2][nS++] = detail::find_idx<D>(t, l, -h, 0);
          t.sub[3][nN++] = detail::find_idx<D>(t, -l, h, 0);
          t.sub[3][nN++] = detail::find_idx<D>(t, l, h, 0);
        }
      }
    }
  } else {
    const int rm = rad - 1;
    for (int i = -rm; i <= rm; i++)
      for (int j = -rm; j <= rm; j++)
        for (int k = -rm; k <= rm; k++)
          if (i * i + j * j + k * k <= rm * rm) {
            t.off[nC][0] = i; t.off[nC][1] = j; t.off[nC][2] = k; nC++;
          }
    for (int i = -rad; i <= rad; i++)
      for (int j = -rad; j <= rad; j++)
        for (int k = -rad; k <= rad; k++) {
          const int rsq = i * i + j * j + k * k;
          if (rsq > rm * rm && rsq <= rad * rad) {
            t.off[nC][0] = i; t.off[nC][1] = j; t.off[nC][2] = k; nC++;
          }
        }
    int cnt[6] = {0, 0, 0, 0, 0, 0};
    for (int n = 0; n < SInfo<D>::C; n++) {
      const int l = t.off[n][0], m = t.off[n][1], q = t.off[n][2];
      using detail::iabs;
      if (-l >= iabs(m) && -l >= iabs(q)) t.sub[0][cnt[0]++] = n;
      if (l >= iabs(m) && l >= iabs(q)) t.sub[1][cnt[1]++] = n;
      if (-m >= iabs(l) && -m >= iabs(q)) t.sub[2][cnt[2]++] = n;
      if (m >= iabs(l) && m >= iabs(q)) t.sub[3][cnt[3]++] = n;
      if (-q >= iabs(l) && -q >= iabs(m)) t.sub[4][cnt[4]++] = n;
      if (q >= iabs(l) && q >= iabs(m)) t.sub[5][cnt[5]++] = n;
    }
  }
  return t;
}

// Number of monomials of total degree <= t in D dimensions.
// Face points per pass of the linear reconstruction kernel: both faces of
// one axis, or one face when that would hold too many accumulators (3D with
// 3x3 points).  The axis-major weight copy is laid out per pass.
KFVM_HD constexpr int linear_pass_points(int dim, int nq) {
  return dim == 3 ? (nq == 3 ? nq * nq : 2 * nq * nq) : 2 * nq;
}

// Interior quadrature points per pass of the GLM-MHD linear kernel (all
// NQ^D points, or one NQ^2 layer of the 3D 3x3x3 set).  The weight copy
// of that kernel holds the face passes, then the interior passes.
KFVM_HD constexpr int mhd_interior_pass_points(int dim, int nq) { return dim == 3 && nq == 2 ? 8 : nq * nq; }

KFVM_HD constexpr int num_monos(int D, int t) {
  return D == 2 ? (t + 1) * (t + 2) / 2 : (t + 1) * (t + 2) * (t + 3) / 6;
}

// Exponents of the monomials of total degree <= tDeg in the reference's
// enumeration order (Monomials.cpp:33-57); entry 0 is the constant.
template <int D, int TDEG>
struct MonoEnum {
  int deg[num_monos(D, TDEG)][3];
};

template <int D, int TDEG>
constexpr MonoEnum<D, TDEG> make_mono_enum() {
  MonoEnum<D, TDEG> m{};
  int n = 0;
  for (int t = 0; t <= TDEG; t++) {
    if (D == 2) {
      for (int i = 0; i <= t; i++) {
        m.deg[n][0] = i; m.deg[n][1] = t - i; m.deg[n][2] = 0; n++;
      }
    } else {
      for (int i = 0; i <= t; i++)
        for (int j = 0; j <= t - i; j++) {
          m.deg[n][0] = i; m.deg[n][1] = j; m.deg[n][2] = t - i - j; n++;
        }
    }
  }
  return m;
}

}  // namespace kfvm
