// HIP (gfx950) kernels of one KFVM-WENO RK stage.
//
// Data layout in HBM (FP64, structure of arrays, x fastest):
//   U[v][cell]                       cell averages, no ghost cells: boundary
//                                    conditions are applied as index maps
//                                    (mirror / periodic / signed mirror) when
//                                    stencils are gathered
//   FS, value (f*Q + q)*NV + v       reconstructed state of each cell on its own
//                                    face f (W,E,S,N,B,T) at face point q, in
//                                    blocks of FSB cells (fs_base)
//   WI, value f*Q + q                KXRCF indicator p/rho^gamma at those points,
//                                    same blocked layout
//   theta[cell], wflag[cell]         positivity theta and WENO flag
//
// One stage = linear recon (+indicator)  -> KXRCF select (compacted list)
//           -> KFVM-WENO on the list (or all cells) -> positivity limiter
//           -> Riemann solves + flux divergence + RK register update (+ error
//              norm on the FSAL stage), with the max wave speed reduced on
//              the device.
//
// Reference correspondence (per kernel below):
//   k_linear     Stencil_K.H:32-90     k_select   Numeric_K.H:292-398
//   k_weno       Stencil_K.H:164-283, 416-454, 506-589
//   k_pp         Common_K.H:299-643    k_flux_rk  Common_K.H:180-297,
//                                      Numeric_K.H:109-290
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../host/stencil_geom.hpp"
#include "device_api.hpp"
#include "physics.hpp"

namespace kfvm {
namespace dev {

constexpr int BLK = 256;

// ------------------------------------------------------------ index helpers
// Source index along one axis for a (possibly ghost) cell index i, |i| <= n:
// periodic wrap or mirror; side = 0 / 1 when mirrored at the low / high end.
// Written with selects so a gather compiles without branches.
__device__ inline int map_axis(int i, int n, int bcLo, int bcHi, int &side) {
  const bool lo = i < 0, hi = i >= n;
  const int per = lo ? i + n : (hi ? i - n : i);
  const int mir = lo ? -i - 1 : (hi ? 2 * n - 1 - i : i);
  const bool perLo = bcLo == 0, perHi = bcHi == 0;
  const bool usePer = (lo && perLo) || (hi && perHi);
  side = (lo && !perLo) ? 0 : ((hi && !perHi) ? 1 : -1);
  return usePer ? per : mir;
}

// Source of a (possibly ghost) cell: linear index of the interior cell it
// copies and the mask of variables the wall mirrors negate.  The per-face
// masks are picked with selects: a lane-dependent index into the Grid
// argument would turn every gather into extra memory loads.
struct CellRef {
  size_t idx;
  unsigned neg;
};

__device__ inline unsigned side_mask(int side, unsigned lo, unsigned hi) {
  return side == 0 ? lo : (side == 1 ? hi : 0u);
}

template <int D, int NV>
__device__ inline CellRef map_cell(const Grid &g, int i, int j, int k) {
  int sx, sy, sz = -1;
  const int ii = map_axis(i, g.nx, g.bc[0], g.bc[1], sx);
  const int jj = map_axis(j, g.ny, g.bc[2], g.bc[3], sy);
  // 2D: k is the slice of a batch of independent problems (0 for one problem)
  int kk = k;
  if (D == 3) kk = map_axis(k, g.nz, g.bc[4], g.bc[5], sz);
  CellRef r;
  r.idx = (size_t(kk) * g.ny + jj) * g.nx + ii;
  r.neg = 0u;
  if (NV >= 5) {
    r.neg = side_mask(sx, g.negmask[0], g.negmask[1]) ^ side_mask(sy, g.negmask[2], g.negmask[3]);
    if (D == 3) r.neg ^= side_mask(sz, g.negmask[4], g.negmask[5]);
  }
  return r;
}

__device__ inline double ld(const double *__restrict__ U, size_t nc, const CellRef &r, int v) {
  const double x = U[size_t(v) * nc + r.idx];
  return ((r.neg >> v) & 1u) ? -x : x;
}

// Cell average of variable v at (possibly ghost) cell (i,j,k).
template <int D, int NV>
__device__ inline double load_u(const double *__restrict__ U, const Grid &g, int i, int j, int k,
                                int v) {
  return ld(U, g.ncell, map_cell<D, NV>(g, i, j, k), v);
}

template <int D>
__device__ inline void cell_ijk(int c, const Grid &g, int &i, int &j, int &k) {
  i = c % g.nx;
  const int r = c / g.nx;
  j = r % g.ny;
  k = r / g.ny;  // 2D: the slice of a batch (nz = number of problems)
}

template <int D, int NQ>
struct Shape {
  static constexpr int F = 2 * D;
  static constexpr int Q = D == 3 ? NQ * NQ : NQ;
  static constexpr int C = SInfo<D>::C;
};

// ------------------------------------------------------ positivity limiter
// Bounds of the Zhang-Shu-type limiter from the 3^D neighbourhood, widened by
// the flattener (Common_K.H:329-387).
struct PPBounds {
  double rDens, maxDens, minDens, minEint;
};

template <int D>
__device__ inline PPBounds pp_bounds(const double *__restrict__ U, const Grid &g, int i, int j,
                                     int k, const double *ref) {
  PPBounds b;
  b.rDens = ref[0];
  double maxDens = -1.7976931348623157e308, minDens = 1.7976931348623157e308;
  double minEint = 1.7976931348623157e308, minSdSp = 1.7976931348623157e308;
  constexpr int zl = D == 3 ? -1 : 0, zh = D == 3 ? 1 : 0;
#pragma unroll 1
  for (int xo = -1; xo <= 1; xo++)
#pragma unroll 1
    for (int yo = -1; yo <= 1; yo++)
#pragma unroll
      for (int zo = zl; zo <= zh; zo++) {
        double s[5];
#pragma unroll
        for (int v = 0; v < 5; v++) s[v] = load_u<D, 5>(U, g, i + xo, j + yo, k + zo, v);
        maxDens = fmax(maxDens, s[0]);
        minDens = fmin(minDens, s[0]);
        const double ei = Hydro::eint(s);
        const double sd = Hydro::sound(g.gamma, s);
        minEint = fmin(minEint, ei);
        minSdSp = fmin(minSdSp, sd);
      }
  const double dxU = load_u<D, 5>(U, g, i + 1, j, k, 1) / load_u<D, 5>(U, g, i + 1, j, k, 0) -
                     load_u<D, 5>(U, g, i - 1, j, k, 1) / load_u<D, 5>(U, g, i - 1, j, k, 0);
  const double dyV = load_u<D, 5>(U, g, i, j + 1, k, 2) / load_u<D, 5>(U, g, i, j + 1, k, 0) -
                     load_u<D, 5>(U, g, i, j - 1, k, 2) / load_u<D, 5>(U, g, i, j - 1, k, 0);
  double dzW;
  if (D == 3)
    dzW = load_u<D, 5>(U, g, i, j, k + 1, 3) / load_u<D, 5>(U, g, i, j, k + 1, 0) -
          load_u<D, 5>(U, g, i, j, k - 1, 3) / load_u<D, 5>(U, g, i, j, k - 1, 0);
  else
    dzW = ref[3] / ref[0] - ref[3] / ref[0];
  const double divV = (dxU + dyV + dzW) / 2.0;
  double kappa = 0.2;
  const double eta = fmin(1.0, fmax(0.0, -(kappa * minSdSp + divV) / (kappa * minSdSp)));
  kappa *= (1.0 - eta);
  b.maxDens = maxDens * (1.0 + kappa);
  b.minDens = minDens * (1.0 - kappa);
  b.minEint = minEint * (1.0 - kappa);
  return b;
}

__device__ inline double dens_theta(double th, double rho, const PPBounds &b) {
  th = rho > b.minDens ? th : fmin(th, (b.rDens - b.minDens) / (b.rDens - rho));
  th = rho < b.maxDens ? th : fmin(th, (b.maxDens - b.rDens) / (rho - b.rDens));
  return th;
}

__device__ inline double eint_pos(double emin, const double *ref, const double *s) {
  double a = 0.0, b = 1.0, c;
  double st[5];
  for (int it = 0; it < 10; it++) {
    c = 0.5 * (a + b);
#pragma unroll
    for (int v = 0; v < 5; v++) st[v] = ref[v] + c * (s[v] - ref[v]);
    const double dec = Hydro::eint(st) - emin;
    if (dec > 0.0) a = c;
    else b = c;
  }
  return a;
}

// The states of one cell are reached as P[(fq * 5 + v) * ps]: its face-state
// block (P = FS + fs_base(c, np), ps = FSB) or its compact WENO row (ps = 1).
template <int D, int NQ>
__device__ inline void pp_scale(double *__restrict__ P, size_t ps, const double *ref, double th, bool replace) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
#pragma unroll 1
  for (int fq = 0; fq < FQ; fq++)
#pragma unroll
    for (int v = 0; v < 5; v++) {
      double *p = P + (size_t(fq) * 5 + v) * ps;
      *p = replace ? ref[v] : ref[v] + th * (*p - ref[v]);
    }
}

// The limiter on the face states of one cell, in the reference's order:
// density theta, mix, internal-energy theta (10-step bisection), mix.
template <int D, int NQ>
__device__ inline double pp_apply(double *__restrict__ P, size_t ps, const double *ref, const PPBounds &b) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  double thD = 1.0;
#pragma unroll 1
  for (int fq = 0; fq < FQ; fq++) thD = dens_theta(thD, P[(size_t(fq) * 5) * ps], b);
  if (thD > 0.0 && thD < 1.0) pp_scale<D, NQ>(P, ps, ref, thD, false);
  else if (thD < 1.0) pp_scale<D, NQ>(P, ps, ref, thD, true);
  double thP = 1.0;
#pragma unroll 1
  for (int fq = 0; fq < FQ; fq++) {
    double s[5];
#pragma unroll
    for (int v = 0; v < 5; v++) s[v] = P[(size_t(fq) * 5 + v) * ps];
    const double ei = Hydro::eint(s);
    thP = ei > b.minEint ? thP : fmin(thP, eint_pos(b.minEint, ref, s));
  }
  if (thP > 0.0 && thP < 1.0) pp_scale<D, NQ>(P, ps, ref, thP, false);
  else if (thP < 1.0) pp_scale<D, NQ>(P, ps, ref, thP, true);
  return fmin(thD, thP);
}

// ----------------------------------------- linear reconstruction + limiter
// Face states from the full-stencil weights in conserved variables
// (Stencil_K.H:32-90), the KXRCF indicator p/rho^gamma of each (unlimited)
// state, and the positivity limiter: a fast path keeps the states when no
// bound is violated, otherwise the exact reference sequence runs on FS.
template <int D, int NQ, int NV>
__global__ __launch_bounds__(BLK) void k_linear(const double *__restrict__ U,
                                                double *__restrict__ FS,
                                                double *__restrict__ WI,
                                                double *__restrict__ thetaNew, Grid g,
                                                const double *__restrict__ W) {
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = Shape<D, NQ>::C, F = Shape<D, NQ>::F, Q = Shape<D, NQ>::Q, FQ = F * Q;
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int i, j, k;
  cell_ijk<D>(c, g, i, j, k);
  const size_t nc = g.ncell;
  double ref[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) ref[v] = U[size_t(v) * nc + c];
  PPBounds pb;
  if (NV == 5) pb = pp_bounds<D>(U, g, i, j, k, ref);
  double rhoF[FQ], msq[FQ];
  double thD = 1.0;
  bool okE = true;
#pragma unroll 1
  for (int v = 0; v < NV; v++) {
    double vs[C];
#pragma unroll
    for (int n = 0; n < C; n++)
      vs[n] = load_u<D, NV>(U, g, i + ST.off[n][0], j + ST.off[n][1], k + ST.off[n][2], v);
#pragma unroll
    for (int fq = 0; fq < FQ; fq++) {
      double acc = 0.0;
#pragma unroll
      for (int n = 0; n < C; n++) acc = fma(W[fq * C + n], vs[n], acc);
      FS[fs_base(c, FQ * NV) + (size_t(fq) * NV + v) * FSB] = acc;
      if (NV == 5) {
        if (v == 0) {
          rhoF[fq] = acc;
          thD = dens_theta(thD, acc, pb);
        } else if (v == 1) {
          msq[fq] = acc * acc;
        } else if (v < 4) {
          msq[fq] = msq[fq] + acc * acc;
        } else {
          const double ei = acc - 0.5 * msq[fq] / rhoF[fq];
          okE = okE && ei > pb.minEint;
          WI[fs_base(c, FQ) + size_t(fq) * FSB] = (g.gamma - 1.0) * ei / pow(rhoF[fq], g.gamma);
        }
      }
    }
  }
  if (NV == 5) {
    double th = 1.0;
    if (!(thD == 1.0 && okE)) th = pp_apply<D, NQ>(FS + fs_base(c, FQ * NV), FSB, ref, pb);
    thetaNew[c] = th;
  }
}

// Wave-aggregated append of the flagged cells to the compacted list (one
// atomic per wave).  With pos, every live cell also records its list position
// when that is below cap (its WENO states then stay in the compact rows of
// WS), -1 otherwise.
__device__ inline void list_append(bool live, bool flagged, int c, int *__restrict__ list,
                                   int *__restrict__ counter, int *__restrict__ pos, int cap) {
  const unsigned long long mask = __ballot(flagged);
  int idx = -1;
  if (mask != 0ull) {
    const int lane = threadIdx.x & 63;
    const int cnt = __popcll(mask);
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, cnt);
    base = __shfl(base, leader);
    if (flagged) {
      idx = base + __popcll(mask & ((1ull << lane) - 1ull));
      list[idx] = c;
    }
  }
  if (pos && live) pos[c] = idx >= 0 && idx < cap ? idx : -1;
}

// ------------------------------------------------------------ KXRCF selector
template <int D, int NQ, class P = Hydro>
__global__ __launch_bounds__(BLK) void k_select(const double *__restrict__ U,
                                                const double *__restrict__ WI,
                                                const double *__restrict__ theta,
                                                double *__restrict__ wflag, int *__restrict__ list,
                                                int *__restrict__ counter, double wThresh,
                                                Grid g, int *__restrict__ pos, int cap) {
  constexpr int F = Shape<D, NQ>::F, Q = Shape<D, NQ>::Q;
  const int c = blockIdx.x * BLK + threadIdx.x;
  bool flagged = false;
  if (c < g.ncell) {
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    const size_t nc = g.ncell;
    double maxJump = 0.0;
#pragma unroll
    for (int f = 0; f < F; f++) {
      const int ax = f >> 1, hi = f & 1;
      int ni = i, nj = j, nk = k;
      if (ax == 0) ni += hi ? 1 : -1;
      if (ax == 1) nj += hi ? 1 : -1;
      if (ax == 2) nk += hi ? 1 : -1;
      const int n_ax = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
      const int pos = ax == 0 ? ni : (ax == 1 ? nj : nk);
      bool inside = pos >= 0 && pos < n_ax;
      if (!inside && g.bc[f] == 0) {
        inside = true;
        if (ax == 0) ni = (ni + g.nx) % g.nx;
        if (ax == 1) nj = (nj + g.ny) % g.ny;
        if (ax == 2) nk = (nk + g.nz) % g.nz;
      }
      const size_t cn = (size_t(nk) * g.ny + nj) * g.nx + ni;
      const int fo = f ^ 1;  // opposite face of the neighbour
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const double win = WI[fs_base(c, F * Q) + size_t(f * Q + q) * FSB];
        const double wout = inside ? WI[fs_base(cn, F * Q) + size_t(fo * Q + q) * FSB] : win;
        maxJump = fmax(maxJump, fabs(wout - win));
      }
    }
    double uc[P::NV];
#pragma unroll
    for (int v = 0; v < P::NV; v++) uc[v] = U[size_t(v) * nc + c];
    const double ind = maxJump / P::weno(g.gamma, uc);
    flagged = ind >= wThresh || theta[c] < 1.0;
    wflag[c] = flagged ? 1.0 : 0.0;
  }
  list_append(c < g.ncell, flagged, c, list, counter, pos, cap);
}

// ---------------------------------------------------------------- KFVM-WENO
// One cell: linearised-primitive stencil values, smoothness indicators of the
// full / centre / biased substencils, WENO-AO weights, face states, back to
// conserved variables.  The conserved energy is accumulated in FS in the
// reference's operand order, so no per-point arrays stay live in registers.
template <int D, int NQ, int NV>
__device__ inline void weno_cell(const double *__restrict__ U, double *__restrict__ FS, const Grid &g,
                                 const double *__restrict__ FW, const double *__restrict__ DW,
                                 int c) {
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, Cc = SInfo<D>::Cc, Cb = SInfo<D>::Cb;
  constexpr int nSub = SInfo<D>::nSub, I = SInfo<D>::I;
  constexpr int F = Shape<D, NQ>::F, Q = Shape<D, NQ>::Q, FQ = F * Q;
  int i, j, k;
  cell_ijk<D>(c, g, i, j, k);
  const size_t nc = g.ncell;

  double ref[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) ref[v] = U[size_t(v) * nc + c];
  double rho = 1.0, uu[3] = {0.0, 0.0, 0.0}, vsq = 0.0, gamm = 1.0;
  if (NV == 5) {
    LinPrim lp(ref, g.gamma);
    rho = lp.rho;
    uu[0] = lp.u;
    uu[1] = lp.v;
    uu[2] = lp.w;
    vsq = lp.vsq;
    gamm = lp.gamm;
  }
  // linear weights of the WENO-AO combination
  constexpr double gHi = 0.7, gLo = 0.7;
  constexpr double g0 = gHi, g1 = (1.0 - gHi) * gLo, gb = (1.0 - gHi) * (1.0 - gLo) / double(2 * D);

#pragma unroll 1
  for (int v = 0; v < NV; v++) {
    // stencil values of variable v in linearised primitive form
    double vs[C];
    double uv = 0.0, ruv = 0.0;
    if (NV == 5 && v >= 1 && v <= 3) {
      uv = v == 1 ? uu[0] : (v == 2 ? uu[1] : uu[2]);
      ruv = rho * uv;
    }
#pragma unroll
    for (int n = 0; n < C; n++) {
      const int a = i + ST.off[n][0], b = j + ST.off[n][1], d = k + ST.off[n][2];
      if (NV == 5) {
        const double s0 = load_u<D, NV>(U, g, a, b, d, 0);
        if (v == 0) {
          vs[n] = s0;
        } else if (v < 4) {
          const double sv = load_u<D, NV>(U, g, a, b, d, v);
          vs[n] = (sv - uv * s0) / rho;
        } else {
          const double s1 = load_u<D, NV>(U, g, a, b, d, 1);
          const double s2 = load_u<D, NV>(U, g, a, b, d, 2);
          const double s3 = load_u<D, NV>(U, g, a, b, d, 3);
          const double s4 = load_u<D, NV>(U, g, a, b, d, 4);
          vs[n] = gamm * (vsq * s0 - uu[0] * s1 - uu[1] * s2 - uu[2] * s3 + s4);
        }
      } else {
        vs[n] = load_u<D, NV>(U, g, a, b, d, v);
      }
    }
    // smoothness indicators
    double om[nSub];
#pragma unroll
    for (int s = 0; s < nSub; s++) {
      const int n = s == 0 ? C : (s == 1 ? Cc : Cb);
      double beta = 0.0;
#pragma unroll
      for (int ii = 0; ii < I; ii++) {
        const double *dd = DW + (s * I + ii) * C;
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < n; m++) {
          const int cc = s < 2 ? m : ST.sub[s - 2][m];
          t = fma(dd[m], vs[cc], t);
        }
        beta = fma(t, t, beta);
      }
      om[s] = beta;
    }
    // WENO-JS weights, normalised, then the WENO-AO transform
    double tot = 0.0;
#pragma unroll
    for (int s = 0; s < nSub; s++) {
      const double gs = s == 0 ? g0 : (s == 1 ? g1 : gb);
      om[s] = gs / (om[s] * om[s] + 1.e-40);
      tot += om[s];
    }
#pragma unroll
    for (int s = 0; s < nSub; s++) om[s] /= tot;
    om[0] /= g0;
#pragma unroll
    for (int s = 1; s < nSub; s++) om[s] -= om[0] * (s == 1 ? g1 : gb);

    // face values sum_s omega_s (W_s . V_s), converted back to conserved
#pragma unroll 1
    for (int f = 0; f < F; f++) {
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int fq = f * Q + q;
        double acc = 0.0;
#pragma unroll
        for (int s = 0; s < nSub; s++) {
          const int n = s == 0 ? C : (s == 1 ? Cc : Cb);
          const double *w = FW + (s * FQ + fq) * C;
          double t = 0.0;
#pragma unroll
          for (int m = 0; m < n; m++) {
            const int cc = s < 2 ? m : ST.sub[s - 2][m];
            t = fma(w[m], vs[cc], t);
          }
          acc = fma(om[s], t, acc);
        }
        double *out = FS + fs_base(c, FQ * NV) + (size_t(fq) * NV + v) * FSB;
        if (NV == 5) {
          double *dens = FS + fs_base(c, FQ * NV) + (size_t(fq) * NV + 0) * FSB;
          double *etot = FS + fs_base(c, FQ * NV) + (size_t(fq) * NV + 4) * FSB;
          if (v == 0) {
            *out = acc;
            *etot = vsq * acc;
          } else if (v < 4) {
            *out = rho * acc + uv * (*dens);
            *etot = *etot + ruv * acc;
          } else {
            *out = *etot + acc / gamm;
          }
        } else {
          *out = acc;
        }
      }
    }
  }
}

template <int D, int NQ, int NV>
__global__ __launch_bounds__(BLK) void k_weno(const double *__restrict__ U,
                                              double *__restrict__ FS,
                                              double *__restrict__ thetaNew, Grid g,
                                              const double *__restrict__ FW,
                                              const double *__restrict__ DW,
                                              const int *__restrict__ list,
                                              const int *__restrict__ counter) {
  const int n = list ? *counter : g.ncell;
  for (int t = blockIdx.x * BLK + threadIdx.x; t < n; t += gridDim.x * BLK) {
    const int c = list ? list[t] : t;
    weno_cell<D, NQ, NV>(U, FS, g, FW, DW, c);
    if (NV == 5) {
      int i, j, k;
      cell_ijk<D>(c, g, i, j, k);
      double ref[5];
#pragma unroll
      for (int v = 0; v < 5; v++) ref[v] = U[size_t(v) * g.ncell + c];
      const PPBounds pb = pp_bounds<D>(U, g, i, j, k, ref);
      constexpr int NPL = Shape<D, NQ>::F * Shape<D, NQ>::Q * 5;
      thetaNew[c] = pp_apply<D, NQ>(FS + fs_base(c, NPL), FSB, ref, pb);
    }
  }
}

// -------------------------------------------- Riemann + divergence + RK update
template <int NV, int DIR>
__device__ inline double riemann(const Grid &g, double ch, const double *uL, const double *uR,
                                 double *Fl) {
  if (NV == 9) {
    if (g.rs == 4) return Mhd::solve_kepes<DIR>(g.gamma, ch, uL, uR, Fl);
    return Mhd::solve_hll<DIR>(g.gamma, ch, g.rs == 0, uL, uR, Fl);
  } else if (NV == 5) {
    if (g.rs == 2) return Hydro::solve_hllc<DIR>(g.gamma, uL, uR, Fl);
    return Hydro::solve_hll<DIR>(g.gamma, g.rs == 0, uL, uR, Fl);
  } else {
    if (!g.haveVel && DIR == 2) {
      Fl[0] = 0.0;
      return 0.0;
    }
    const double a = g.vel[DIR];
    Fl[0] = a > 0.0 ? a * uL[0] : a * uR[0];
    return fabs(a);
  }
}

// RK register update of one cell from its RHS (Numeric_K.H:198-290): the
// FSAL prime stores K; a stage forms U_next (and Utmp, Uhat); the FSAL stage
// finishes the embedded solution, keeps K and adds the cell's term of the
// error norm, with the positivity minimum (NaN: -1).  write = false (PRIME /
// FSAL of a batch slice that keeps its accepted step) skips the stores.
template <int NV>
__device__ inline void rk_epilogue(const RKArgs &a, const double *__restrict__ U, size_t c, size_t nc,
                                   const double *rhs, double betaDt, double bhatDt, bool write, double &errLoc,
                                   double &posLoc) {
  if (a.mode == RK_PRIME) {
    if (write) {
#pragma unroll
      for (int v = 0; v < NV; v++) a.K[size_t(v) * nc + c] = rhs[v];
    }
  } else if (a.mode == RK_STAGE) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const size_t o = size_t(v) * nc + c;
      const double u = U[o];
      double tmp = 0.0;
      if (a.Utmp) {
        tmp = a.Utmp[o] + a.delta * u;
        a.Utmp[o] = tmp;
      }
      a.Unext[o] = a.g1 * u + a.g2 * tmp + a.g3 * a.Uprev[o] + betaDt * rhs[v];
      a.Uhat[o] = a.Uhat[o] + bhatDt * rhs[v];
    }
  } else if (write) {  // RK_FSAL: finish the embedded solution, keep K, error norm
    bool hasNan = false;
    double st[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const size_t o = size_t(v) * nc + c;
      const double u = U[o];
      const double uh = a.Uhat[o] + bhatDt * rhs[v];
      a.Uhat[o] = uh;
      a.K[o] = rhs[v];
      st[v] = u;
      hasNan = hasNan || isnan(u);
      const double e = (u - uh) / (a.atol + a.rtol * fmax(fabs(u), fabs(uh)));
      errLoc += e * e;  // the reference's pow(e, 2), as a product (rounding-level difference)
    }
    posLoc = hasNan ? -1.0 : (NV == 5 ? Hydro::positivity(st) : (NV == 9 ? fmin(st[0], Mhd::eint(st)) : 1.0));
  }
}

// Block-wide reductions with a fixed combination order (deterministic).
__device__ inline double block_sum(double x, double *sh, int tid) {
  sh[tid] = x;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ inline double block_min(double x, double *sh, int tid) {
  sh[tid] = x;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double o = sh[tid + s];
      sh[tid] = sh[tid] > o ? o : sh[tid];
    }
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ inline double block_max(double x, double *sh, int tid) {
  sh[tid] = x;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double o = sh[tid + s];
      sh[tid] = o > sh[tid] ? o : sh[tid];
    }
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumbering block b of nb so that every XCD works on one contiguous range of
// bricks keeps neighbouring bricks (and their shared halo cells) in the same L2.
__device__ inline int xcd_block(int b, int nb) {
  if (nb % 8) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// Brick of cells handled by one workgroup of the flux kernel.  16 cells in x
// make every row of a face-state plane one full 128-byte line, so no line is
// fetched twice by the two bricks that share it.  2D has two more shapes of
// the same 256 cells, S = 1 (32x8) and S = 2 (64x4), picked per grid size by
// the tile autotuner (kfvm_amd/autotune.py; init key tune:fluxbrick).
template <int D, int S = 0>
struct Brick {
  static constexpr int X = D == 3 || S == 0 ? 16 : (S == 1 ? 32 : 64);
  static constexpr int Y = D == 3 ? 4 : (S == 0 ? 16 : (S == 1 ? 8 : 4));
  static constexpr int Z = D == 3 ? 4 : 1;
  static_assert(X * Y * Z == BLK, "one cell per thread");
  static constexpr int NXF = (X + 1) * Y * Z, NYF = X * (Y + 1) * Z, NZF = D == 3 ? X * Y * (Z + 1) : 0;
  static constexpr int NF = NXF + NYF + NZF;
  // face work items with each direction padded to whole waves, so that no
  // wave runs the Riemann solvers of two directions
  static constexpr int pad64(int n) { return (n + 63) / 64 * 64; }
  static constexpr int PXF = pad64(NXF), PYF = pad64(NYF), PZF = pad64(NZF), PF = PXF + PYF + PZF;
};

// Face state of cell (ci,cj,ck) on its face f at point q; cells outside the
// domain resolve through periodic wrap or the face boundary condition (copy of
// the inner state, with the mirror's sign for reflecting / noslip walls).
// row >= 0 (Euler): the cell's states are row `row` of the compact WENO
// staging ws[t][F*Q*5] instead of its FS slots.
template <int D, int NQ, int NV>
__device__ inline void face_state(const double *__restrict__ FS, const Grid &g, int ci, int cj,
                                  int ck, int f, int q, const double *inner, double *out, int row = -1,
                                  const double *__restrict__ ws = nullptr) {
  constexpr int Q = Shape<D, NQ>::Q;
  const int ax = f >> 1;
  const int n_ax = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
  int pos = ax == 0 ? ci : (ax == 1 ? cj : ck);
  if (pos < 0 || pos >= n_ax) {
    const bool low = pos < 0;
    const int bcLo = ax == 0 ? g.bc[0] : (ax == 1 ? g.bc[2] : g.bc[4]);
    const int bcHi = ax == 0 ? g.bc[1] : (ax == 1 ? g.bc[3] : g.bc[5]);
    if ((low ? bcLo : bcHi) != 0) {
      const unsigned mLo = ax == 0 ? g.negmask[0] : (ax == 1 ? g.negmask[2] : g.negmask[4]);
      const unsigned mHi = ax == 0 ? g.negmask[1] : (ax == 1 ? g.negmask[3] : g.negmask[5]);
      const unsigned m = NV >= 5 ? (low ? mLo : mHi) : 0u;
#pragma unroll
      for (int v = 0; v < NV; v++) out[v] = ((m >> v) & 1u) ? -inner[v] : inner[v];
      return;
    }
    pos = pos < 0 ? pos + n_ax : pos - n_ax;
    if (ax == 0) ci = pos;
    else if (ax == 1) cj = pos;
    else ck = pos;
  }
  const size_t cell = (size_t(ck) * g.ny + cj) * g.nx + ci;
  const size_t nc = g.ncell;
  if (NV == 5 && row >= 0) {
    const double *r = ws + size_t(row) * (Shape<D, NQ>::F * Q * 5) + (f * Q + q) * 5;
#pragma unroll
    for (int v = 0; v < NV; v++) out[v] = r[v];
    return;
  }
#pragma unroll
  for (int v = 0; v < NV; v++)
    out[v] = FS[fs_base(cell, Shape<D, NQ>::F * Q * NV) + (size_t(f * Q + q) * NV + v) * FSB];
}

// GLM-MHD source terms of one cell integrated into rhs (defined with the
// GLM-MHD kernels).
template <int D, int NQ>
__device__ void mhd_source_rhs(const double *__restrict__ U, const double *__restrict__ FS,
                               const double *__restrict__ SRC, const double *__restrict__ theta, const Grid &g,
                               const double *__restrict__ Dm, size_t c, int i, int j, int k, double divB,
                               double *rhs);

template <int D, int NQ, int NV, int S = 0>
__global__ __launch_bounds__(BLK) void k_flux_rk(const double *__restrict__ U,
                                                 const double *__restrict__ FS, Grid g,
                                                 RKArgs a) {
  using Bk = Brick<D, S>;
  constexpr int Q = Shape<D, NQ>::Q;
  __shared__ double fl[NV][Bk::NF];
  // GLM-MHD: normal field averaged over the two states of each face (div B)
  __shared__ double bvf[NV == 9 ? Bk::NF : 1];
  // GLM-MHD: this cell's div B terms of each face point (x, y, z), summed
  // after the face loop in the source kernel's point order
  constexpr int QT = NV == 9 ? Q : 1;
  double divT[QT][3];
#pragma unroll
  for (int qq = 0; qq < QT; qq++) divT[qq][0] = divT[qq][1] = divT[qq][2] = 0.0;
  // compact WENO rows of the brick's cells and their face neighbours (Euler
  // sparse WENO; -1: states in FS)
  constexpr int RX = Bk::X + 2, RY = Bk::Y + 2, RZ = D == 3 ? Bk::Z + 2 : 1;
  __shared__ int rowTab[NV == 5 ? RX * RY * RZ : 1];
  __shared__ double sh[BLK];
  const int tid = threadIdx.x + Bk::X * (threadIdx.y + Bk::Y * threadIdx.z);
  // one-dimensional grid of bricks, x fastest, renumbered per XCD
  const int nbx = (g.nx + Bk::X - 1) / Bk::X, nby = (g.ny + Bk::Y - 1) / Bk::Y;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int bzi = b / (nbx * nby);
  const int i0 = (b % nbx) * Bk::X, j0 = ((b / nbx) % nby) * Bk::Y, k0 = bzi * Bk::Z;
  const int ci = i0 + threadIdx.x, cj = j0 + threadIdx.y, ck = k0 + threadIdx.z;
  const bool live = ci < g.nx && cj < g.ny && (D == 2 || ck < g.nz);
  const size_t nc = g.ncell;
  const size_t c = live ? (size_t(ck) * g.ny + cj) * g.nx + ci : 0;
  double vmax = 0.0, errLoc = 0.0, posLoc = 1.0;
  // 2D batches: the brick's slice (blockIdx.z) and its step size
  const int slice = D == 2 ? bzi : 0;
  const double betaDt = a.sliceDt ? a.beta * a.sliceDt[slice] : a.betaDt;
  const double bhatDt = a.sliceDt ? a.bhat * a.sliceDt[slice] : a.bhatDt;
  // GLM cleaning speed of this evaluation (NV == 9)
  const double ch = NV == 9 ? __longlong_as_double((long long)a.chBits[0]) -
                                  __longlong_as_double((long long)a.chBits[1])
                            : 0.0;
  double rhs[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) rhs[v] = 0.0;
  const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy, rdz = 1.0 / g.dz;

  const bool useRows = NV == 5 && a.pos != nullptr;
  if (useRows) {
    for (int e = tid; e < RX * RY * RZ; e += BLK) {
      const int x = i0 - 1 + e % RX, y = j0 - 1 + (e / RX) % RY, z = D == 3 ? k0 - 1 + e / (RX * RY) : k0;
      // a cell beyond a non-periodic edge is never read (its face takes the
      // boundary state); a periodic one wraps
      const bool wall = (x < 0 && g.bc[0] != 0) || (x >= g.nx && g.bc[1] != 0) || (y < 0 && g.bc[2] != 0) ||
                        (y >= g.ny && g.bc[3] != 0) ||
                        (D == 3 && ((z < 0 && g.bc[4] != 0) || (z >= g.nz && g.bc[5] != 0)));
      int r = -1;
      if (!wall) {
        const int xw = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
        const int yw = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
        const int zw = D == 3 ? (z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z)) : z;
        if (xw >= 0 && xw < g.nx && yw >= 0 && yw < g.ny && zw >= 0 && (D == 2 || zw < g.nz))
          r = a.pos[(size_t(zw) * g.ny + yw) * g.nx + xw];
      }
      if (NV == 5) rowTab[e] = r;
    }
    __syncthreads();
  }
  auto row_of = [&](int x, int y, int z) {
    if (!useRows) return -1;
    const int e = (x - i0 + 1) + RX * ((y - j0 + 1) + RY * (D == 3 ? z - k0 + 1 : 0));
    return NV == 5 ? rowTab[e] : -1;
  };

  // Face items (the same for every face point q): d0 = t | dir << 20 |
  // needed << 30 (-1: padding), and for each side the source of its states:
  // a cell index (>= 0, its FS slots), a compact WENO row r (-(r + 2)), or the
  // mirror of the other side (-1: a wall).  Euler and linear advection
  // resolve them once per brick into LDS (each thread reads back only the
  // items it wrote); GLM-MHD, whose flux planes leave no LDS for them at two
  // workgroups per CU, resolves them at every point.
  constexpr bool FDL = NV != 9;
  __shared__ int fd0[FDL ? Bk::PF : 1], fdL[FDL ? Bk::PF : 1], fdR[FDL ? Bk::PF : 1];
  constexpr int MIRROR = -1;
  auto describe = [&](int w, int &d0, int &sL, int &sR) {
    // padded work item w -> face t of the unpadded numbering (x, then y, then z faces)
    int dir, fi, fj, fk, t;
    bool pad = false;
    if (w < Bk::PXF) {
      pad = w >= Bk::NXF;
      t = w;
      dir = 0;
      fi = t % (Bk::X + 1);
      fj = (t / (Bk::X + 1)) % Bk::Y;
      fk = t / ((Bk::X + 1) * Bk::Y);
    } else if (w < Bk::PXF + Bk::PYF) {
      const int u = w - Bk::PXF;
      pad = u >= Bk::NYF;
      t = Bk::NXF + u;
      dir = 1;
      fi = u % Bk::X;
      fj = (u / Bk::X) % (Bk::Y + 1);
      fk = u / (Bk::X * (Bk::Y + 1));
    } else {
      const int u = w - Bk::PXF - Bk::PYF;
      pad = u >= Bk::NZF;
      t = Bk::NXF + Bk::NYF + u;
      dir = 2;
      fi = u % Bk::X;
      fj = (u / Bk::X) % Bk::Y;
      fk = u / (Bk::X * Bk::Y);
    }
    sL = sR = MIRROR;
    if (pad) {
      d0 = -1;
      return;
    }
    // right cell (the face is its lower face) and left cell
    const int ri = i0 + fi, rj = j0 + fj, rk = k0 + fk;
    const int li = ri - (dir == 0), lj = rj - (dir == 1), lk = rk - (dir == 2);
    const bool rIn = ri < g.nx && rj < g.ny && (D == 2 || rk < g.nz);
    const bool lIn = li >= 0 && lj >= 0 && lk >= 0 && li < g.nx && lj < g.ny && (D == 2 || lk < g.nz);
    // a face is needed when one of its cells is a live cell of this brick
    const bool rBrick = fi < Bk::X && fj < Bk::Y && fk < Bk::Z && rIn;
    const bool lBrick = (dir == 0 ? fi > 0 : (dir == 1 ? fj > 0 : fk > 0)) && lIn;
    const bool needed = rBrick || lBrick;
    // the side's source; a cell off the grid along dir wraps (periodic) or
    // mirrors the other side (wall), as face_state resolves it
    auto source = [&](int x, int y, int z) {
      const int n_ax = dir == 0 ? g.nx : (dir == 1 ? g.ny : g.nz);
      int pos = dir == 0 ? x : (dir == 1 ? y : z);
      const int row = row_of(x, y, z);
      if (pos < 0 || pos >= n_ax) {
        const bool low = pos < 0;
        if ((low ? g.bc[2 * dir] : g.bc[2 * dir + 1]) != 0) return MIRROR;
        pos = low ? pos + n_ax : pos - n_ax;
        if (dir == 0) x = pos;
        else if (dir == 1) y = pos;
        else z = pos;
      }
      if (NV == 5 && row >= 0) return -(row + 2);
      return (z * g.ny + y) * g.nx + x;
    };
    if (needed) {
      sR = source(ri, rj, rk);
      sL = source(li, lj, lk);
    }
    d0 = t | (dir << 20) | (needed ? 1 << 30 : 0);
  };
  if constexpr (FDL) {
    for (int w = tid; w < Bk::PF; w += BLK) describe(w, fd0[w], fdL[w], fdR[w]);
  }
  // states of face f at point q from a side's source
  auto load_side = [&](int src, int f, int q, double *out) {
    if (NV == 5 && src <= -2) {
      const double *r = a.ws + size_t(-src - 2) * (Shape<D, NQ>::F * Q * 5) + (f * Q + q) * 5;
#pragma unroll
      for (int v = 0; v < NV; v++) out[v] = r[v];
    } else {
      const double *p = FS + fs_base(size_t(src), Shape<D, NQ>::F * Q * NV) + size_t(f * Q + q) * NV * FSB;
#pragma unroll
      for (int v = 0; v < NV; v++) out[v] = p[size_t(v) * FSB];
    }
  };
  auto mirror_of = [&](const double *inner, unsigned m, double *out) {
    const unsigned mm = NV >= 5 ? m : 0u;
#pragma unroll
    for (int v = 0; v < NV; v++) out[v] = ((mm >> v) & 1u) ? -inner[v] : inner[v];
  };

#pragma unroll 1
  for (int q = 0; q < Q; q++) {
    // every face of the brick at point q, solved once
#pragma unroll 1
    for (int w = tid; w < Bk::PF; w += BLK) {
      int d0, sL, sR;
      if constexpr (FDL) {
        d0 = fd0[w];
        sL = fdL[w];
        sR = fdR[w];
      } else {
        describe(w, d0, sL, sR);
      }
      if (d0 < 0) continue;
      const int t = d0 & 0xFFFFF, dir = (d0 >> 20) & 3;
      double Fl[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) Fl[v] = 0.0;
      double bv = 0.0;
      if (d0 >> 30) {
        double uL[NV], uR[NV];
        const int lo = 2 * dir + 1, hi = 2 * dir;  // left cell's upper face, right cell's lower face
        if (sR != MIRROR) {
          load_side(sR, hi, q, uR);
          if (sL == MIRROR) mirror_of(uR, g.negmask[2 * dir], uL);
          else load_side(sL, lo, q, uL);
        } else {
          load_side(sL, lo, q, uL);
          mirror_of(uL, g.negmask[2 * dir + 1], uR);
        }
        double sp;
        if (dir == 0) {
          sp = riemann<NV, 0>(g, ch, uL, uR, Fl);
          if constexpr (NV == 9) bv = 0.5 * (uL[5] + uR[5]);
        } else if (dir == 1) {
          sp = riemann<NV, 1>(g, ch, uL, uR, Fl);
          if constexpr (NV == 9) bv = 0.5 * (uL[6] + uR[6]);
        } else {
          sp = riemann<NV, 2>(g, ch, uL, uR, Fl);
          if constexpr (NV == 9) bv = 0.5 * (uL[7] + uR[7]);
        }
        vmax = sp > vmax ? sp : vmax;
      }
#pragma unroll
      for (int v = 0; v < NV; v++) fl[v][t] = Fl[v];
      if constexpr (NV == 9) bvf[t] = bv;
    }
    __syncthreads();
    if (live) {
      const int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
      const int xW = tx + (Bk::X + 1) * (ty + Bk::Y * tz), xE = xW + 1;
      const int yS = Bk::NXF + tx + Bk::X * (ty + (Bk::Y + 1) * tz), yN = yS + Bk::X;
      const int zB = Bk::NXF + Bk::NYF + tx + Bk::X * (ty + Bk::Y * tz), zT = zB + Bk::X * Bk::Y;
      // the reference divides each term by the spacing; multiplying by its
      // reciprocal differs by rounding only and saves 12 NV divisions
      if (D == 2) {
        const double w = g.wt[q];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          rhs[v] += w * (fl[v][xW] - fl[v][xE]) * rdx;
          rhs[v] += w * (fl[v][yS] - fl[v][yN]) * rdy;
        }
      } else {
        const double w = g.wt[q / NQ] * g.wt[q % NQ];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          rhs[v] += w * (fl[v][xW] - fl[v][xE]) * rdx;
          rhs[v] += w * (fl[v][yS] - fl[v][yN]) * rdy;
          rhs[v] += w * (fl[v][zB] - fl[v][zT]) * rdz;
        }
      }
      if constexpr (NV == 9) {
        // div B terms of point q as SourceTerms_K forms them (the face's
        // average is the same from either side: the sum commutes)
        const double twt = D == 2 ? g.wt[q] : g.wt[q % NQ] * g.wt[q / NQ];
        const double tx = twt * (bvf[xE] - bvf[xW]) / g.dx;
        const double ty = twt * (bvf[yN] - bvf[yS]) / g.dy;
        const double tz = D == 3 ? twt * (bvf[zT] - bvf[zB]) / g.dz : 0.0;
#pragma unroll
        for (int qq = 0; qq < QT; qq++) {
          divT[qq][0] = q == qq ? tx : divT[qq][0];
          divT[qq][1] = q == qq ? ty : divT[qq][1];
          divT[qq][2] = q == qq ? tz : divT[qq][2];
        }
      }
    }
    __syncthreads();
  }

  if constexpr (NV == 9) {
    if (live) {
    // div B in the order of SourceTerms_K's point loop (MHD_GLM_K.H:316-547),
    // then the source terms at the interior points, integrated (IntegrateRHS_K)
    constexpr int NZ = D == 3 ? NQ : 1;
    double divB = 0.0;
#pragma unroll
    for (int nQ = 0; nQ < NQ; nQ++)
#pragma unroll
      for (int nS = 0; nS < NZ; nS++) {
        const int q = nS * NQ + nQ;
        divB += divT[q][0];
        divB += divT[q][1];
        if (D == 3) divB += divT[q][2];
      }
    mhd_source_rhs<D, NQ>(U, FS, a.src, a.theta, g, a.Dm, c, ci, cj, ck, divB, rhs);
    }
  }
  const bool write = !a.sliceWrite || a.sliceWrite[slice] != 0;
  if (live) rk_epilogue<NV>(a, U, c, nc, rhs, betaDt, bhatDt, write, errLoc, posLoc);
  const int bid = b;  // slice-major: the bricks of a 2D slice are one contiguous range
  const double bmax = block_max(vmax, sh, tid);
  if (tid == 0 && bmax > 0.0) atomicMax(a.vmaxBits + slice, (unsigned long long)__double_as_longlong(bmax));
  if (a.mode == RK_FSAL) {
    const double es = block_sum(errLoc, sh, tid);
    const double pm = block_min(posLoc, sh, tid);
    if (tid == 0) {
      a.errPartial[bid] = es;
      a.posPartial[bid] = pm;
    }
  }
}

}  // namespace dev
}  // namespace kfvm

#include "recon5.hpp"
#include "fused5.hpp"
#include "mhd.hpp"
#include "indicator.hpp"

namespace kfvm {
// blocks of BLK threads covering n items
inline unsigned stage_blocks(int64_t n) { return unsigned((n + dev::BLK - 1) / dev::BLK); }
}  // namespace kfvm

namespace kfvm {
namespace dev {

// Flow statistics (FlowStats<Hydro> Hydro_K.H:103-173, FlowStats<MHD_GLM>
// MHD_GLM_K.H:108-207): per cell, with ifac = dvol / volume,
//   ekin += ifac |m|^2 / (2 rho)
//   enst += ifac rho |omega|^2 / 2, omega = curl u from central differences
//   heli += ifac B . curl B (GLM-MHD; central differences of B)
// in the reference's operation order; the neighbours come through the
// boundary index maps, as the reference's ghost cells.  One partial sum per
// block: out[3 blockIdx + s].
template <int D, int NV>
__global__ __launch_bounds__(BLK) void k_flow_stats(const double *__restrict__ U, Grid g, double ifac,
                                                    double *__restrict__ out) {
  __shared__ double sh[BLK];
  const int c = blockIdx.x * BLK + threadIdx.x;
  double ek = 0.0, et = 0.0, he = 0.0;
  if (c < g.ncell) {
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    auto u = [&](int di, int dj, int dk, int v) { return load_u<D, NV>(U, g, i + di, j + dj, k + dk, v); };
    const size_t nc = g.ncell;
    const double rho = U[c], mx = U[nc + c], my = U[2 * nc + c], mz = U[3 * nc + c];
    const double ekin = 0.5 * (mx * mx + my * my + mz * mz) / rho;
    const double dyu = 0.5 * (u(0, 1, 0, 1) / u(0, 1, 0, 0) - u(0, -1, 0, 1) / u(0, -1, 0, 0)) / g.dy;
    const double dxv = 0.5 * (u(1, 0, 0, 2) / u(1, 0, 0, 0) - u(-1, 0, 0, 2) / u(-1, 0, 0, 0)) / g.dx;
    double dzu = 0.0, dzv = 0.0, dxw = 0.0, dyw = 0.0;
    if (D == 3) {
      dzu = 0.5 * (u(0, 0, 1, 1) / u(0, 0, 1, 0) - u(0, 0, -1, 1) / u(0, 0, -1, 0)) / g.dz;
      dzv = 0.5 * (u(0, 0, 1, 2) / u(0, 0, 1, 0) - u(0, 0, -1, 2) / u(0, 0, -1, 0)) / g.dz;
      dxw = 0.5 * (u(1, 0, 0, 3) / u(1, 0, 0, 0) - u(-1, 0, 0, 3) / u(-1, 0, 0, 0)) / g.dx;
      dyw = 0.5 * (u(0, 1, 0, 3) / u(0, 1, 0, 0) - u(0, -1, 0, 3) / u(0, -1, 0, 0)) / g.dy;
    }
    const double omx = dyw - dzv, omy = dzu - dxw, omz = dxv - dyu;
    const double enst = 0.5 * rho * (omx * omx + omy * omy + omz * omz);
    ek = ifac * ekin;
    et = ifac * enst;
    if (NV == 9) {
      const double dybx = 0.5 * (u(0, 1, 0, 5) - u(0, -1, 0, 5)) / g.dy;
      const double dxby = 0.5 * (u(1, 0, 0, 6) - u(-1, 0, 0, 6)) / g.dx;
      double dzbx = 0.0, dzby = 0.0, dxbz = 0.0, dybz = 0.0;
      if (D == 3) {
        dzbx = 0.5 * (u(0, 0, 1, 5) - u(0, 0, -1, 5)) / g.dz;
        dzby = 0.5 * (u(0, 0, 1, 6) - u(0, 0, -1, 6)) / g.dz;
        dxbz = 0.5 * (u(1, 0, 0, 7) - u(-1, 0, 0, 7)) / g.dx;
        dybz = 0.5 * (u(0, 1, 0, 7) - u(0, -1, 0, 7)) / g.dy;
      }
      const double cbx = dybz - dzby, cby = dzbx - dxbz, cbz = dxby - dybx;
      const double bcb = U[5 * nc + c] * cbx + U[6 * nc + c] * cby + U[7 * nc + c] * cbz;
      he = ifac * bcb;
    }
  }
  const double se = block_sum(ek, sh, threadIdx.x);
  const double st = block_sum(et, sh, threadIdx.x);
  const double sb = block_sum(he, sh, threadIdx.x);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x] = se;
    out[3 * blockIdx.x + 1] = st;
    out[3 * blockIdx.x + 2] = sb;
  }
}

#ifdef KFVM_SMALL_KERNELS
// ------------------------------------------------------ small register kernels
template <int NV>
__global__ __launch_bounds__(BLK) void k_stage_pre(const double *__restrict__ Uprev,
                                                   const double *__restrict__ K,
                                                   double *__restrict__ U, double *__restrict__ Uhat,
                                                   double *__restrict__ Utmp, double betaDt,
                                                   double bhatDt, int ncell) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= ncell) return;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const size_t o = size_t(v) * ncell + c;
    const double up = Uprev[o], kk = K[o];
    U[o] = up + betaDt * kk;
    Uhat[o] = up + bhatDt * kk;
    if (Utmp) Utmp[o] = up;
  }
}

// Stage start of a batch of slices, each with its own step size.
template <int NV>
__global__ __launch_bounds__(BLK) void k_stage_pre_slices(const double *__restrict__ Uprev,
                                                          const double *__restrict__ K,
                                                          double *__restrict__ U, double *__restrict__ Uhat,
                                                          double *__restrict__ Utmp,
                                                          const double *__restrict__ sliceDt, double beta,
                                                          double bhat, int ncell, int cellsPerSlice) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= ncell) return;
  const double dts = sliceDt[c / cellsPerSlice];
  const double betaDt = beta * dts, bhatDt = bhat * dts;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const size_t o = size_t(v) * ncell + c;
    const double up = Uprev[o], kk = K[o];
    U[o] = up + betaDt * kk;
    Uhat[o] = up + bhatDt * kk;
    if (Utmp) Utmp[o] = up;
  }
}

// Accepted slices: U -> Uprev.
__global__ __launch_bounds__(BLK) void k_commit_slices(const double *__restrict__ U, double *__restrict__ Uprev,
                                                       const int *__restrict__ mask, int nv, int ncell,
                                                       int cellsPerSlice) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= ncell || !mask[c / cellsPerSlice]) return;
  for (int v = 0; v < nv; v++) Uprev[size_t(v) * ncell + c] = U[size_t(v) * ncell + c];
}

// Per-slice final reduction: one block per slice over its blocks' partials.
__global__ __launch_bounds__(BLK) void k_finalize_slices(const double *__restrict__ errPartial,
                                                         const double *__restrict__ posPartial,
                                                         int blocksPerSlice, double *__restrict__ out) {
  __shared__ double se[BLK], sp[BLK];
  const size_t b0 = size_t(blockIdx.x) * blocksPerSlice;
  double e = 0.0, p = 1.0;
  for (int t = threadIdx.x; t < blocksPerSlice; t += BLK) {
    e += errPartial[b0 + t];
    const double o = posPartial[b0 + t];
    p = p > o ? o : p;
  }
  se[threadIdx.x] = e;
  sp[threadIdx.x] = p;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      const double o = sp[threadIdx.x + s];
      sp[threadIdx.x] = sp[threadIdx.x] > o ? o : sp[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = se[0];
    out[2 * blockIdx.x + 1] = sp[0];
  }
}

// Deterministic final reduction of the per-block partials.
__global__ __launch_bounds__(1024) void k_finalize(const double *__restrict__ errPartial,
                                                   const double *__restrict__ posPartial,
                                                   int nPart, double *__restrict__ out) {
  __shared__ double se[1024], sp[1024];
  double e = 0.0, p = 1.0;
  for (int t = threadIdx.x; t < nPart; t += 1024) {
    e += errPartial[t];
    const double o = posPartial[t];
    p = p > o ? o : p;
  }
  se[threadIdx.x] = e;
  sp[threadIdx.x] = p;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      const double o = sp[threadIdx.x + s];
      sp[threadIdx.x] = sp[threadIdx.x] > o ? o : sp[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sp[0];
  }
}

__global__ void k_fill(double *__restrict__ p, double v, size_t n) {
  for (size_t t = size_t(blockIdx.x) * BLK + threadIdx.x; t < n; t += size_t(gridDim.x) * BLK) p[t] = v;
}

// Zero n 32-bit words (counters, speeds): a kernel, so that steps captured
// into HIP graphs hold no memset nodes.
__global__ void k_zero32(unsigned *__restrict__ p, size_t n) {
  for (size_t t = size_t(blockIdx.x) * BLK + threadIdx.x; t < n; t += size_t(gridDim.x) * BLK) p[t] = 0u;
}

#endif  // KFVM_SMALL_KERNELS

}  // namespace dev

}  // namespace kfvm
