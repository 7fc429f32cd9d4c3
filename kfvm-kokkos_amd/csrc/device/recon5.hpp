// Reconstruction kernels of the 5-variable (Euler) equations.
//
//   k_linrec_brick  Stencil_K.H:32-90: linear GP face states in conserved
//                   variables.  A block owns a brick of cells; each conserved
//                   variable of the brick plus its radius-2 halo is staged in
//                   LDS with coalesced loads (boundary conditions applied by
//                   the index map of the load), then every thread reads its 33
//                   stencil values at compile-time LDS offsets and forms the
//                   F*Q dot products with the uniform weights.
//   k_cell_props    per-cell internal energy and sound speed, the two
//                   neighbourhood quantities of the limiter bounds.
//   k_wenorec5      Stencil_K.H:164-283, 416-454, 506-589: one lane per
//                   (cell, variable) over the compacted WENO list.
//   k_linear_post   Numeric_K.H:315-379 indicator p/rho^gamma of every
//                   unlimited face state + Common_K.H:299-643 limiter.
//   k_weno_post     Hydro_K.H:53-64 back-conversion + Common_K.H:299-643.
#pragma once

namespace kfvm {
namespace dev {

constexpr int LPC = 5;                 // lanes per cell (k_wenorec5)
constexpr int CPW = 12;                // cells per wave
constexpr int CPB = CPW * (BLK / 64);  // cells per block

struct Group {
  int lane, slot, v;
  __device__ inline Group() {
    lane = threadIdx.x & 63;
    slot = lane / LPC;
    v = lane - LPC * slot;
    if (slot >= CPW) v = 0;  // idle lanes shadow variable 0
  }
};

// ----------------------------------------------------------- brick geometry
template <int D>
struct RBox {
  static constexpr int H = 2;  // stencil radius
  static constexpr int BX = D == 3 ? 8 : 16, BY = D == 3 ? 8 : 16, BZ = D == 3 ? 8 : 1;
  static constexpr int LX = BX + 2 * H, LY = BY + 2 * H, LZ = D == 3 ? BZ + 2 * H : 1;
  static constexpr int N = LX * LY * LZ;   // staged cells per variable
  static constexpr int NT = BX * BY * BZ;  // threads per block
};

// Workgroups are dispatched round-robin over the 8 XCDs; handing each XCD a
// contiguous range of bricks keeps neighbouring bricks (and their shared halo
// cells) in the same L2.
__device__ inline int xcd_block(int b, int nb) {
  if (nb % 8) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// distance of a box coordinate outside the brick interior [H, H+B)
template <int B, int H>
__device__ inline int box_out(int t) {
  return t < H ? H - t : (t >= B + H ? t - (B + H) + 1 : 0);
}

template <int D, int NQ>
__global__ __launch_bounds__(RBox<D>::NT) void k_linrec_brick(const double *__restrict__ U,
                                                              double *__restrict__ FS,
                                                              double *__restrict__ props, Grid g,
                                                              const double *__restrict__ W) {
  using Bx = RBox<D>;
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q, H = Bx::H;
  __shared__ double lds[Bx::N];
  const int nbx = (g.nx + Bx::BX - 1) / Bx::BX, nby = (g.ny + Bx::BY - 1) / Bx::BY;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (b % nbx) * Bx::BX, j0 = ((b / nbx) % nby) * Bx::BY, k0 = (b / (nbx * nby)) * Bx::BZ;
  const int tid = threadIdx.x;
  const int tx = tid % Bx::BX, ty = (tid / Bx::BX) % Bx::BY, tz = tid / (Bx::BX * Bx::BY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  const bool live = i < g.nx && j < g.ny && (D == 2 || k < g.nz);
  const size_t nc = g.ncell;
  const size_t cell = (size_t(k) * g.ny + j) * g.nx + i;
  const double *p = lds + (tz * Bx::LY + ty) * Bx::LX + tx;  // stencil box corner (-H,-H,-H)
  constexpr int ctr = H + Bx::LX * H + (D == 3 ? Bx::LX * Bx::LY * H : 0);
  double ref[5];
#pragma unroll 1
  for (int v = 0; v < 5; v++) {
    for (int idx = tid; idx < Bx::N; idx += Bx::NT) {
      const int a = idx % Bx::LX, r = idx / Bx::LX, bb = r % Bx::LY, c = r / Bx::LY;
      const int ex = box_out<Bx::BX, H>(a), ey = box_out<Bx::BY, H>(bb);
      const int ez = D == 3 ? box_out<Bx::BZ, H>(c) : 0;
      const int gi = i0 + a - H, gj = j0 + bb - H, gk = D == 3 ? k0 + c - H : 0;
      // cells no stencil of a live cell reads (box corners, beyond the halo)
      const bool need = ex * ex + ey * ey + ez * ez <= H * H && gi < g.nx + H && gj < g.ny + H &&
                        (D == 2 || gk < g.nz + H);
      if (need) lds[idx] = load_u<D, 5>(U, g, gi, gj, gk, v);
    }
    __syncthreads();
    if (live) {
      double vs[C];
#pragma unroll
      for (int n = 0; n < C; n++)
        vs[n] = p[(ST.off[n][0] + H) + Bx::LX * (ST.off[n][1] + H) +
                  (D == 3 ? Bx::LX * Bx::LY * (ST.off[n][2] + H) : 0)];
      ref[v] = p[ctr];
#pragma unroll
      for (int fq = 0; fq < FQ; fq++) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n < C; n++) acc = fma(W[fq * C + n], vs[n], acc);
        FS[(size_t(fq) * 5 + v) * nc + cell] = acc;
      }
    }
    __syncthreads();
  }
  if (live) {
    props[cell] = Hydro::eint(ref);
    props[nc + cell] = Hydro::sound(g.gamma, ref);
  }
}

template <int D>
__global__ __launch_bounds__(BLK) void k_cell_props(const double *__restrict__ U,
                                                    double *__restrict__ props, Grid g) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  const size_t nc = g.ncell;
  double s[5];
#pragma unroll
  for (int w = 0; w < 5; w++) s[w] = U[size_t(w) * nc + c];
  props[c] = Hydro::eint(s);
  props[nc + c] = Hydro::sound(g.gamma, s);
}

// Limiter bounds (Common_K.H:329-387) from the per-cell density, internal
// energy and sound speed of the 3^D neighbourhood.  Both are even in the
// momenta, so the mirror images of wall boundaries need no sign handling.
template <int D>
__device__ inline PPBounds pp_bounds_p(const double *__restrict__ U,
                                       const double *__restrict__ props, const Grid &g, int i,
                                       int j, int k, const double *ref) {
  const size_t nc = g.ncell;
  double maxDens = -1.7976931348623157e308, minDens = 1.7976931348623157e308;
  double minEint = 1.7976931348623157e308, minSdSp = 1.7976931348623157e308;
  constexpr int zl = D == 3 ? -1 : 0, zh = D == 3 ? 1 : 0;
#pragma unroll 1
  for (int xo = -1; xo <= 1; xo++) {
    int sx;
    const int ii = map_axis(i + xo, g.nx, g.bc[0], g.bc[1], sx);
#pragma unroll
    for (int yo = -1; yo <= 1; yo++) {
      int sy;
      const int jj = map_axis(j + yo, g.ny, g.bc[2], g.bc[3], sy);
#pragma unroll
      for (int zo = zl; zo <= zh; zo++) {
        int sz;
        const int kk = D == 3 ? map_axis(k + zo, g.nz, g.bc[4], g.bc[5], sz) : 0;
        const size_t c = (size_t(kk) * g.ny + jj) * g.nx + ii;
        const double rho = U[c];
        maxDens = fmax(maxDens, rho);
        minDens = fmin(minDens, rho);
        minEint = fmin(minEint, props[c]);
        minSdSp = fmin(minSdSp, props[nc + c]);
      }
    }
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
  PPBounds b;
  b.rDens = ref[0];
  b.maxDens = maxDens * (1.0 + kappa);
  b.minDens = minDens * (1.0 - kappa);
  b.minEint = minEint * (1.0 - kappa);
  return b;
}

// --------------------------------------------------- KFVM-WENO (faces only)
// One lane per (cell, variable): linearised-primitive stencil values,
// smoothness indicators of the full / centre / biased substencils, WENO-AO
// weights and the weighted face values, still in linearised primitive
// variables (k_weno_post converts them back).
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_wenorec5(const double *__restrict__ U,
                                                  double *__restrict__ FS, Grid g,
                                                  const double *__restrict__ FW,
                                                  const double *__restrict__ DW,
                                                  const int *__restrict__ list,
                                                  const int *__restrict__ counter) {
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, Cc = SInfo<D>::Cc, Cb = SInfo<D>::Cb;
  constexpr int nSub = SInfo<D>::nSub, I = SInfo<D>::I;
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  constexpr double gHi = 0.7, gLo = 0.7;
  constexpr double g0 = gHi, g1 = (1.0 - gHi) * gLo, gb = (1.0 - gHi) * (1.0 - gLo) / double(2 * D);
  const Group G;
  if (G.slot >= CPW) return;
  const int n = list ? *counter : g.ncell;
  const int waveId = blockIdx.x * (BLK / 64) + (threadIdx.x >> 6);
  const int nWaves = gridDim.x * (BLK / 64);
  const size_t nc = g.ncell;
#pragma unroll 1
  for (int t = waveId * CPW + G.slot; t < n; t += nWaves * CPW) {
    const int c = list ? list[t] : t;
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const LinPrim lp(ref, g.gamma);
    const double uv = G.v == 1 ? lp.u : (G.v == 2 ? lp.v : lp.w);
    double vs[C];
#pragma unroll
    for (int m = 0; m < C; m++) {
      const int a = i + ST.off[m][0], b = j + ST.off[m][1], d = k + ST.off[m][2];
      const double s0 = load_u<D, 5>(U, g, a, b, d, 0);
      if (G.v == 0) {
        vs[m] = s0;
      } else if (G.v < 4) {
        const double sv = load_u<D, 5>(U, g, a, b, d, G.v);
        vs[m] = (sv - uv * s0) / lp.rho;
      } else {
        const double s1 = load_u<D, 5>(U, g, a, b, d, 1);
        const double s2 = load_u<D, 5>(U, g, a, b, d, 2);
        const double s3 = load_u<D, 5>(U, g, a, b, d, 3);
        const double s4 = load_u<D, 5>(U, g, a, b, d, 4);
        vs[m] = lp.gamm * (lp.vsq * s0 - lp.u * s1 - lp.v * s2 - lp.w * s3 + s4);
      }
    }
    double om[nSub];
#pragma unroll
    for (int s = 0; s < nSub; s++) {
      const int ns = s == 0 ? C : (s == 1 ? Cc : Cb);
      double beta = 0.0;
#pragma unroll
      for (int ii = 0; ii < I; ii++) {
        const double *dd = DW + (s * I + ii) * C;
        double tt = 0.0;
#pragma unroll
        for (int m = 0; m < ns; m++) tt = fma(dd[m], vs[s < 2 ? m : ST.sub[s - 2][m]], tt);
        beta = fma(tt, tt, beta);
      }
      om[s] = beta;
    }
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
#pragma unroll
    for (int fq = 0; fq < FQ; fq++) {
      double acc = 0.0;
#pragma unroll
      for (int s = 0; s < nSub; s++) {
        const int ns = s == 0 ? C : (s == 1 ? Cc : Cb);
        const double *wq = FW + (s * FQ + fq) * C;
        double tt = 0.0;
#pragma unroll
        for (int m = 0; m < ns; m++) tt = fma(wq[m], vs[s < 2 ? m : ST.sub[s - 2][m]], tt);
        acc = fma(om[s], tt, acc);
      }
      FS[(size_t(fq) * 5 + G.v) * nc + c] = acc;
    }
  }
}

// ------------------------------------------- per-cell post-processing passes
// After linear reconstruction: KXRCF indicator p/rho^gamma of every unlimited
// face state, then the positivity limiter (fast test, exact reference
// sequence when a bound is violated).
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_linear_post(const double *__restrict__ U,
                                                     double *__restrict__ FS,
                                                     double *__restrict__ WI,
                                                     double *__restrict__ thetaNew,
                                                     const double *__restrict__ props, Grid g) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int i, j, k;
  cell_ijk<D>(c, g, i, j, k);
  const size_t nc = g.ncell;
  double ref[5];
#pragma unroll
  for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
  const PPBounds pb = pp_bounds_p<D>(U, props, g, i, j, k, ref);
  double thD = 1.0;
  bool okE = true;
#pragma unroll 2
  for (int fq = 0; fq < FQ; fq++) {
    double s[5];
#pragma unroll
    for (int w = 0; w < 5; w++) s[w] = FS[(size_t(fq) * 5 + w) * nc + c];
    thD = dens_theta(thD, s[0], pb);
    const double ei = Hydro::eint(s);
    okE = okE && ei > pb.minEint;
    WI[size_t(fq) * nc + c] = Hydro::weno(g.gamma, s);
  }
  double th = 1.0;
  if (!(thD == 1.0 && okE)) th = pp_apply<D, NQ>(FS, nc, c, ref, pb);
  thetaNew[c] = th;
}

// After WENO: back to conserved variables about the cell's own average
// (Hydro_K.H:53-64), then the positivity limiter.
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_weno_post(const double *__restrict__ U,
                                                   double *__restrict__ FS,
                                                   double *__restrict__ thetaNew,
                                                   const double *__restrict__ props, Grid g,
                                                   const int *__restrict__ list,
                                                   const int *__restrict__ counter) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  const int n = list ? *counter : g.ncell;
  const size_t nc = g.ncell;
  for (int t = blockIdx.x * BLK + threadIdx.x; t < n; t += gridDim.x * BLK) {
    const int c = list ? list[t] : t;
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const LinPrim lp(ref, g.gamma);
    const double ru = lp.rho * lp.u, rvv = lp.rho * lp.v, rw = lp.rho * lp.w;
    double thD = 1.0;
    bool okE = true;
    const PPBounds pb = pp_bounds_p<D>(U, props, g, i, j, k, ref);
#pragma unroll 2
    for (int fq = 0; fq < FQ; fq++) {
      double *p = FS + (size_t(fq) * 5) * nc + c;
      const double d = p[0], mx = p[nc], my = p[2 * nc], mz = p[3 * nc], e = p[4 * nc];
      double s[5];
      s[4] = (lp.vsq * d + ru * mx + rvv * my + rw * mz + e / lp.gamm);
      s[1] = (lp.rho * mx + lp.u * d);
      s[2] = (lp.rho * my + lp.v * d);
      s[3] = (lp.rho * mz + lp.w * d);
      s[0] = d;
#pragma unroll
      for (int w = 1; w < 5; w++) p[w * nc] = s[w];
      thD = dens_theta(thD, d, pb);
      okE = okE && Hydro::eint(s) > pb.minEint;
    }
    double th = 1.0;
    if (!(thD == 1.0 && okE)) th = pp_apply<D, NQ>(FS, nc, c, ref, pb);
    thetaNew[c] = th;
  }
}

}  // namespace dev
}  // namespace kfvm
