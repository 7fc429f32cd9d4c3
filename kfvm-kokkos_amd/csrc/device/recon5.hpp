// Reconstruction kernels of the 5-variable (Euler) equations.
//
//   k_linear_fused  Stencil_K.H:32-90 linear GP face states in conserved
//                   variables, the Numeric_K.H:315-379 indicator p/rho^gamma
//                   of every unlimited face state, and the Common_K.H:299-643
//                   limiter.  A block owns a brick of cells; the five
//                   conserved variables of the brick plus its radius-2 halo
//                   are staged in LDS with coalesced loads (boundary
//                   conditions applied by the index map of the load).
//   k_pp_lanes      the limiter's exact sequence for the non-WENO cells that
//                   failed the fast test of k_linear_fused.
//   k_cell_props    per-cell internal energy and sound speed, the two
//                   neighbourhood quantities of the limiter bounds.
//   k_wenorec5      Stencil_K.H:164-283, 416-454, 506-589: one lane per
//                   (cell, variable) over the compacted WENO list.
//   k_weno_post     Hydro_K.H:53-64 back-conversion + Common_K.H:299-643
//                   over the listed cells (k_weno_post_cells: every cell).
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

// The stencil tables as a static constexpr member, so that fully unrolled
// loops fold the offsets into immediates.
template <int D>
struct StTab {
  static constexpr StencilTables<D> ST = make_stencil_tables<D>();
};

// Read-only data in the constant address space: uniform loads become scalar
// (s_load) loads.
typedef __attribute__((address_space(4))) double cdouble;

// ----------------------------------------------------------- brick geometry
template <int D>
struct RBox {
  static constexpr int H = 2;  // stencil radius
  // 16 cells along x: a wave (16x4 cells) reads LDS rows of 16 doubles
  // without bank conflicts and stores whole 128-byte lines
  static constexpr int BX = 16, BY = D == 3 ? 4 : 16, BZ = D == 3 ? 4 : 1;
  static constexpr int LX = BX + 2 * H, LY = BY + 2 * H, LZ = D == 3 ? BZ + 2 * H : 1;
  static constexpr int N = LX * LY * LZ;   // staged cells per variable
  static constexpr int NT = BX * BY * BZ;  // threads per block
};

// distance of a box coordinate outside the brick interior [H, H+B)
template <int B, int H>
__device__ inline int box_out(int t) {
  return t < H ? H - t : (t >= B + H ? t - (B + H) + 1 : 0);
}

// Limiter bounds (Common_K.H:329-387) inside a staged brick: densities and
// velocities from LDS, the neighbours' internal energy and sound speed from
// the per-cell property planes (both are even in the momenta, so the mirror
// images of wall boundaries need no sign handling).  Bx: the brick shape
// (the LDS holds each variable's box as one plane of Bx::N cells).
template <int D, class Bx = RBox<D>>
__device__ inline PPBounds pp_bounds_lds(const double *__restrict__ lds, int c0,
                                         const double *__restrict__ props, const Grid &g, int i,
                                         int j, int k, const double *ref) {
  constexpr int SX = 1, SY = Bx::LX, SZ = Bx::LX * Bx::LY, NP = Bx::N;
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
        const int kk = D == 3 ? map_axis(k + zo, g.nz, g.bc[4], g.bc[5], sz) : k;
        const size_t c = (size_t(kk) * g.ny + jj) * g.nx + ii;
        const double rho = lds[c0 + xo * SX + yo * SY + zo * SZ];
        maxDens = fmax(maxDens, rho);
        minDens = fmin(minDens, rho);
        minEint = fmin(minEint, props[c]);
        minSdSp = fmin(minSdSp, props[nc + c]);
      }
    }
  }
  const double *r0 = lds + c0, *m1 = lds + NP + c0, *m2 = lds + 2 * NP + c0, *m3 = lds + 3 * NP + c0;
  const double dxU = m1[SX] / r0[SX] - m1[-SX] / r0[-SX];
  const double dyV = m2[SY] / r0[SY] - m2[-SY] / r0[-SY];
  double dzW;
  if (D == 3)
    dzW = m3[SZ] / r0[SZ] - m3[-SZ] / r0[-SZ];
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

// The linear states on the brick's outer faces (each face of the brick: the
// cells next to it, all Q points of their face on it), for the neighbouring
// bricks' flux kernel.  Work items are (face, variable, cell of the face)
// with the face's items padded to whole waves, so the weights of a wave are
// uniform (scalar loads); each item sums its Q points over the stencil in the
// same order as the main passes, so the values equal theirs bit for bit.
template <int D>
struct SurfItems {
  using Bx = RBox<D>;
  static constexpr int T[3] = {Bx::BY * Bx::BZ, Bx::BX * Bx::BZ, Bx::BX * Bx::BY};
  static constexpr int pad(int n) { return (n + 63) / 64 * 64; }
  static constexpr int slots(int f) { return pad(5 * T[f >> 1]); }
  static constexpr int total() {
    int s = 0;
    for (int f = 0; f < 2 * D; f++) s += slots(f);
    return s;
  }
};

template <int D, int NQ>
__device__ inline void surface_states(const double *__restrict__ lds, const Grid &g, const double *__restrict__ WT,
                                      double *__restrict__ SB, const Surf &sb, int i0, int j0, int k0, int tid) {
  using Bx = RBox<D>;
  using SI = SurfItems<D>;
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, Q = Shape<D, NQ>::Q, H = Bx::H, NP = Bx::N;
  constexpr int P = linear_pass_points(D, NQ);
  const int lastX = min(Bx::BX, g.nx - i0) - 1, lastY = min(Bx::BY, g.ny - j0) - 1;
  const int lastZ = D == 3 ? min(Bx::BZ, g.nz - k0) - 1 : 0;
#pragma unroll 1
  for (int slot = tid; slot < SI::total(); slot += Bx::NT) {
    // face f of the brick (uniform over the wave) and item e within it
    int f = 0, e = slot;
#pragma unroll
    for (int ff = 0; ff < 2 * D - 1; ff++)
      if (f == ff && e >= SI::slots(ff)) {
        e -= SI::slots(ff);
        f = ff + 1;
      }
    const int ax = f >> 1, Tf = SI::T[ax];
    if (e >= 5 * Tf) continue;
    const int v = e / Tf, tl = e % Tf;
    int x, y, z;
    if (ax == 0) {
      x = (f & 1) ? lastX : 0;
      y = tl % Bx::BY;
      z = tl / Bx::BY;
    } else if (ax == 1) {
      x = tl % Bx::BX;
      y = (f & 1) ? lastY : 0;
      z = tl / Bx::BX;
    } else {
      x = tl % Bx::BX;
      y = tl / Bx::BX;
      z = (f & 1) ? lastZ : 0;
    }
    if (x > lastX || y > lastY || z > lastZ) continue;
    const int corner = (z * Bx::LY + y) * Bx::LX + x;
    const double *p = lds + v * NP + corner;
    // this face's Q points sit at positions p0 .. p0 + Q - 1 of pass fq0 / P
    // f is the same on every lane of the wave (faces are padded to whole waves)
    const int fu = __builtin_amdgcn_readfirstlane(f);
    const int fq0 = fu * Q, pass = fq0 / P, p0 = fq0 % P;
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) acc[q] = 0.0;
    // the stencil in chunks, the weight pointer made opaque per chunk so that
    // at most one chunk's weights are in scalar registers at a time
    constexpr int CH = 3;
#pragma unroll
    for (int n0 = 0; n0 < C; n0 += CH) {
      const double *wt0 = WT + size_t(pass) * C * P + p0 + n0 * P;
      asm volatile("" : "+s"(wt0));
      const cdouble *wt = (const cdouble *)wt0;
#pragma unroll
      for (int m = 0; m < CH && n0 + m < C; m++) {
        const int n = n0 + m;
        const double xv = p[(ST.off[n][0] + H) + Bx::LX * (ST.off[n][1] + H) +
                            (D == 3 ? Bx::LX * Bx::LY * (ST.off[n][2] + H) : 0)];
#pragma unroll
        for (int q = 0; q < Q; q++) acc[q] = fma(wt[m * P + q], xv, acc[q]);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, Q, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    double *d = SB + surf_index<D, Q>(sb, g, f, i0 + x, j0 + y, k0 + z, 0, v);
    const size_t T = size_t(sb.T[ax]);
#pragma unroll
    for (int q = 0; q < Q; q++) d[size_t(q) * 5 * T] = acc[q];
  }
}

// The learned WENO selection on the linear pass' staged brick (Euler);
// defined with the learned indicator (indicator.hpp).
template <int D>
__device__ void learned_tail(double *lds, const LearnArgs &la, const Grid &g, int i0, int j0, int k0, int tid);

// Linear reconstruction + KXRCF indicator + limiter, one thread per cell of a
// staged brick.  All five conserved variables of the brick and its halo sit
// in LDS.  The face points are produced in passes of P points (both faces of
// an axis, or one face for 3D 3x3 points): for each variable the thread reads
// its stencil values at compile-time LDS offsets, forms the dot products with
// the uniform weights and stores the states; density and |m|^2 of the pass'
// points stay in registers until the energy completes them, which gives
// p/rho^gamma and the limiter's fast test without reading the states back.
//
// SURF = true (the fused flux path): no face state is stored except those of
// the brick's outer faces (surface_states), into the surface planes SB (see
// Surf), from which the neighbouring bricks' flux kernel reads them.
//
// LEARN = true: the learned WENO selection follows on the same brick
// (learned_tail), in place of a separate selection kernel that would stage
// the brick's density and pressure again.
template <int D, int NQ, bool SURF = false, bool LEARN = false>
__global__ __launch_bounds__(RBox<D>::NT, 3) void k_linear_fused(
    const double *__restrict__ U, double *__restrict__ FS, double *__restrict__ WI,
    double *__restrict__ thetaNew, const double *__restrict__ props, Grid g,
    const double *__restrict__ WT, int *__restrict__ ppList, int *__restrict__ ppCount,
    double *__restrict__ SB = nullptr, Surf sb = Surf{}, LearnArgs la = LearnArgs{}) {
  using Bx = RBox<D>;
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, Q = Shape<D, NQ>::Q, H = Bx::H, NP = Bx::N;
  constexpr int P = linear_pass_points(D, NQ), NPASS = 2 * D * Q / P;
  __shared__ double lds[5 * NP];
  const int nbx = (g.nx + Bx::BX - 1) / Bx::BX, nby = (g.ny + Bx::BY - 1) / Bx::BY;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (b % nbx) * Bx::BX, j0 = ((b / nbx) % nby) * Bx::BY, k0 = (b / (nbx * nby)) * Bx::BZ;
  const int tid = threadIdx.x;
  const size_t nc = g.ncell;
  for (int idx = tid; idx < NP; idx += Bx::NT) {
    const int a = idx % Bx::LX, r = idx / Bx::LX, bb = r % Bx::LY, c = r / Bx::LY;
    const int ex = box_out<Bx::BX, H>(a), ey = box_out<Bx::BY, H>(bb);
    const int ez = D == 3 ? box_out<Bx::BZ, H>(c) : 0;
    const int gi = i0 + a - H, gj = j0 + bb - H, gk = D == 3 ? k0 + c - H : k0;
    // cells no stencil of a live cell reads (box corners, beyond the halo)
    const bool need = ex * ex + ey * ey + ez * ez <= H * H && gi < g.nx + H && gj < g.ny + H &&
                      (D == 2 || gk < g.nz + H);
    if (need) {
      const CellRef cr = map_cell<D, 5>(g, gi, gj, gk);
#pragma unroll
      for (int w = 0; w < 5; w++) lds[w * NP + idx] = ld(U, nc, cr, w);
    }
  }
  __syncthreads();
  const int tx = tid % Bx::BX, ty = (tid / Bx::BX) % Bx::BY, tz = tid / (Bx::BX * Bx::BY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  if (i < g.nx && j < g.ny && (D == 2 || k < g.nz)) {  // no barrier inside
  const int cell = (k * g.ny + j) * g.nx + i;
  const int corner = (tz * Bx::LY + ty) * Bx::LX + tx;  // stencil box corner (-H,-H,-H)

  const int c0 = corner + H + Bx::LX * H + (D == 3 ? Bx::LX * Bx::LY * H : 0);
  double ref[5];
#pragma unroll
  for (int w = 0; w < 5; w++) ref[w] = lds[w * NP + c0];
  const PPBounds pb = pp_bounds_lds<D>(lds, c0, props, g, i, j, k, ref);
  const double gm1 = g.gamma - 1.0;
  double thD = 1.0;
  bool okE = true;
  // One pass: the P points of all five variables.  The stencil walk is
  // outermost (each accumulator still sums n = 0..C-1 in order), so the P
  // weights of one stencil cell -- one scalar load, shared by the five
  // variables -- feed 5P FMAs.  The next stencil cell's weights and its five
  // LDS values are fetched before this cell's FMAs issue (software
  // pipelining): the wait for them falls after 5P FMAs instead of before each
  // group of FMAs.  The weight pointer is made opaque once per stencil cell,
  // which keeps the scalar loads of later cells from being hoisted (they
  // would park their values in VGPR lanes).
  auto pass = [&](int ax, double (&acc)[5][P]) {
    const double *pc = lds + corner;
    const double *wbase = WT + ax * C * P;
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
      for (int q = 0; q < P; q++) acc[v][q] = 0.0;
    double w[P], x[5];
    {
      const double *w0 = wbase;
      asm volatile("" : "+s"(w0));
      const cdouble *wc = (const cdouble *)w0;
#pragma unroll
      for (int q = 0; q < P; q++) w[q] = wc[q];
      constexpr int o = ST.off[0][0] + H + Bx::LX * (ST.off[0][1] + H) +
                        (D == 3 ? Bx::LX * Bx::LY * (ST.off[0][2] + H) : 0);
#pragma unroll
      for (int v = 0; v < 5; v++) x[v] = pc[v * NP + o];
    }
#pragma unroll
    for (int n = 0; n < C; n++) {
      double wn[P], xn[5];
      if (n + 1 < C) {
        // The next cell's addresses pass through an empty asm that takes this
        // cell's weights and values as inputs: the wait for this cell's loads
        // (one counter serves scalar and LDS loads, and scalar loads may
        // return out of order, so a wait with a scalar load outstanding is a
        // full one) falls before the next cell's loads are issued, not after.
        const double *w1 = wbase + (n + 1) * P;
        int cn = corner;
        asm volatile("" : "+s"(w1), "+v"(cn) : "s"(w[0]), "s"(w[P - 1]), "v"(x[0]), "v"(x[1]), "v"(x[2]),
                     "v"(x[3]), "v"(x[4]));
        const cdouble *wc = (const cdouble *)w1;
#pragma unroll
        for (int q = 0; q < P; q++) wn[q] = wc[q];
        const int o = (ST.off[n + 1][0] + H) + Bx::LX * (ST.off[n + 1][1] + H) +
                      (D == 3 ? Bx::LX * Bx::LY * (ST.off[n + 1][2] + H) : 0);
#pragma unroll
        for (int v = 0; v < 5; v++) xn[v] = lds[v * NP + cn + o];
      }
#pragma unroll
      for (int q = 0; q < P; q++)
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v][q] = fma(w[q], x[v], acc[v][q]);
      if (n + 1 < C) {
        // the next cell's LDS reads, then this cell's FMAs
        __builtin_amdgcn_sched_group_barrier(0x100, 5, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 5 * P, 0);
#pragma unroll
        for (int q = 0; q < P; q++) w[q] = wn[q];
#pragma unroll
        for (int v = 0; v < 5; v++) x[v] = xn[v];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };
#pragma unroll 1
  for (int ax = 0; ax < NPASS; ax++) {
    double acc[5][P];
    pass(ax, acc);
    if (FS) {
#pragma unroll
      for (int q = 0; q < P; q++)
#pragma unroll
        for (int v = 0; v < 5; v++)
          FS[fs_base(cell, Shape<D, NQ>::F * Shape<D, NQ>::Q * 5) + (size_t(ax * P + q) * 5 + v) * FSB] =
              acc[v][q];
    }
#pragma unroll
    for (int q = 0; q < P; q++) {
      thD = dens_theta(thD, acc[0][q], pb);
      // 0 + m^2 == m^2 exactly: the same sums as Hydro_EOS.H
      double msq = 0.0;
      msq = msq + acc[1][q] * acc[1][q];
      msq = msq + acc[2][q] * acc[2][q];
      msq = msq + acc[3][q] * acc[3][q];
      // Hydro::eint / pres / weno of the point (Hydro_EOS.H)
      const double ei = acc[4][q] - 0.5 * msq / acc[0][q];
      okE = okE && ei > pb.minEint;
      if (WI) {  // the KXRCF indicator's input; not needed by the learned selection
        WI[fs_base(cell, Shape<D, NQ>::F * Shape<D, NQ>::Q) + size_t(ax * P + q) * FSB] =
            (gm1 * ei) / pow(acc[0][q], g.gamma);
        __builtin_amdgcn_sched_barrier(0);  // one pow at a time: bounded registers
      }
    }
  }
  // Cells whose states violate a bound go to a list for k_pp_lanes (the
  // limiter's exact sequence re-reads all states; keeping it out of this
  // kernel keeps its waves convergent).
  const bool slow = !(thD == 1.0 && okE);
  if (!slow) thetaNew[cell] = 1.0;
  const unsigned long long mask = __ballot(slow);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(ppCount, __popcll(mask));
    base = __shfl(base, leader);
    if (slow) ppList[base + __popcll(mask & ((1ull << lane) - 1ull))] = cell;
  }
  }
  // every thread of the block, live or not, takes surface items
  if (SURF) surface_states<D, NQ>(lds, g, WT, SB, sb, i0, j0, k0, tid);
  // every thread of the block, live or not, runs the selection (barriers, MFMA)
  if (LEARN) learned_tail<D>(lds, la, g, i0, j0, k0, tid);
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

// --------------------------------------------------- KFVM-WENO (faces only)
// One lane per (cell, variable): linearised-primitive stencil values,
// smoothness indicators of the full / centre / biased substencils, WENO-AO
// weights and the weighted face values, still in linearised primitive
// variables (k_weno_post converts them back).  Every matrix-vector product
// walks the stencil in its outer loop with one accumulator per output, so
// each accumulator still sums the stencil cells in ascending order (the
// reference's unblocked GEMV/GEMM order) while one stencil cell's weights
// arrive as one contiguous scalar load.  Weight layouts (stencil-major):
// FT[nSub][C][F*Q], DT[nSub][C][I].
// Per-substencil pieces as templates, so that each one is a small fully
// unrolled body with compile-time stencil indices (one fully unrolled loop
// nest over all substencils exceeds the unroller's limits and would leave the
// stencil values in scratch).

template <int D, int S>
struct SubSten {
  static constexpr int n = S == 0 ? SInfo<D>::C : (S == 1 ? SInfo<D>::Cc : SInfo<D>::Cb);
  __device__ static constexpr int idx(int m) {
    if constexpr (S < 2) return m;
    else return StTab<D>::ST.sub[S - 2][m];
  }
};

// First row of substencil S in a weight table: the padded layout gives every
// substencil C rows ([nSub][C][...]); the compact one stores only its n_S
// rows, substencils back to back.
template <int D, int S>
struct SubOff {
  static constexpr int value = SubOff<D, S - 1>::value + SubSten<D, S - 1>::n;
};
template <int D>
struct SubOff<D, 0> {
  static constexpr int value = 0;
};
template <int D, int S, bool COMPACT>
struct WRow {
  static constexpr int value = COMPACT ? SubOff<D, S>::value : S * SInfo<D>::C;
};

// beta_S = sum_i (D_S V_S)_i^2 (findSmoothness, Stencil_K.H:506-554); dt is a
// scalar-cache (cdouble) or LDS table, padded or compact
template <int D, int S, bool COMPACT = false, class WP = const cdouble *>
__device__ __attribute__((always_inline)) inline double weno_beta(const double (&vs)[SInfo<D>::C], WP dt) {
  constexpr int I = SInfo<D>::I, ns = SubSten<D, S>::n, r0 = WRow<D, S, COMPACT>::value;
  double acc[I];
#pragma unroll
  for (int ii = 0; ii < I; ii++) acc[ii] = 0.0;
#pragma unroll
  for (int m = 0; m < ns; m++) {
    const double x = vs[SubSten<D, S>::idx(m)];
#pragma unroll
    for (int ii = 0; ii < I; ii++) acc[ii] = fma(dt[(r0 + m) * I + ii], x, acc[ii]);
  }
  double beta = 0.0;
#pragma unroll
  for (int ii = 0; ii < I; ii++) beta = fma(acc[ii], acc[ii], beta);
  return beta;
}

// out[h] += omega_S * (W_S V_S)[h] over HF face points (fillRiemannStates,
// Stencil_K.H:416-454)
template <int D, int S, int HF, int FQ, bool COMPACT = false, class WP = const cdouble *>
__device__ __attribute__((always_inline)) inline void weno_faces(const double (&vs)[SInfo<D>::C], WP fc, double om,
                                                               double (&out)[HF]) {
  constexpr int ns = SubSten<D, S>::n, r0 = WRow<D, S, COMPACT>::value;
  double acc[HF];
#pragma unroll
  for (int h = 0; h < HF; h++) acc[h] = 0.0;
#pragma unroll
  for (int m = 0; m < ns; m++) {
    const double x = vs[SubSten<D, S>::idx(m)];
#pragma unroll
    for (int h = 0; h < HF; h++) acc[h] = fma(fc[(r0 + m) * FQ + h], x, acc[h]);
  }
#pragma unroll
  for (int h = 0; h < HF; h++) out[h] = fma(om, acc[h], out[h]);
}

template <int D, int S, int N, bool COMPACT = false>
struct WenoSubs {
  template <class WP>
  __device__ __attribute__((always_inline)) static inline void beta(const double (&vs)[SInfo<D>::C], WP dt,
                                                                    double *om) {
    om[S] = weno_beta<D, S, COMPACT>(vs, dt);
    __builtin_amdgcn_sched_barrier(0);
    WenoSubs<D, S + 1, N, COMPACT>::beta(vs, dt, om);
  }
  template <int HF, int FQ, class WP>
  __device__ __attribute__((always_inline)) static inline void faces(const double (&vs)[SInfo<D>::C], WP fc,
                                                                     const double *om, double (&out)[HF]) {
    weno_faces<D, S, HF, FQ, COMPACT>(vs, fc, om[S], out);
    __builtin_amdgcn_sched_barrier(0);
    WenoSubs<D, S + 1, N, COMPACT>::template faces<HF, FQ>(vs, fc, om, out);
  }
};
template <int D, int N, bool COMPACT>
struct WenoSubs<D, N, N, COMPACT> {
  template <class WP>
  __device__ static inline void beta(const double (&)[SInfo<D>::C], WP, double *) {}
  template <int HF, int FQ, class WP>
  __device__ static inline void faces(const double (&)[SInfo<D>::C], WP, const double *, double (&)[HF]) {}
};

template <int D, int NQ>
struct WenoChunk {  // face points per accumulation chunk
  static constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  static constexpr int H = D == 3 ? (NQ == 2 ? 12 : 9) : FQ;
};

// Listed cells at list positions t < cap write their states compactly to
// WS[t][point * 5 + var] (whole cache lines, read back coalesced by
// k_weno_post); the others, and the unlisted full pass, write FS in place.
//
// The weights (compact: WC = face rows [SUB][F*Q], then derivative rows
// [SUB][I], SUB = C + Cc + B*Cb) are staged in LDS once per workgroup and read
// from there: the scalar cache (16 KB) cannot hold them (36 KB in 3D at NQ=2),
// so scalar loads of them kept missing to L2 with the next FMAs waiting on
// them.  The grid is persistent (at most wenoBlocks workgroups striding over
// the list), so the staging is paid once per workgroup.  Every lane streams
// one weight per FMA out of LDS, so the kernel lives on LDS latency: three
// waves per SIMD (at most 168 VGPRs, no spills) run it 11-14 % faster than the
// two that 256 VGPRs allow; four (128 VGPRs) spill.
template <int D, int NQ>
struct WenoLds {
  static constexpr int SUB = SInfo<D>::C + SInfo<D>::Cc + (SInfo<D>::nSub - 2) * SInfo<D>::Cb;
  static constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  static constexpr int NFACE = SUB * FQ, NW = SUB * (FQ + SInfo<D>::I);
};

template <int D, int NQ>
__global__ __launch_bounds__(BLK, 3) void k_wenorec5(const double *__restrict__ U,
                                                  double *__restrict__ FS, Grid g,
                                                  const double *__restrict__ WC,
                                                  const int *__restrict__ list,
                                                  const int *__restrict__ counter,
                                                  double *__restrict__ WS, int cap) {
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, Cc = SInfo<D>::Cc, Cb = SInfo<D>::Cb;
  constexpr int nSub = SInfo<D>::nSub, I = SInfo<D>::I;
  constexpr int FQ = WenoChunk<D, NQ>::FQ, HF = WenoChunk<D, NQ>::H;
  constexpr double gHi = 0.7, gLo = 0.7;
  constexpr double g0 = gHi, g1 = (1.0 - gHi) * gLo, gb = (1.0 - gHi) * (1.0 - gLo) / double(2 * D);
  using WL = WenoLds<D, NQ>;
  static_assert(WL::SUB == SubOff<D, nSub>::value, "compact rows of all substencils");
  __shared__ double wl[WL::NW];
  for (int e = threadIdx.x; e < WL::NW; e += BLK) wl[e] = WC[e];
  __syncthreads();
  const Group G;
  if (G.slot >= CPW) return;  // no barrier follows
  const int n = list ? *counter : g.ncell;
  const int waveId = blockIdx.x * (BLK / 64) + (threadIdx.x >> 6);
  const int nWaves = gridDim.x * (BLK / 64);
  const size_t nc = g.ncell;
#pragma unroll 1
  for (int t = waveId * CPW + G.slot; t < n; t += nWaves * CPW) {
    const int c = list ? list[t] : t;
    const bool compact = list && t < cap;
    // state q of this cell at P[q * ps]
    double *const P = compact ? WS + size_t(t) * (FQ * 5) : FS + fs_base(c, FQ * 5);
    const size_t ps = compact ? 1 : FSB;
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const LinPrim lp(ref, g.gamma);
    const double uv = G.v == 1 ? lp.u : (G.v == 2 ? lp.v : lp.w);
    double vs[C];
    // each lane loads its own variable of the stencil cells; the density and
    // momenta the other variables' conversions need come from the cell's
    // other lanes (lanes 5 slot + w hold variable w)
    const int gbase = LPC * G.slot;
#pragma unroll
    for (int m = 0; m < C; m++) {
      const CellRef r = map_cell<D, 5>(g, i + ST.off[m][0], j + ST.off[m][1], k + ST.off[m][2]);
      const double sv = ld(U, nc, r, G.v);
      const double s0 = __shfl(sv, gbase);
      const double s1 = __shfl(sv, gbase + 1), s2 = __shfl(sv, gbase + 2), s3 = __shfl(sv, gbase + 3);
      if (G.v == 4) {
        vs[m] = lp.gamm * (lp.vsq * s0 - lp.u * s1 - lp.v * s2 - lp.w * s3 + sv);
      } else {
        vs[m] = G.v == 0 ? s0 : (sv - uv * s0) / lp.rho;
      }
    }
    // smoothness indicators (findSmoothness)
    const double *dt = wl + WL::NFACE;
    double om[nSub];
    WenoSubs<D, 0, nSub, true>::beta(vs, dt, om);
    // createOmegas
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
    // fillRiemannStates, HF face points at a time
#pragma unroll 1
    for (int h0 = 0; h0 < FQ; h0 += HF) {
      const double *fc = wl + h0;
      double out[HF];
#pragma unroll
      for (int h = 0; h < HF; h++) out[h] = 0.0;
      WenoSubs<D, 0, nSub, true>::template faces<HF, FQ>(vs, fc, om, out);
#pragma unroll
      for (int h = 0; h < HF; h++) P[(size_t(h0 + h) * 5 + G.v) * ps] = out[h];
    }
  }
}

// KFVM-WENO on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  The weight
// products of the lane-per-(cell, var) kernel above stream one LDS weight
// per FMA and every lane of a wave reads the same weight, so that kernel is
// bound by LDS bandwidth.  Here a wave takes 16 listed cells (the 16 columns
// of the products) one variable at a time and each substencil S multiplies
// its weight matrix A_S (rows: the I smoothness derivatives, then the F*Q face
// points; columns: the substencil's cells) by the cells' linearised values:
//   acc_S = A_S V_S,  beta_S = sum_i acc_S[i]^2 (i < I),
//   out   = sum_S om_S acc_S[I + h]
// with the WENO-AO weights om_S of the scalar kernel rewritten as
//   om_S = c_S / tot, c_0 = w_0 / g_0, c_S = w_S - c_0 g_S, w_S = g_S / (beta_S^2 + 1e-40),
//   tot = sum_S w_S,
// so each substencil's part is added as soon as its beta is known.  The
// weight fragments (A: lane l holds row l % 16, column l / 16 of a 16x4
// block) are staged in LDS once per workgroup in lane order; the values (B:
// lane l holds stencil row l / 16 of column l % 16) are written to a per-wave
// LDS plane by the lanes that load them; element q of the result holds row
// 4 q + l / 16 of column l % 16.  Sums run in a different order than
// the reference's loops: results agree to rounding.
template <int D, int NQ>
struct WenoMfma {
  static constexpr int C = SInfo<D>::C, I = SInfo<D>::I, nSub = SInfo<D>::nSub;
  static constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  static constexpr int RT = (I + FQ + 15) / 16;  // 16-row tiles of [derivatives; faces]
  static constexpr int CP = (C + 3) / 4 * 4;     // stencil rows of the value plane
  static constexpr int kt(int S) { return ((S == 0 ? C : (S == 1 ? SInfo<D>::Cc : SInfo<D>::Cb)) + 3) / 4; }
  static constexpr int fragOff(int S) { return S == 0 ? 0 : fragOff(S - 1) + RT * kt(S - 1); }
  static constexpr int NFRAG = fragOff(nSub);
  static constexpr int WPB = BLK / 64;  // waves per block
  static constexpr int LDS = (NFRAG * 64 + WPB * CP * 16) * int(sizeof(double));
  // 3D at NQ = 2 (two workgroups per CU): 17 % faster than the scalar kernel
  // on the listed cells of the 256^3 planar Riemann run.  In 2D (13-cell
  // stencils: 14 small products per variable) and at 3D NQ = 3 (74 KB of
  // fragments, one workgroup per CU) the scalar kernel is faster.
  static constexpr bool USE = D == 3 && NQ == 2 && LDS <= 80 * 1024;
};

typedef double v4d __attribute__((ext_vector_type(4)));

// stencil index of row m of substencil S (m beyond the substencil: the
// centre cell, whose weight there is zero)
template <int D, int S>
__device__ constexpr int msub(int m) {
  return m < SubSten<D, S>::n ? SubSten<D, S>::idx(m) : 0;
}

template <int D, int NQ, int S>
struct MfmaSub {
  using M = WenoMfma<D, NQ>;
  // acc = A_S V_S over the substencil's column tiles; vp: this wave's value
  // plane, grp = lane / 16, col = lane % 16
  __device__ __attribute__((always_inline)) static inline void run(const double *wf, const double *vp, int lane,
                                                                   int grp, int col, v4d (&acc)[M::RT]) {
#pragma unroll
    for (int r = 0; r < M::RT; r++) acc[r] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < M::kt(S); k++) {
      const int m0 = msub<D, S>(4 * k), m1 = msub<D, S>(4 * k + 1), m2 = msub<D, S>(4 * k + 2),
                m3 = msub<D, S>(4 * k + 3);
      const int m = grp == 0 ? m0 : (grp == 1 ? m1 : (grp == 2 ? m2 : m3));
      const double b = vp[m * 16 + col];
#pragma unroll
      for (int r = 0; r < M::RT; r++) {
        const double a = wf[(M::fragOff(S) + k * M::RT + r) * 64 + lane];
        acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[r], 0, 0, 0);
      }
    }
  }
};

// Offsets of stencil rows g, g + 4, g + 8, ... (rows past the stencil: the
// centre), three rows per word at 9 bits each ((dx + 2) | (dy + 2) << 3 |
// (dz + 2) << 6), so a lane keeps its rows' offsets in three registers.
template <int D>
struct MfmaRows {
  static constexpr int C = SInfo<D>::C, NJ = (C + 3) / 4, NW = (NJ + 2) / 3;
  unsigned w[4][NW];
};
template <int D>
constexpr MfmaRows<D> make_mfma_rows() {
  MfmaRows<D> t{};
  constexpr auto ST = make_stencil_tables<D>();
  for (int g = 0; g < 4; g++)
    for (int jj = 0; jj < MfmaRows<D>::NJ; jj++) {
      const int m = g + 4 * jj < MfmaRows<D>::C ? g + 4 * jj : 0;
      const unsigned code = unsigned(ST.off[m][0] + 2) | unsigned(ST.off[m][1] + 2) << 3 |
                            unsigned(ST.off[m][2] + 2) << 6;
      t.w[g][jj / 3] |= code << (9 * (jj % 3));
    }
  return t;
}

template <int D, int NQ>
__global__ __launch_bounds__(BLK, 2) void k_wenorec5_mfma(const double *__restrict__ U,
                                                        double *__restrict__ FS, Grid g,
                                                        const double *__restrict__ WC,
                                                        const int *__restrict__ list,
                                                        const int *__restrict__ counter,
                                                        double *__restrict__ WS, int cap) {
  using M = WenoMfma<D, NQ>;
  constexpr int C = M::C, I = M::I, nSub = M::nSub, FQ = M::FQ, RT = M::RT, CP = M::CP;
  constexpr int SUB = WenoLds<D, NQ>::SUB;
  constexpr double gHi = 0.7, gLo = 0.7;
  constexpr double g0 = gHi, g1 = (1.0 - gHi) * gLo, gb = (1.0 - gHi) * (1.0 - gLo) / double(2 * D);
  __shared__ double wf[M::NFRAG * 64];
  __shared__ double vplane[M::WPB][CP * 16];
  // weight fragments in lane order: fragment (S, k, r) holds rows 16 r .. 16 r + 15
  // and substencil columns 4 k .. 4 k + 3
  for (int e = threadIdx.x; e < M::NFRAG * 64; e += BLK) {
    const int f = e >> 6, l = e & 63;
    int S = 0;
#pragma unroll
    for (int q = 1; q < nSub; q++) S += f >= M::fragOff(q) ? 1 : 0;
    const int ns = S == 0 ? C : (S == 1 ? SInfo<D>::Cc : SInfo<D>::Cb);
    const int off = S == 0 ? 0 : (S == 1 ? C : C + SInfo<D>::Cc + (S - 2) * SInfo<D>::Cb);
    const int fl = f - M::fragOff(S), k = fl / RT, r = fl % RT;
    const int row = 16 * r + (l & 15), m = 4 * k + (l >> 4);
    double w = 0.0;
    if (m < ns) {
      if (row < I) w = WC[size_t(SUB) * FQ + size_t(off + m) * I + row];
      else if (row < I + FQ) w = WC[size_t(off + m) * FQ + (row - I)];
    }
    wf[e] = w;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, grp = lane >> 4, col = lane & 15;
  double *const vp = vplane[threadIdx.x >> 6];
  const int n = list ? *counter : g.ncell;
  const int waveId = blockIdx.x * M::WPB + (threadIdx.x >> 6);
  const int nWaves = gridDim.x * M::WPB;
  const size_t nc = g.ncell;
  constexpr int NJ = (C + 3) / 4;  // stencil rows per lane group
  constexpr auto RW = make_mfma_rows<D>();
  unsigned rowWord[MfmaRows<D>::NW];
#pragma unroll
  for (int q = 0; q < MfmaRows<D>::NW; q++)
    rowWord[q] = grp == 0 ? RW.w[0][q] : (grp == 1 ? RW.w[1][q] : (grp == 2 ? RW.w[2][q] : RW.w[3][q]));
#pragma unroll 1
  for (int t0 = waveId * 16; t0 < n; t0 += nWaves * 16) {
    const int t = t0 + col;
    const bool valid = t < n;
    const int tt = valid ? t : n - 1;
    const int c = list ? list[tt] : tt;
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const LinPrim lp(ref, g.gamma);
    // the stencil offsets are unpacked again per group of cells: kept
    // unpacked across the loop they would take 27 registers and spill
    unsigned rw[MfmaRows<D>::NW];
#pragma unroll
    for (int q = 0; q < MfmaRows<D>::NW; q++) {
      rw[q] = rowWord[q];
      asm volatile("" : "+v"(rw[q]));
    }
    // linearised values of all five variables at this lane's stencil rows
    // (grp, grp + 4, ...), loaded together once per group of cells
    double lv[5][NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; jj++) {
      const unsigned code = (rw[jj / 3] >> (9 * (jj % 3))) & 511u;
      const int dx = int(code & 7u) - 2, dy = int((code >> 3) & 7u) - 2, dz = int(code >> 6) - 2;
      const CellRef r = map_cell<D, 5>(g, i + dx, j + dy, k + dz);
      const double s0 = ld(U, nc, r, 0), s1 = ld(U, nc, r, 1), s2 = ld(U, nc, r, 2), s3 = ld(U, nc, r, 3),
                   s4 = ld(U, nc, r, 4);
      lv[0][jj] = s0;
      lv[1][jj] = (s1 - lp.u * s0) / lp.rho;
      lv[2][jj] = (s2 - lp.v * s0) / lp.rho;
      lv[3][jj] = (s3 - lp.w * s0) / lp.rho;
      lv[4][jj] = lp.gamm * (lp.vsq * s0 - lp.u * s1 - lp.v * s2 - lp.w * s3 + s4);
    }
    const bool compact = list && t < cap;
    double *const P = compact ? WS + size_t(tt) * (FQ * 5) : FS + fs_base(c, FQ * 5);
    const size_t ps = compact ? 1 : FSB;
#pragma unroll 1
    for (int v = 0; v < 5; v++) {
      int gq = grp;
      asm volatile("" : "+v"(gq));
      double vals[NJ];
#pragma unroll
      for (int jj = 0; jj < NJ; jj++)
        vals[jj] = v == 0 ? lv[0][jj] : (v == 1 ? lv[1][jj] : (v == 2 ? lv[2][jj] : (v == 3 ? lv[3][jj] : lv[4][jj])));
      // the previous variable's reads of the plane are done before it is rewritten
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
      for (int jj = 0; jj < NJ; jj++)
        if (grp + 4 * jj < CP) vp[(grp + 4 * jj) * 16 + col] = vals[jj];
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      v4d out[RT], acc[RT];
#pragma unroll
      for (int r = 0; r < RT; r++) out[r] = v4d{0.0, 0.0, 0.0, 0.0};
      double tot = 0.0, c0 = 0.0;
      auto fold = [&](int S) {
        // beta of this column: rows < I of acc, summed over the lane group
        double b = 0.0;
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int row = 16 * r + 4 * q + grp;
            b = row < I ? fma(acc[r][q], acc[r][q], b) : b;
          }
        b += __shfl_xor(b, 16);
        b += __shfl_xor(b, 32);
        const double gs = S == 0 ? g0 : (S == 1 ? g1 : gb);
        const double w = gs / (b * b + 1.e-40);
        tot += w;
        double cs;
        if (S == 0) {
          c0 = w / g0;
          cs = c0;
        } else {
          cs = w - c0 * gs;
        }
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
          for (int q = 0; q < 4; q++) out[r][q] = fma(cs, acc[r][q], out[r][q]);
      };
      __builtin_amdgcn_sched_barrier(0);
      MfmaSub<D, NQ, 0>::run(wf, vp, lane, gq, col, acc);
      fold(0);
      __builtin_amdgcn_sched_barrier(0);
      MfmaSub<D, NQ, 1>::run(wf, vp, lane, gq, col, acc);
      fold(1);
      __builtin_amdgcn_sched_barrier(0);
      MfmaSub<D, NQ, 2>::run(wf, vp, lane, gq, col, acc);
      fold(2);
      __builtin_amdgcn_sched_barrier(0);
      MfmaSub<D, NQ, 3>::run(wf, vp, lane, gq, col, acc);
      fold(3);
      __builtin_amdgcn_sched_barrier(0);
      MfmaSub<D, NQ, 4>::run(wf, vp, lane, gq, col, acc);
      fold(4);
      __builtin_amdgcn_sched_barrier(0);
      MfmaSub<D, NQ, 5>::run(wf, vp, lane, gq, col, acc);
      fold(5);
      if constexpr (nSub > 6) {
        __builtin_amdgcn_sched_barrier(0);
        MfmaSub<D, NQ, 6>::run(wf, vp, lane, gq, col, acc);
        fold(6);
        __builtin_amdgcn_sched_barrier(0);
        MfmaSub<D, NQ, 7>::run(wf, vp, lane, gq, col, acc);
        fold(7);
      }
      const double it = 1.0 / tot;
      if (valid) {
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int h = 16 * r + 4 * q + grp - I;
            if (h >= 0 && h < FQ) P[(size_t(h) * 5 + v) * ps] = out[r][q] * it;
          }
      }
    }
  }
}

// Positivity limiter (Common_K.H:299-643), exact sequence, for the cells
// whose linear states failed the fast test of k_linear_fused, skipping the
// cells the selector handed to WENO (k_weno_post limits their new states).
// Most listed cells are skipped, so a group of LPG lanes works on one cell, one lane
// per face point; the bounds' 3^D neighbourhood is spread over the group and
// every min / max over points is a group reduction (min and max do not depend
// on the order of their operands, so the result is that of the sequential
// loops).
template <int FQ>
struct PPGroup {
  static constexpr int LPG = FQ <= 8 ? 8 : (FQ <= 16 ? 16 : (FQ <= 32 ? 32 : 64));
};

template <int LPG>
__device__ inline double grp_min(double x) {
#pragma unroll
  for (int o = LPG / 2; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o));
  return x;
}
template <int LPG>
__device__ inline double grp_max(double x) {
#pragma unroll
  for (int o = LPG / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
  return x;
}

// Limiter bounds (Common_K.H:329-387) from the per-cell density, internal
// energy and sound speed of the 3^D neighbourhood, spread over a group of LPG
// lanes (lane pl of the group); every lane returns the bounds.  Both
// quantities are even in the momenta, so the mirror images of wall
// boundaries need no sign handling.
template <int D, int LPG>
__device__ inline PPBounds grp_pp_bounds(const double *__restrict__ U, const double *__restrict__ props,
                                         const Grid &g, int i, int j, int k, const double *ref, int pl) {
  constexpr int NB = D == 3 ? 27 : 9;
  const size_t nc = g.ncell;
  double maxDens = -1.7976931348623157e308, minDens = 1.7976931348623157e308;
  double minEint = 1.7976931348623157e308, minSdSp = 1.7976931348623157e308;
  for (int nb = pl; nb < NB; nb += LPG) {
    const int xo = D == 3 ? nb / 9 - 1 : nb / 3 - 1;
    const int yo = D == 3 ? (nb / 3) % 3 - 1 : nb % 3 - 1;
    const int zo = D == 3 ? nb % 3 - 1 : 0;
    const CellRef r = map_cell<D, 5>(g, i + xo, j + yo, k + zo);
    const double rho = U[r.idx];
    maxDens = fmax(maxDens, rho);
    minDens = fmin(minDens, rho);
    minEint = fmin(minEint, props[r.idx]);
    minSdSp = fmin(minSdSp, props[nc + r.idx]);
  }
  maxDens = grp_max<LPG>(maxDens);
  minDens = grp_min<LPG>(minDens);
  minEint = grp_min<LPG>(minEint);
  minSdSp = grp_min<LPG>(minSdSp);
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
  PPBounds pb;
  pb.rDens = ref[0];
  pb.maxDens = maxDens * (1.0 + kappa);
  pb.minDens = minDens * (1.0 - kappa);
  pb.minEint = minEint * (1.0 - kappa);
  return pb;
}

template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_pp_lanes(const double *__restrict__ U,
                                                  double *__restrict__ FS,
                                                  double *__restrict__ thetaNew,
                                                  const double *__restrict__ props,
                                                  const double *__restrict__ wflag, Grid g,
                                                  const int *__restrict__ list,
                                                  const int *__restrict__ count) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q, LPG = PPGroup<FQ>::LPG;
  constexpr int GPW = 64 / LPG;
  const int n = *count;
  const int lane = threadIdx.x & 63, pl = lane % LPG;
  const int gid = (blockIdx.x * (BLK / 64) + (threadIdx.x >> 6)) * GPW + lane / LPG;
  const int ngroups = gridDim.x * (BLK / 64) * GPW;
  const size_t nc = g.ncell;
  for (int t = gid; t < n; t += ngroups) {
    const int c = list[t];
    if (wflag[c] != 0.0) continue;  // uniform over the group
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const PPBounds pb = grp_pp_bounds<D, LPG>(U, props, g, i, j, k, ref, pl);
    // pp_apply with one lane per face point
    const bool own = pl < FQ;
    double st[5];
    double *p = FS + fs_base(c, FQ * 5) + (size_t(own ? pl : 0) * 5) * FSB;
#pragma unroll
    for (int w = 0; w < 5; w++) st[w] = own ? p[w * FSB] : ref[w];
    const double thD = grp_min<LPG>(own ? dens_theta(1.0, st[0], pb) : 1.0);
    const bool mixD = thD > 0.0 && thD < 1.0, repD = !mixD && thD < 1.0;
#pragma unroll
    for (int w = 0; w < 5; w++) st[w] = repD ? ref[w] : (mixD ? ref[w] + thD * (st[w] - ref[w]) : st[w]);
    double thE = 1.0;
    if (own && !(Hydro::eint(st) > pb.minEint)) thE = eint_pos(pb.minEint, ref, st);
    const double thP = grp_min<LPG>(thE);
    const bool mixP = thP > 0.0 && thP < 1.0, repP = !mixP && thP < 1.0;
#pragma unroll
    for (int w = 0; w < 5; w++) st[w] = repP ? ref[w] : (mixP ? ref[w] + thP * (st[w] - ref[w]) : st[w]);
    if (own && (thD < 1.0 || thP < 1.0)) {
#pragma unroll
      for (int w = 0; w < 5; w++) p[w * FSB] = st[w];
    }
    if (pl == 0) thetaNew[c] = fmin(thD, thP);
  }
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
        const int kk = D == 3 ? map_axis(k + zo, g.nz, g.bc[4], g.bc[5], sz) : k;
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

// ------------------------------------------------- post-WENO pass
// After WENO: back to conserved variables about the cell's own average
// (Hydro_K.H:53-64), then the positivity limiter (Common_K.H:299-643) in its
// exact sequence, a group of G lanes per listed cell, each lane with the
// face points pl, pl + G, ... (G = 8 for up to 24 points: 8 cells per wave in
// flight; 16 for NQ = 3 in 3D).  The min over points of each theta is a lane
// fold then a group reduction, which equals the sequential fold, and when no
// bound is violated both thetas stay 1 and the states pass unchanged, as
// after the reference's early exit.  States of list positions t < cap come
// from WS[t] (k_wenorec5) and stay there (the flux kernel reads them through
// the per-cell list positions); the other listed cells are limited in FS.
template <int FQ>
struct PostGroup {
  static constexpr int G = FQ <= 24 ? 8 : 16;
  static constexpr int PPL = (FQ + G - 1) / G;  // points per lane
};

template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_weno_post(const double *__restrict__ U,
                                                   double *__restrict__ FS,
                                                   double *__restrict__ thetaNew,
                                                   const double *__restrict__ props, Grid g,
                                                   const int *__restrict__ list,
                                                   const int *__restrict__ counter,
                                                   double *__restrict__ WS, int cap) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  constexpr int G = PostGroup<FQ>::G, PPL = PostGroup<FQ>::PPL, GPW = 64 / G;
  const int n = list ? *counter : g.ncell;
  const int lane = threadIdx.x & 63, pl = lane % G;
  const int gid = (blockIdx.x * (BLK / 64) + (threadIdx.x >> 6)) * GPW + lane / G;
  const int ngroups = gridDim.x * (BLK / 64) * GPW;
  const size_t nc = g.ncell;
  for (int t = gid; t < n; t += ngroups) {
    const int c = list ? list[t] : t;
    const bool compact = list && t < cap;
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const PPBounds pb = grp_pp_bounds<D, G>(U, props, g, i, j, k, ref, pl);
    const LinPrim lp(ref, g.gamma);
    const double ru = lp.rho * lp.u, rvv = lp.rho * lp.v, rw = lp.rho * lp.w;
    double st[PPL][5];
    double thDl = 1.0;
#pragma unroll
    for (int e = 0; e < PPL; e++) {
      const int pt = pl + G * e;
      if (pt < FQ) {
        double a[5];
        if (compact) {
          const double *q = WS + size_t(t) * (FQ * 5) + pt * 5;
#pragma unroll
          for (int w = 0; w < 5; w++) a[w] = q[w];
        } else {
          const double *q = FS + fs_base(c, FQ * 5) + (size_t(pt) * 5) * FSB;
#pragma unroll
          for (int w = 0; w < 5; w++) a[w] = q[w * FSB];
        }
        const double d = a[0], mx = a[1], my = a[2], mz = a[3], en = a[4];
        st[e][4] = (lp.vsq * d + ru * mx + rvv * my + rw * mz + en / lp.gamm);
        st[e][1] = (lp.rho * mx + lp.u * d);
        st[e][2] = (lp.rho * my + lp.v * d);
        st[e][3] = (lp.rho * mz + lp.w * d);
        st[e][0] = d;
        thDl = fmin(thDl, dens_theta(1.0, d, pb));
      } else {
#pragma unroll
        for (int w = 0; w < 5; w++) st[e][w] = ref[w];
      }
    }
    const double thD = grp_min<G>(thDl);
    const bool mixD = thD > 0.0 && thD < 1.0, repD = !mixD && thD < 1.0;
    double thEl = 1.0;
#pragma unroll
    for (int e = 0; e < PPL; e++) {
#pragma unroll
      for (int w = 0; w < 5; w++)
        st[e][w] = repD ? ref[w] : (mixD ? ref[w] + thD * (st[e][w] - ref[w]) : st[e][w]);
      if (pl + G * e < FQ && !(Hydro::eint(st[e]) > pb.minEint))
        thEl = fmin(thEl, eint_pos(pb.minEint, ref, st[e]));
    }
    const double thP = grp_min<G>(thEl);
    const bool mixP = thP > 0.0 && thP < 1.0, repP = !mixP && thP < 1.0;
#pragma unroll
    for (int e = 0; e < PPL; e++) {
      const int pt = pl + G * e;
      if (pt < FQ) {
        double *const q = compact ? WS + size_t(t) * (FQ * 5) + pt * 5 : FS + fs_base(c, FQ * 5) + (size_t(pt) * 5) * FSB;
        const size_t qs = compact ? 1 : FSB;
#pragma unroll
        for (int w = 0; w < 5; w++)
          q[w * qs] = repP ? ref[w] : (mixP ? ref[w] + thP * (st[e][w] - ref[w]) : st[e][w]);
      }
    }
    if (pl == 0) thetaNew[c] = fmin(thD, thP);
  }
}

// The same pass over every cell (full WENO, no list): one thread per cell,
// so that each access of a state plane covers consecutive cells.
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_weno_post_cells(const double *__restrict__ U,
                                                         double *__restrict__ FS,
                                                         double *__restrict__ thetaNew,
                                                         const double *__restrict__ props, Grid g) {
  constexpr int FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  const size_t nc = g.ncell;
  for (int c = blockIdx.x * BLK + threadIdx.x; c < g.ncell; c += gridDim.x * BLK) {
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
      double *p = FS + fs_base(c, FQ * 5) + (size_t(fq) * 5) * FSB;
      const double d = p[0], mx = p[FSB], my = p[2 * FSB], mz = p[3 * FSB], e = p[4 * FSB];
      double st[5];
      st[4] = (lp.vsq * d + ru * mx + rvv * my + rw * mz + e / lp.gamm);
      st[1] = (lp.rho * mx + lp.u * d);
      st[2] = (lp.rho * my + lp.v * d);
      st[3] = (lp.rho * mz + lp.w * d);
      st[0] = d;
#pragma unroll
      for (int w = 1; w < 5; w++) p[w * FSB] = st[w];
      thD = dens_theta(thD, d, pb);
      okE = okE && Hydro::eint(st) > pb.minEint;
    }
    // no bound violated: the states stay as they are (theta 1)
    double th = 1.0;
    if (!(thD == 1.0 && okE)) th = pp_apply<D, NQ>(FS + fs_base(c, FQ * 5), FSB, ref, pb);
    thetaNew[c] = th;
  }
}

}  // namespace dev
}  // namespace kfvm
