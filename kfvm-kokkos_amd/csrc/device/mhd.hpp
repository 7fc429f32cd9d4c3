// Stage kernels of the GLM-MHD equations (9 variables: rho, mx, my, mz, E,
// Bx, By, Bz, psi).  The face states (FS) and the interior quadrature-point
// states (SRC, needed by the GLM source terms) are planes
// [(point * 9 + var)][cell] (FS in blocks of FSB cells).
//
//   k_mhd_linear     KernelLinearRecon_K incl. the source branch
//                    (Stencil_K.H:32-118) on a staged brick, with the KXRCF
//                    input p/rho^gamma of every face state (Numeric_K.H:315-379),
//                    the limiter's fast test and the cleaning-speed maxima
//   k_mhd_wenorec    KernelWenoRecon_K + fillInterior (Stencil_K.H:164-504),
//                    one lane per (cell, variable)
//   k_mhd_weno_post  VarConvert<MHD_GLM> back-conversion (MHD_GLM_K.H:24-69),
//                    then the limiter of the WENO cells: lanes per listed
//                    cell (k_mhd_weno_post_cells: thread per cell, full WENO)
//   k_mhd_pp         PositivityPreserve_K with interior points
//                    (Common_K.H:299-643) for the cells that failed the fast
//                    test, then SpeedEstimate_K (MHD_GLM_K.H:213-310) on the
//                    limited face states
//   mhd_source_rhs   SourceTerms_K (MHD_GLM_K.H:316-547): div B and psi-gradient
//                    terms at the interior points, integrated inside the flux
//                    kernel (k_flux_rk)
//   k_mhd_chmax      the cleaning speed c_h = max(c_f + |u|) - max(|u|) over
//                    all face states, x-components on every face
#pragma once

namespace kfvm {
namespace dev {

template <int D, int NQ>
struct MShape {
  static constexpr int F = 2 * D, Q = D == 3 ? NQ * NQ : NQ, FQ = F * Q;
  static constexpr int NI = D == 3 ? NQ * NQ * NQ : NQ * NQ;  // interior points
  static constexpr int NZ = D == 3 ? NQ : 1;
};

// linearised primitive context about the centre cell (VarConvert<MHD_GLM>)
struct MhdLin {
  double gamm, rho, u, v, w, bx, by, bz, rPsi, vsq;
  __device__ inline MhdLin(const double *r, double gamma)
      : gamm(gamma - 1.0), rho(r[0]), u(r[1] / r[0]), v(r[2] / r[0]), w(r[3] / r[0]), bx(r[5]),
        by(r[6]), bz(r[7]), rPsi(r[8]), vsq(0.0) {
    vsq = (u * u + v * v + w * w) / 2.0;
  }
};

// ------------------------------------------------------------ linear pass
// Brick of the linear pass: the nine variables of 16x8x4 cells (3D) and their
// radius-2 halo in LDS (138 KB: one workgroup of 8 waves per CU).
template <int D>
struct MBox {
  static constexpr int H = 2;
  static constexpr int BX = 16, BY = D == 3 ? 8 : 16, BZ = D == 3 ? 4 : 1;
  static constexpr int LX = BX + 2 * H, LY = BY + 2 * H, LZ = D == 3 ? BZ + 2 * H : 1;
  static constexpr int N = LX * LY * LZ;   // staged cells per variable
  static constexpr int NT = BX * BY * BZ;  // threads per block
};

// PP points of all nine variables from the staged brick.  The stencil walk is
// outermost, so each accumulator sums n = 0..C-1 in order (the reference's
// dot product), and the PP weights of one stencil cell -- one scalar load --
// feed 9 PP FMAs.  The next stencil cell's weights and LDS values are fetched
// before this cell's FMAs issue; their addresses pass through an empty asm
// that takes this cell's operands as inputs, so the wait for this cell's
// loads is not merged with the next cell's (one counter serves scalar and LDS
// loads, and a wait with a scalar load outstanding is a full one).
template <int D, class Bx, int PP>
__device__ __forceinline__ void mhd_pass(const double *__restrict__ lds, int corner, const double *wbase,
                                         double (&acc)[9][PP]) {
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, H = Bx::H, NP = Bx::N;
#pragma unroll
  for (int v = 0; v < 9; v++)
#pragma unroll
    for (int q = 0; q < PP; q++) acc[v][q] = 0.0;
  double w[PP], x[9];
  {
    const double *w0 = wbase;
    asm volatile("" : "+s"(w0));
    const cdouble *wc = (const cdouble *)w0;
#pragma unroll
    for (int q = 0; q < PP; q++) w[q] = wc[q];
    constexpr int o = ST.off[0][0] + H + Bx::LX * (ST.off[0][1] + H) +
                      (D == 3 ? Bx::LX * Bx::LY * (ST.off[0][2] + H) : 0);
#pragma unroll
    for (int v = 0; v < 9; v++) x[v] = lds[v * NP + corner + o];
  }
#pragma unroll
  for (int n = 0; n < C; n++) {
    double wn[PP], xn[9];
    if (n + 1 < C) {
      const double *w1 = wbase + (n + 1) * PP;
      int cn = corner;
      asm volatile("" : "+s"(w1), "+v"(cn) : "s"(w[0]), "s"(w[PP - 1]), "v"(x[0]), "v"(x[1]), "v"(x[2]),
                   "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]));
      const cdouble *wc = (const cdouble *)w1;
#pragma unroll
      for (int q = 0; q < PP; q++) wn[q] = wc[q];
      const int o = (ST.off[n + 1][0] + H) + Bx::LX * (ST.off[n + 1][1] + H) +
                    (D == 3 ? Bx::LX * Bx::LY * (ST.off[n + 1][2] + H) : 0);
#pragma unroll
      for (int v = 0; v < 9; v++) xn[v] = lds[v * NP + cn + o];
    }
#pragma unroll
    for (int q = 0; q < PP; q++)
#pragma unroll
      for (int v = 0; v < 9; v++) acc[v][q] = fma(w[q], x[v], acc[v][q]);
    if (n + 1 < C) {
      __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 9 * PP, 0);
#pragma unroll
      for (int q = 0; q < PP; q++) w[q] = wn[q];
#pragma unroll
      for (int v = 0; v < 9; v++) x[v] = xn[v];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

// Fastest magnetosonic speed + |u| and |u| of one face state, x-components on
// every face (SpeedEstimate_K, MHD_GLM_K.H:213-310).
__device__ inline void mhd_face_speed(double gamma, const double *s, double &maxLam, double &maxVel) {
  const double vel = fabs(s[1] / s[0]);
  const double lam = Mhd::fast(gamma, s, s[5]) + vel;
  maxVel = vel > maxVel ? vel : maxVel;
  maxLam = lam > maxLam ? lam : maxLam;
}

// Linear face and interior states (KernelLinearRecon_K incl. the source
// branch, Stencil_K.H:32-118), one thread per cell of a staged brick, in
// passes of P face points / PI interior points of all nine variables.  While
// a pass' states are in registers the kernel also forms
//   - the KXRCF indicator p/rho^gamma of each face state (WI, when given;
//     Numeric_K.H:315-379),
//   - the limiter's fast test: every density inside the widened bounds and
//     every internal energy above the bound (Common_K.H:299-643) -- a cell
//     that passes keeps its states and theta = 1, one that fails goes to
//     ppList for the exact sequence (k_mhd_pp),
//   - the cleaning-speed maxima of the face states (SpeedEstimate_K) into the
//     per-cell plane spd [maxLam | maxVel]; cells that the limiter or WENO
//     change later overwrite theirs.
// The learned selection stays a kernel of its own here (k_learned_select):
// run on this 138 KB brick, which leaves 8 waves per CU, the network's
// matrix-core phase cost more than the separate kernel's re-staging of
// density and pressure.
template <int D, int NQ>
__global__ __launch_bounds__(MBox<D>::NT, 2) void k_mhd_linear(
    const double *__restrict__ U, double *__restrict__ FS, double *__restrict__ SRC, double *__restrict__ WI,
    double *__restrict__ thetaNew, double *__restrict__ spd, const double *__restrict__ props, Grid g,
    const double *__restrict__ WT, int *__restrict__ ppList, int *__restrict__ ppCount) {
  using Bx = MBox<D>;
  constexpr int C = SInfo<D>::C, H = Bx::H, NP = Bx::N;
  constexpr int FQ = MShape<D, NQ>::FQ, NI = MShape<D, NQ>::NI;
  constexpr int P = linear_pass_points(D, NQ), PI = mhd_interior_pass_points(D, NQ);
  static_assert(FQ % P == 0 && NI % PI == 0, "passes must tile the points");
  __shared__ double lds[9 * NP];
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
      const CellRef cr = map_cell<D, 9>(g, gi, gj, gk);
#pragma unroll
      for (int w = 0; w < 9; w++) lds[w * NP + idx] = ld(U, nc, cr, w);
    }
  }
  __syncthreads();
  const int tx = tid % Bx::BX, ty = (tid / Bx::BX) % Bx::BY, tz = tid / (Bx::BX * Bx::BY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  if (i < g.nx && j < g.ny && (D == 2 || k < g.nz)) {  // no barrier inside
  const int cell = (k * g.ny + j) * g.nx + i;
  const int corner = (tz * Bx::LY + ty) * Bx::LX + tx;  // stencil box corner (-H,-H,-H)
  const int c0 = corner + H + Bx::LX * H + (D == 3 ? Bx::LX * Bx::LY * H : 0);
  double ref[9];
#pragma unroll
  for (int w = 0; w < 9; w++) ref[w] = lds[w * NP + c0];
  const PPBounds pb = pp_bounds_lds<D, Bx>(lds, c0, props, g, i, j, k, ref);
  double thD = 1.0, maxLam = 0.0, maxVel = 0.0;
  bool okE = true;
  const size_t fb = fs_base(cell, FQ * 9);
#pragma unroll 1
  for (int ps = 0; ps < FQ / P; ps++) {
    double acc[9][P];
    mhd_pass<D, Bx, P>(lds, corner, WT + size_t(ps) * C * P, acc);
#pragma unroll
    for (int q = 0; q < P; q++)
#pragma unroll
      for (int v = 0; v < 9; v++) FS[fb + (size_t(ps * P + q) * 9 + v) * FSB] = acc[v][q];
#pragma unroll
    for (int q = 0; q < P; q++) {
      double s[9];
#pragma unroll
      for (int v = 0; v < 9; v++) s[v] = acc[v][q];
      thD = dens_theta(thD, s[0], pb);
      okE = okE && Mhd::eint(s) > pb.minEint;
      mhd_face_speed(g.gamma, s, maxLam, maxVel);
      if (WI) {
        WI[fs_base(cell, FQ) + size_t(ps * P + q) * FSB] = Mhd::weno(g.gamma, s);
        __builtin_amdgcn_sched_barrier(0);  // one pow at a time: bounded registers
      }
    }
  }
  const double *wi = WT + size_t(FQ) * C;
#pragma unroll 1
  for (int ps = 0; ps < NI / PI; ps++) {
    double acc[9][PI];
    mhd_pass<D, Bx, PI>(lds, corner, wi + size_t(ps) * C * PI, acc);
#pragma unroll
    for (int q = 0; q < PI; q++)
#pragma unroll
      for (int v = 0; v < 9; v++) SRC[(size_t(ps * PI + q) * 9 + v) * nc + cell] = acc[v][q];
#pragma unroll
    for (int q = 0; q < PI; q++) {
      double s[9];
#pragma unroll
      for (int v = 0; v < 9; v++) s[v] = acc[v][q];
      thD = dens_theta(thD, s[0], pb);
      okE = okE && Mhd::eint(s) > pb.minEint;
    }
  }
  spd[cell] = maxLam;
  spd[nc + cell] = maxVel;
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
}

// KFVM-WENO of the listed cells (list == nullptr: all), one lane per (cell,
// variable): 7 cells of 9 lanes per wave.  The lane gathers its variable's
// linearised-primitive stencil values, forms the 8 smoothness indicators and
// the WENO-AO weights, and writes the weighted face and interior values, still
// linearised primitive (k_mhd_weno_post converts them back).  The products
// walk the stencil in their outer loop with one accumulator per output (the
// Euler kernel's pieces: WenoSubs over the compact table WC in LDS, or the
// stencil-major tables FT [nSub][C][F*Q], DT [nSub][C][I], CT [nSub][C][NQ^D]
// by scalar loads; see MWenoLds), so each output still sums the stencil cells
// in ascending order.  The grid is persistent when the table is in LDS (at
// most 2048 workgroups striding over the list), so the staging is paid once
// per workgroup.
constexpr int MLPC = 9;                  // lanes per cell
constexpr int MCPW = 7;                  // cells per wave
constexpr int MCPB = MCPW * (BLK / 64);  // cells per block

template <int D, int NQ>
struct MhdChunk {  // interior points per accumulation chunk
  static constexpr int NI = MShape<D, NQ>::NI;
  static constexpr int H = NI > 12 ? 9 : NI;
};

// The compact weight table of the GLM-MHD WENO kernel: the Euler kernel's
// face rows [SUB][F*Q] and derivative rows [SUB][I], then interior-point rows
// [SUB][NQ^D] (compact_weno_weights).  Staged in LDS once per workgroup
// where it fits in 64 KB (the scalar cache cannot hold it: 43 KB in 3D at
// NQ = 2); the 3D NQ = 3 table (85 KB) stays with the stencil-major tables
// read by scalar loads.  Like the Euler kernel it runs on LDS latency: three
// waves per SIMD (168 VGPRs) beat the two that 256 VGPRs allow by 10 %.
template <int D, int NQ>
struct MWenoLds {
  static constexpr int SUB = WenoLds<D, NQ>::SUB, FQ = MShape<D, NQ>::FQ, I = SInfo<D>::I;
  static constexpr int NI = MShape<D, NQ>::NI;
  static constexpr int NDER = SUB * FQ, NCELL = SUB * (FQ + I), NW = SUB * (FQ + I + NI);
  static constexpr bool USE = NW * int(sizeof(double)) <= 65536;
};

template <int D, int NQ>
__global__ __launch_bounds__(BLK, 3) void k_mhd_wenorec(const double *__restrict__ U,
                                                     double *__restrict__ FS,
                                                     double *__restrict__ SRC, Grid g,
                                                     const double *__restrict__ FT,
                                                     const double *__restrict__ DT,
                                                     const double *__restrict__ CT,
                                                     const int *__restrict__ list,
                                                     const int *__restrict__ counter,
                                                     const double *__restrict__ WC) {
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, nSub = SInfo<D>::nSub;
  constexpr int FQ = MShape<D, NQ>::FQ, NI = MShape<D, NQ>::NI;
  constexpr int HF = WenoChunk<D, NQ>::H, HI = MhdChunk<D, NQ>::H;
  static_assert(FQ % HF == 0 && NI % HI == 0, "chunks must tile the points");
  constexpr double gHi = 0.7, gLo = 0.7;
  constexpr double g0 = gHi, g1 = (1.0 - gHi) * gLo, gb = (1.0 - gHi) * (1.0 - gLo) / double(2 * D);
  using WL = MWenoLds<D, NQ>;
  __shared__ double wl[WL::USE ? WL::NW : 1];
  if constexpr (WL::USE) {
    static_assert(WL::SUB == SubOff<D, nSub>::value, "compact rows of all substencils");
    for (int e = threadIdx.x; e < WL::NW; e += BLK) wl[e] = WC[e];
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int slot = lane / MLPC;
  if (slot >= MCPW) return;  // lane 63 idles; no barrier follows
  const int v = lane - MLPC * slot;
  const int n = list ? *counter : g.ncell;
  const int waveId = blockIdx.x * (BLK / 64) + (threadIdx.x >> 6);
  const int nWaves = gridDim.x * (BLK / 64);
  const size_t nc = g.ncell;
#pragma unroll 1
  for (int t = waveId * MCPW + slot; t < n; t += nWaves * MCPW) {
    const int c = list ? list[t] : t;
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[9];
#pragma unroll
    for (int w = 0; w < 9; w++) ref[w] = U[size_t(w) * nc + c];
    const MhdLin lp(ref, g.gamma);
    const double uv = v == 1 ? lp.u : (v == 2 ? lp.v : lp.w);
    // stencil values of variable v, linearised primitive (cons2LinPrim): each
    // lane loads its own variable, the others come from the cell's lanes
    // (lane 9 slot + w holds variable w)
    const int gbase = MLPC * slot;
    double vs[C];
#pragma unroll
    for (int m = 0; m < C; m++) {
      const CellRef r = map_cell<D, 9>(g, i + ST.off[m][0], j + ST.off[m][1], k + ST.off[m][2]);
      const double sv = ld(U, nc, r, v);
      double q[9];
#pragma unroll
      for (int w = 0; w < 9; w++) q[w] = __shfl(sv, gbase + w);
      if (v == 4) {
        vs[m] = lp.gamm * (lp.vsq * q[0] + q[4] - lp.u * q[1] - lp.v * q[2] - lp.w * q[3] - lp.bx * q[5] -
                           lp.by * q[6] - lp.bz * q[7] - lp.rPsi * q[8]);
      } else if (v >= 1 && v <= 3) {
        vs[m] = (sv - uv * q[0]) / lp.rho;
      } else {
        vs[m] = sv;
      }
    }
    // smoothness indicators and WENO-AO weights
    double om[nSub];
    const double *ft0 = FT, *ct0 = CT;
    if constexpr (WL::USE) {
      WenoSubs<D, 0, nSub, true>::beta(vs, (const double *)(wl + WL::NDER), om);
    } else {
      // weight tables opaque to loop-invariant code motion (see k_linear_fused)
      const double *dt0 = DT;
      asm volatile("" : "+s"(dt0), "+s"(ft0), "+s"(ct0));
      WenoSubs<D, 0, nSub>::beta(vs, (const cdouble *)dt0, om);
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
    // face points, then interior points, a chunk at a time
#pragma unroll 1
    for (int h0 = 0; h0 < FQ; h0 += HF) {
      double out[HF];
#pragma unroll
      for (int h = 0; h < HF; h++) out[h] = 0.0;
      if constexpr (WL::USE) {
        WenoSubs<D, 0, nSub, true>::template faces<HF, FQ>(vs, (const double *)(wl + h0), om, out);
      } else {
        const double *fc0 = ft0 + h0;
        asm volatile("" : "+s"(fc0));
        WenoSubs<D, 0, nSub>::template faces<HF, FQ>(vs, (const cdouble *)fc0, om, out);
      }
#pragma unroll
      for (int h = 0; h < HF; h++) FS[fs_base(c, FQ * 9) + (size_t(h0 + h) * 9 + v) * FSB] = out[h];
    }
#pragma unroll 1
    for (int p0 = 0; p0 < NI; p0 += HI) {
      double out[HI];
#pragma unroll
      for (int h = 0; h < HI; h++) out[h] = 0.0;
      if constexpr (WL::USE) {
        WenoSubs<D, 0, nSub, true>::template faces<HI, NI>(vs, (const double *)(wl + WL::NCELL + p0), om, out);
      } else {
        const double *cc0 = ct0 + p0;
        asm volatile("" : "+s"(cc0));
        WenoSubs<D, 0, nSub>::template faces<HI, NI>(vs, (const cdouble *)cc0, om, out);
      }
#pragma unroll
      for (int h = 0; h < HI; h++) SRC[(size_t(p0 + h) * 9 + v) * nc + c] = out[h];
    }
  }
}

// ------------------------------------------------------------------ limiter
__device__ inline double mhd_eint_pos(double emin, const double *ref, const double *s) {
  double a = 0.0, b = 1.0, c;
  double st[9];
  for (int it = 0; it < 10; it++) {
    c = 0.5 * (a + b);
#pragma unroll
    for (int v = 0; v < 9; v++) st[v] = ref[v] + c * (s[v] - ref[v]);
    const double dec = Mhd::eint(st) - emin;
    if (dec > 0.0) a = c;
    else b = c;
  }
  return a;
}

template <int D, int NQ>
__device__ inline void mhd_scale(double *__restrict__ FS, double *__restrict__ SRC, size_t nc,
                                 int c, const double *ref, double th, bool replace) {
  constexpr int FQ = MShape<D, NQ>::FQ, NI = MShape<D, NQ>::NI;
#pragma unroll 1
  for (int p = 0; p < FQ + NI; p++) {
    double *b = p < FQ ? FS + fs_base(c, FQ * 9) + size_t(p) * 9 * FSB : SRC + size_t(p - FQ) * 9 * nc + c;
    const size_t bs = p < FQ ? size_t(FSB) : nc;
#pragma unroll
    for (int v = 0; v < 9; v++) b[v * bs] = replace ? ref[v] : ref[v] + th * (b[v * bs] - ref[v]);
  }
}

// Per-cell internal energy and sound speed (MHD_GLM_EOS.H), the two
// neighbourhood quantities of the limiter bounds; both are even in the
// momenta and the field, so mirrored wall cells need no sign.
template <int D>
__global__ __launch_bounds__(BLK) void k_mhd_props(const double *__restrict__ U, double *__restrict__ props,
                                                   Grid g) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  const size_t nc = g.ncell;
  double s[9];
#pragma unroll
  for (int v = 0; v < 9; v++) s[v] = U[size_t(v) * nc + c];
  props[c] = Mhd::eint(s);
  props[nc + c] = Mhd::sound(g.gamma, s);
}

// The limiter's exact sequence on one cell (PositivityPreserve_K with the
// interior points, Common_K.H:299-643), then the cleaning-speed maxima of its
// limited face states into spd: the energy pass already holds each face
// state, so the speeds come from those registers unless that pass rescales
// the states again.
template <int D, int NQ>
__device__ inline void mhd_pp_cell(const double *__restrict__ U, double *__restrict__ FS, double *__restrict__ SRC,
                                   double *__restrict__ thetaNew, double *__restrict__ spd,
                                   const double *__restrict__ props, const Grid &g, int c) {
  constexpr int FQ = MShape<D, NQ>::FQ, NI = MShape<D, NQ>::NI;
  int i, j, k;
  cell_ijk<D>(c, g, i, j, k);
  const size_t nc = g.ncell;
  double ref[9];
#pragma unroll
  for (int v = 0; v < 9; v++) ref[v] = U[size_t(v) * nc + c];
  // ranges over the 3^D neighbourhood (Common_K.H:329-387)
  double maxDens = -1.7976931348623157e308, minDens = 1.7976931348623157e308;
  double minEint = 1.7976931348623157e308, minSdSp = 1.7976931348623157e308;
  constexpr int zl = D == 3 ? -1 : 0, zh = D == 3 ? 1 : 0;
#pragma unroll 1
  for (int xo = -1; xo <= 1; xo++)
#pragma unroll 1
    for (int yo = -1; yo <= 1; yo++)
#pragma unroll 1
      for (int zo = zl; zo <= zh; zo++) {
        const size_t q = map_cell<D, 9>(g, i + xo, j + yo, k + zo).idx;
        const double rho = U[q];
        maxDens = fmax(maxDens, rho);
        minDens = fmin(minDens, rho);
        minEint = fmin(minEint, props[q]);
        minSdSp = fmin(minSdSp, props[nc + q]);
      }
  const double dxU = load_u<D, 9>(U, g, i + 1, j, k, 1) / load_u<D, 9>(U, g, i + 1, j, k, 0) -
                     load_u<D, 9>(U, g, i - 1, j, k, 1) / load_u<D, 9>(U, g, i - 1, j, k, 0);
  const double dyV = load_u<D, 9>(U, g, i, j + 1, k, 2) / load_u<D, 9>(U, g, i, j + 1, k, 0) -
                     load_u<D, 9>(U, g, i, j - 1, k, 2) / load_u<D, 9>(U, g, i, j - 1, k, 0);
  double dzW;
  if (D == 3)
    dzW = load_u<D, 9>(U, g, i, j, k + 1, 3) / load_u<D, 9>(U, g, i, j, k + 1, 0) -
          load_u<D, 9>(U, g, i, j, k - 1, 3) / load_u<D, 9>(U, g, i, j, k - 1, 0);
  else
    dzW = ref[3] / ref[0] - ref[3] / ref[0];
  const double divV = (dxU + dyV + dzW) / 2.0;
  double kappa = 0.2;
  const double eta = fmin(1.0, fmax(0.0, -(kappa * minSdSp + divV) / (kappa * minSdSp)));
  kappa *= (1.0 - eta);
  maxDens *= (1.0 + kappa);
  minDens *= (1.0 - kappa);
  minEint *= (1.0 - kappa);
  PPBounds b;
  b.rDens = ref[0];
  b.maxDens = maxDens;
  b.minDens = minDens;
  b.minEint = minEint;
  // density: faces then interior points
  double thD = 1.0;
#pragma unroll 1
  for (int p = 0; p < FQ + NI; p++) {
    const double rho = p < FQ ? FS[fs_base(c, FQ * 9) + size_t(p) * 9 * FSB] : SRC[size_t(p - FQ) * 9 * nc + c];
    thD = dens_theta(thD, rho, b);
  }
  if (thD > 0.0 && thD < 1.0) mhd_scale<D, NQ>(FS, SRC, nc, c, ref, thD, false);
  else if (thD < 1.0) mhd_scale<D, NQ>(FS, SRC, nc, c, ref, thD, true);
  // internal energy (and the speeds of the face states as they stand)
  double thP = 1.0, maxLam = 0.0, maxVel = 0.0;
#pragma unroll 1
  for (int p = 0; p < FQ + NI; p++) {
    const double *q = p < FQ ? FS + fs_base(c, FQ * 9) + size_t(p) * 9 * FSB : SRC + size_t(p - FQ) * 9 * nc + c;
    const size_t qs = p < FQ ? size_t(FSB) : nc;
    double s[9];
#pragma unroll
    for (int v = 0; v < 9; v++) s[v] = q[v * qs];
    const double ei = Mhd::eint(s);
    thP = ei > minEint ? thP : fmin(thP, mhd_eint_pos(minEint, ref, s));
    if (p < FQ) mhd_face_speed(g.gamma, s, maxLam, maxVel);
  }
  if (thP < 1.0) {
    mhd_scale<D, NQ>(FS, SRC, nc, c, ref, thP, !(thP > 0.0));
    maxLam = 0.0;
    maxVel = 0.0;
#pragma unroll 1
    for (int fq = 0; fq < FQ; fq++) {
      double s[9];
#pragma unroll
      for (int v = 0; v < 9; v++) s[v] = FS[fs_base(c, FQ * 9) + (size_t(fq) * 9 + v) * FSB];
      mhd_face_speed(g.gamma, s, maxLam, maxVel);
    }
  }
  thetaNew[c] = fmin(thD, thP);
  spd[c] = maxLam;
  spd[nc + c] = maxVel;
}

// Back to conserved variables (linPrim2Cons, MHD_GLM_K.H:24-69) at every face
// and interior point of the listed cells, then the limiter's exact sequence
// on the cell, with one lane per point: a group
// of LPG lanes per cell (PPL points per lane where the cell has more points
// than 64).  The energy sums its terms in the reference's operand order
//   E = vsq d + rho u mx + rho v my + rho w mz + bx Bx + by By + bz Bz
//       + e / (gamma - 1) + psi0 psi.
// The limiter's folds over the points (the density and energy theta, the
// speed maxima) become group reductions: min and max do not depend on the
// order (the NaN-ignoring fmin / fmax and compares of the sequential form),
// so the results equal mhd_pp_cell's.  The limiter bounds come from the
// group's lanes, one neighbour each.
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_mhd_weno_post(const double *__restrict__ U,
                                                       double *__restrict__ FS,
                                                       double *__restrict__ SRC, Grid g,
                                                       const int *__restrict__ list,
                                                       const int *__restrict__ counter,
                                                       double *__restrict__ thetaNew, double *__restrict__ spd,
                                                       const double *__restrict__ props) {
  constexpr int FQ = MShape<D, NQ>::FQ, NI = MShape<D, NQ>::NI, NPT = FQ + NI;
  constexpr int LPG = PPGroup<NPT>::LPG, GPW = 64 / LPG, PPL = (NPT + LPG - 1) / LPG;
  const int n = *counter;
  const size_t nc = g.ncell;
  const int lane = threadIdx.x & 63, pl = lane % LPG;
  const int gid = (blockIdx.x * (BLK / 64) + (threadIdx.x >> 6)) * GPW + lane / LPG;
  const int ngroups = gridDim.x * (BLK / 64) * GPW;
  for (int t = gid; t < n; t += ngroups) {  // uniform over the group
    const int c = list[t];
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[9];
#pragma unroll
    for (int w = 0; w < 9; w++) ref[w] = U[size_t(w) * nc + c];
    const MhdLin lp(ref, g.gamma);
    // this lane's points: pl, pl + LPG, ...
    double st[PPL][9];
    bool own[PPL];
#pragma unroll
    for (int r = 0; r < PPL; r++) {
      const int p = pl + r * LPG;
      own[r] = p < NPT;
      const double *P = p < FQ ? FS + fs_base(c, FQ * 9) + size_t(p) * 9 * FSB
                               : SRC + size_t(own[r] ? p - FQ : 0) * 9 * nc + c;
      const size_t ps = p < FQ ? size_t(FSB) : nc;
      double a[9];
#pragma unroll
      for (int w = 0; w < 9; w++) a[w] = own[r] ? P[w * ps] : ref[w];
      const double d = a[0];
      double e = lp.vsq * d;
      const double mx = lp.rho * a[1] + lp.u * d;
      e = e + lp.rho * lp.u * a[1];
      const double my = lp.rho * a[2] + lp.v * d;
      e = e + lp.rho * lp.v * a[2];
      const double mz = lp.rho * a[3] + lp.w * d;
      e = e + lp.rho * lp.w * a[3];
      e = e + lp.bx * a[5];
      e = e + lp.by * a[6];
      e = e + lp.bz * a[7];
      e = e + a[4] / lp.gamm;
      e = e + lp.rPsi * a[8];
#pragma unroll
      for (int w = 0; w < 9; w++) st[r][w] = a[w];
      st[r][1] = own[r] ? mx : ref[1];
      st[r][2] = own[r] ? my : ref[2];
      st[r][3] = own[r] ? mz : ref[3];
      st[r][4] = own[r] ? e : ref[4];
    }
    // limiter bounds (Common_K.H:329-387), then density, then internal energy
    const PPBounds pb = grp_pp_bounds<D, LPG>(U, props, g, i, j, k, ref, pl);
    double thD = 1.0;
#pragma unroll
    for (int r = 0; r < PPL; r++) thD = own[r] ? dens_theta(thD, st[r][0], pb) : thD;
    thD = grp_min<LPG>(thD);
    const bool mixD = thD > 0.0 && thD < 1.0, repD = !mixD && thD < 1.0;
    double thE = 1.0, maxLam = 0.0, maxVel = 0.0;
#pragma unroll
    for (int r = 0; r < PPL; r++) {
#pragma unroll
      for (int w = 0; w < 9; w++)
        st[r][w] = repD ? ref[w] : (mixD ? ref[w] + thD * (st[r][w] - ref[w]) : st[r][w]);
      if (own[r] && !(Mhd::eint(st[r]) > pb.minEint)) thE = fmin(thE, mhd_eint_pos(pb.minEint, ref, st[r]));
    }
    const double thP = grp_min<LPG>(thE);
    const bool mixP = thP > 0.0 && thP < 1.0, repP = !mixP && thP < 1.0;
#pragma unroll
    for (int r = 0; r < PPL; r++) {
#pragma unroll
      for (int w = 0; w < 9; w++)
        st[r][w] = repP ? ref[w] : (mixP ? ref[w] + thP * (st[r][w] - ref[w]) : st[r][w]);
      if (own[r] && pl + r * LPG < FQ) mhd_face_speed(g.gamma, st[r], maxLam, maxVel);
    }
    maxLam = grp_max<LPG>(maxLam);
    maxVel = grp_max<LPG>(maxVel);
    // the converted momenta and energy, or every variable of a limited cell
    const bool all = thD < 1.0 || thP < 1.0;
#pragma unroll
    for (int r = 0; r < PPL; r++) {
      const int p = pl + r * LPG;
      if (!own[r]) continue;
      double *P = p < FQ ? FS + fs_base(c, FQ * 9) + size_t(p) * 9 * FSB : SRC + size_t(p - FQ) * 9 * nc + c;
      const size_t ps = p < FQ ? size_t(FSB) : nc;
#pragma unroll
      for (int w = 0; w < 9; w++)
        if (all || (w >= 1 && w <= 4)) P[w * ps] = st[r][w];
    }
    if (pl == 0) {
      thetaNew[c] = fmin(thD, thP);
      spd[c] = maxLam;
      spd[nc + c] = maxVel;
    }
  }
}

// k_mhd_weno_post for every cell (full WENO), one thread per cell: the same
// conversion and mhd_pp_cell's sequence; consecutive threads read consecutive
// cells' values, which a list does not give.
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_mhd_weno_post_cells(const double *__restrict__ U,
                                                             double *__restrict__ FS,
                                                             double *__restrict__ SRC, Grid g,
                                                             double *__restrict__ thetaNew, double *__restrict__ spd,
                                                             const double *__restrict__ props) {
  constexpr int FQ = MShape<D, NQ>::FQ, NI = MShape<D, NQ>::NI;
  const int n = g.ncell;
  const size_t nc = g.ncell;
  for (int t = blockIdx.x * BLK + threadIdx.x; t < n; t += gridDim.x * BLK) {
    const int c = t;
    double ref[9];
#pragma unroll
    for (int w = 0; w < 9; w++) ref[w] = U[size_t(w) * nc + c];
    const MhdLin lp(ref, g.gamma);
#pragma unroll 2
    for (int p = 0; p < FQ + NI; p++) {
      double *P = p < FQ ? FS + fs_base(c, FQ * 9) + size_t(p) * 9 * FSB : SRC + size_t(p - FQ) * 9 * nc + c;
      const size_t ps = p < FQ ? size_t(FSB) : nc;
      double a[9];
#pragma unroll
      for (int w = 0; w < 9; w++) a[w] = P[w * ps];
      const double d = a[0];
      double e = lp.vsq * d;
      const double mx = lp.rho * a[1] + lp.u * d;
      e = e + lp.rho * lp.u * a[1];
      const double my = lp.rho * a[2] + lp.v * d;
      e = e + lp.rho * lp.v * a[2];
      const double mz = lp.rho * a[3] + lp.w * d;
      e = e + lp.rho * lp.w * a[3];
      e = e + lp.bx * a[5];
      e = e + lp.by * a[6];
      e = e + lp.bz * a[7];
      e = e + a[4] / lp.gamm;
      e = e + lp.rPsi * a[8];
      P[ps] = mx;
      P[2 * ps] = my;
      P[3 * ps] = mz;
      P[4 * ps] = e;
    }
    mhd_pp_cell<D, NQ>(U, FS, SRC, thetaNew, spd, props, g, c);
  }
}

// The limiter's exact sequence for the cells of the linear pass that failed
// its fast test (list, *count of them); WENO cells are skipped, their
// post-WENO pass limits them.
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_mhd_pp(const double *__restrict__ U, double *__restrict__ FS,
                                                double *__restrict__ SRC, double *__restrict__ thetaNew,
                                                double *__restrict__ spd, const double *__restrict__ props,
                                                const double *__restrict__ wflag, Grid g,
                                                const int *__restrict__ list, const int *__restrict__ count) {
  const int n = *count;
  for (int t = blockIdx.x * BLK + threadIdx.x; t < n; t += gridDim.x * BLK) {
    const int c = list[t];
    if (wflag[c] != 0.0) continue;
    mhd_pp_cell<D, NQ>(U, FS, SRC, thetaNew, spd, props, g, c);
  }
}

// ------------------------------------------------------------ cleaning speed
__device__ inline double glm_ch(const unsigned long long *chBits) {
  return __longlong_as_double((long long)chBits[0]) - __longlong_as_double((long long)chBits[1]);
}

// -------------------------------------------------------------- GLM sources
// Source terms (SourceTerms_K, MHD_GLM_K.H:316-547; no user forcing for the
// registry problems) of cell c at its NQ^D interior points, integrated into
// rhs with the interior quadrature weights (IntegrateRHS_K), point by point
// in ascending order.  SRC holds the limited interior-point states; divB is
// the cell's divergence from the face-averaged normal field (the flux kernel
// forms it from the two states of every face point).  u . grad(psi): the psi
// gradient from the cell's own face and interior values with the
// differentiation matrix Dm [NQ][NQ+2] over the points (-1/2, interior, +1/2)
// (DiffMatLUT, Numeric.cpp:47-99), or from central differences of the cell
// averages where the limiter took theta below 0.75.
template <int D, int NQ>
__device__ void mhd_source_rhs(const double *__restrict__ U, const double *__restrict__ FS,
                               const double *__restrict__ SRC, const double *__restrict__ theta, const Grid &g,
                               const double *__restrict__ Dm, size_t c, int i, int j, int k, double divB,
                               double *rhs) {
  constexpr int Q = MShape<D, NQ>::Q, NI = MShape<D, NQ>::NI, NZ = MShape<D, NQ>::NZ, FQ = MShape<D, NQ>::FQ;
  const size_t nc = g.ncell;
  auto own = [&](int f, int q) { return FS[fs_base(c, FQ * 9) + (size_t(f * Q + q) * 9 + 8) * FSB]; };
  auto src = [&](int p, int v) { return SRC[(size_t(p) * 9 + v) * nc + c]; };
  // u . grad(psi) at the interior points
  double gpsiDu[NI];
  const bool fd = theta[c] < 0.75;
#pragma unroll
  for (int nQ = 0; nQ < NQ; nQ++)
#pragma unroll
    for (int nR = 0; nR < NQ; nR++)
#pragma unroll
      for (int nS = 0; nS < NZ; nS++) {
        const int p = (nQ * NQ + nR) * NZ + nS;
        const double rho = src(p, 0);
        const double vx = src(p, 1) / rho, vy = src(p, 2) / rho, vz = src(p, 3) / rho;
        double dpx, dpy, dpz = 0.0;
        if (fd) {
          const double tdx = 1.0 / (2.0 * g.dx), tdy = 1.0 / (2.0 * g.dy);
          dpx = tdx * (load_u<D, 9>(U, g, i + 1, j, k, 8) - load_u<D, 9>(U, g, i - 1, j, k, 8));
          dpy = tdy * (load_u<D, 9>(U, g, i, j + 1, k, 8) - load_u<D, 9>(U, g, i, j - 1, k, 8));
          if (D == 3) {
            const double tdz = 1.0 / (2.0 * g.dz);
            dpz = tdz * (load_u<D, 9>(U, g, i, j, k + 1, 8) - load_u<D, 9>(U, g, i, j, k - 1, 8));
          }
        } else {
          const int qx = D == 2 ? nR : nR * NQ + nS, qy = D == 2 ? nQ : nQ * NQ + nS;
          dpx = Dm[nQ * (NQ + 2)] * own(0, qx) + Dm[nQ * (NQ + 2) + NQ + 1] * own(1, qx);
#pragma unroll
          for (int n = 1; n <= NQ; n++) dpx += Dm[nQ * (NQ + 2) + n] * src(((n - 1) * NQ + nR) * NZ + nS, 8);
          dpx /= g.dx;
          dpy = Dm[nR * (NQ + 2)] * own(2, qy) + Dm[nR * (NQ + 2) + NQ + 1] * own(3, qy);
#pragma unroll
          for (int n = 1; n <= NQ; n++) dpy += Dm[nR * (NQ + 2) + n] * src((nQ * NQ + n - 1) * NZ + nS, 8);
          dpy /= g.dy;
          if (D == 3) {
            const int qz = nQ * NQ + nR;
            dpz = Dm[nS * (NQ + 2)] * own(4, qz) + Dm[nS * (NQ + 2) + NQ + 1] * own(5, qz);
#pragma unroll
            for (int n = 1; n <= NQ; n++) dpz += Dm[nS * (NQ + 2) + n] * src((nQ * NQ + nR) * NZ + n - 1, 8);
            dpz /= g.dz;
          }
        }
        gpsiDu[p] = vx * dpx + vy * dpy + vz * dpz;
      }
  // the source terms of each point, then their integral
#pragma unroll
  for (int nQ = 0; nQ < NQ; nQ++)
#pragma unroll
    for (int nR = 0; nR < NQ; nR++)
#pragma unroll
      for (int nS = 0; nS < NZ; nS++) {
        const int p = (nQ * NQ + nR) * NZ + nS;
        const double w = D == 3 ? g.wt[nQ] * g.wt[nR] * g.wt[nS] : g.wt[nQ] * g.wt[nR];
        double st[9];
#pragma unroll
        for (int v = 0; v < 9; v++) st[v] = src(p, v);
        const double vx = st[1] / st[0], vy = st[2] / st[0], vz = st[3] / st[0];
        const double uDb = vx * st[5] + vy * st[6] + vz * st[7];
        const double f0 = 0.0;
        double sv[9];
        sv[0] = f0;
        sv[1] = f0 - divB * st[5];
        sv[2] = f0 - divB * st[6];
        sv[3] = f0 - divB * st[7];
        const double eForce = divB * uDb + gpsiDu[p] * st[8];
        sv[4] = f0 - eForce;
        sv[5] = f0 - divB * vx;
        sv[6] = f0 - divB * vy;
        sv[7] = f0 - divB * vz;
        sv[8] = -gpsiDu[p];
#pragma unroll
        for (int v = 0; v < 9; v++) rhs[v] += w * sv[v];
      }
}

// ------------------------------------------------------------ cleaning speed
#ifdef KFVM_SMALL_KERNELS
// The per-cell maxima spd [maxLam | maxVel] folded into chBits (zeroed by the
// caller) with one atomic per block each; the flux kernel reads
// c_h = max(c_f + |u|) - max(|u|) from them.
__global__ __launch_bounds__(BLK) void k_mhd_chmax(const double *__restrict__ spd, int64_t ncell,
                                                   unsigned long long *__restrict__ chBits) {
  __shared__ double sh[BLK];
  double lam = 0.0, vel = 0.0;
  for (int64_t c = int64_t(blockIdx.x) * BLK + threadIdx.x; c < ncell; c += int64_t(gridDim.x) * BLK) {
    const double l = spd[c], v = spd[ncell + c];
    lam = l > lam ? l : lam;
    vel = v > vel ? v : vel;
  }
  const double bl = block_max(lam, sh, threadIdx.x);
  const double bv = block_max(vel, sh, threadIdx.x);
  if (threadIdx.x == 0) {
    if (bl > 0.0) atomicMax(&chBits[0], (unsigned long long)__double_as_longlong(bl));
    if (bv > 0.0) atomicMax(&chBits[1], (unsigned long long)__double_as_longlong(bv));
  }
}
#endif  // KFVM_SMALL_KERNELS

}  // namespace dev
}  // namespace kfvm
