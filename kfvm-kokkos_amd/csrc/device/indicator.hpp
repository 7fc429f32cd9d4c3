// Learned troubled-cell indicator for the Euler equations: features, the
// network, training labels, training (forward / backward / Adam) and the
// a-posteriori numerical-admissibility (NAD) check.  The reference has no such
// component; it replaces the KXRCF selector (WenoSelect_K, Numeric_K.H:292-398)
// as the decision "linear GP reconstruction suffices / full WENO needed".
//
// Features of a cell (Net<D>::NIN = 2C floats), from the C cells of its
// radius-2 stencil in the stencil order, density rho and pressure p, with o
// the centre cell and m = 0..C-2 running over the other cells in order:
//   x[m]        = (rho_m - rho_o) / s_rho
//   x[C-1+m]    = (p_m - p_o) / s_p
//   x[2C-2]     = min(s_rho / |rho_o|, 4),   x[2C-1] = min(s_p / |p_o|, 4)
// with s_q = max_m |q_m - q_o| + 1e-300: the stencil's shape in [-1, 1] and
// its relative range.
//
// Network: x -> 32 (ReLU) -> 32 (ReLU) -> logit z.  Weights are bfloat16
// values (held as floats), sums in FP32 fma order over ascending inputs.
// Parameter layout (floats): W1[H][NIN], b1[H], W2[H][H], b2[H], w3[H], b3.
//
// Label of a cell: 1 if the full-WENO face states (after the limiter) differ
// from the linear ones by more than tau times the stencil's range of rho or E,
// ranges below 1e-3 of the cell's value counting as that floor (so grid-scale
// noise far below the state's magnitude is not a reason for WENO):
//   max_{f,q, v in {rho, E}} |F_weno - F_lin| / (range_v + 1e-3 |U_v| + 1e-300) > tau
#pragma once

namespace kfvm {
namespace dev {

template <int D>
struct Net {
  static constexpr int C = SInfo<D>::C, NIN = 2 * C, H = 32;
  static constexpr int oW1 = 0, oB1 = H * NIN, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H,
                       oB3 = oW3 + H, NP = oB3 + 1;
};

// Internal energy of cell idx: the limiter's per-cell plane (Euler) or from
// the nine GLM-MHD variables (E - |B|^2/2 - |m|^2/(2 rho) - psi^2/2, the
// order of MHD_GLM_EOS.H); both are even in the momenta and the field.
template <int NV>
__device__ inline double cell_eint(const double *__restrict__ U, const double *__restrict__ eint, size_t idx,
                                   size_t nc) {
  if (NV == 9) {
    double s[9];
#pragma unroll
    for (int v = 0; v < 9; v++) s[v] = U[size_t(v) * nc + idx];
    return Mhd::eint(s);
  }
  return eint[idx];
}

// Index of the centre cell in the stencil order.
template <int D>
constexpr int centre_index() {
  for (int n = 0; n < SInfo<D>::C; n++)
    if (StTab<D>::ST.off[n][0] == 0 && StTab<D>::ST.off[n][1] == 0 && StTab<D>::ST.off[n][2] == 0) return n;
  return -1;
}

// Features of cell (i,j,k) into x[NIN] (FP32).  eint: per-cell internal
// energy E - |m|^2 / (2 rho) (the limiter's cell properties), so p = (g-1) eint
// is the pressure of Hydro_EOS.H bit for bit; both are even in the momenta, so
// mirrored ghost cells need no sign.
template <int D, int NV>
__device__ inline void cell_features(const double *__restrict__ U, const double *__restrict__ eint,
                                     const Grid &g, int i, int j, int k, float *x) {
  constexpr auto &ST = StTab<D>::ST;
  constexpr int C = SInfo<D>::C, O = centre_index<D>();
  static_assert(O >= 0, "stencil without its centre");
  const double gm1 = g.gamma - 1.0;
  double rho[C], pr[C];
  // cells whose stencil stays inside the grid index it directly
  const bool inner = i >= 2 && i < g.nx - 2 && j >= 2 && j < g.ny - 2 && (D == 2 || (k >= 2 && k < g.nz - 2));
  const int c = (k * g.ny + j) * g.nx + i;
#pragma unroll
  for (int n = 0; n < C; n++) {
    size_t idx;
    if (inner)
      idx = size_t(c + ST.off[n][0] + g.nx * (ST.off[n][1] + (D == 3 ? g.ny * ST.off[n][2] : 0)));
    else
      idx = map_cell<D, 5>(g, i + ST.off[n][0], j + ST.off[n][1], k + ST.off[n][2]).idx;
    rho[n] = U[idx];
    pr[n] = gm1 * cell_eint<NV>(U, eint, idx, g.ncell);
  }
  double sr = 0.0, sp = 0.0;
#pragma unroll
  for (int n = 0; n < C; n++) {
    if (n == O) continue;
    sr = fmax(sr, fabs(rho[n] - rho[O]));
    sp = fmax(sp, fabs(pr[n] - pr[O]));
  }
  const double ir = 1.0 / (sr + 1e-300), ip = 1.0 / (sp + 1e-300);
#pragma unroll
  for (int n = 0; n < C; n++) {
    if (n == O) continue;
    const int m = n < O ? n : n - 1;
    x[m] = float((rho[n] - rho[O]) * ir);
    x[C - 1 + m] = float((pr[n] - pr[O]) * ip);
  }
  x[2 * C - 2] = float(fmin(sr / fabs(rho[O]), 4.0));
  x[2 * C - 1] = float(fmin(sp / fabs(pr[O]), 4.0));
}

// Logit of the network for features x; params in the layout above.
template <int D>
__device__ inline float net_logit(const float *__restrict__ P, const float *x) {
  using N = Net<D>;
  float h1[N::H], h2[N::H];
#pragma unroll 4
  for (int j = 0; j < N::H; j++) {
    float a = P[N::oB1 + j];
#pragma unroll
    for (int i = 0; i < N::NIN; i++) a = fmaf(P[N::oW1 + j * N::NIN + i], x[i], a);
    h1[j] = fmaxf(a, 0.0f);
  }
#pragma unroll 4
  for (int j = 0; j < N::H; j++) {
    float a = P[N::oB2 + j];
#pragma unroll
    for (int i = 0; i < N::H; i++) a = fmaf(P[N::oW2 + j * N::H + i], h1[i], a);
    h2[j] = fmaxf(a, 0.0f);
  }
  float z = P[N::oB3];
#pragma unroll
  for (int i = 0; i < N::H; i++) z = fmaf(P[N::oW3 + i], h2[i], z);
  return z;
}

// ---- the network on the matrix cores
// v_mfma_f32_16x16x32_bf16: D[16x16] = A[16x32] B[32x16] + C in FP32, lane l
// holding A[l%16][8(l/16)+i], B[8(l/16)+i][l%16] (i < 8) and
// D[4(l/16)+i][l%16] (i < 4).  The hidden units are the rows (A = weights,
// exact in bfloat16), the cells the columns (B = activations).  An FP32
// activation enters as three bfloat16 terms whose sum is the value to 2^-25,
// so each product is exact and only the FP32 accumulation order differs from
// net_logit's.
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ inline void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

// LDS offset (from a cell's brick corner) of the stencil cell behind feature k
// (the rho and p deviations; 0 for the range features and the padding).
template <int D>
constexpr int feature_lds(int k) {
  using Bx = RBox<D>;
  constexpr int C = SInfo<D>::C, O = centre_index<D>();
  int n = -1;
  if (k < C - 1) n = k < O ? k : k + 1;
  else if (k < 2 * C - 2) n = (k - (C - 1)) < O ? k - (C - 1) : k - (C - 1) + 1;
  if (n < 0) return 0;
  const auto &ST = StTab<D>::ST;
  return (ST.off[n][0] + Bx::H) + Bx::LX * (ST.off[n][1] + Bx::H) +
         (D == 3 ? Bx::LX * Bx::LY * (ST.off[n][2] + Bx::H) : 0);
}

// Learned selection: WENO where the network's logit exceeds thr, where the
// previous RHS evaluation limited the cell (theta < 1, as the KXRCF selector),
// or where the last step's admissibility check failed (nadFlag > 0).  One
// block per brick of the linear reconstruction's geometry: rho and p of the
// brick and its radius-2 halo are staged in LDS once.  Each thread takes its
// cell's stencil ranges (the same arithmetic as cell_features); each wave then
// runs the network for its 64 cells on the matrix cores, 16 cells at a time:
// lane l builds features 8(l/16)..+7 of cell l%16 of the group straight into
// the B operand, layer 1 is 2 x ceil(NIN/32) x 3 MFMAs, layer 2 2 x 3 (its
// K order follows layer 1's output layout, with W2's columns permuted to
// match), and the last layer a dot product reduced over the four lane groups.
template <int D, int NV>
__global__ __launch_bounds__(RBox<D>::NT) void k_learned_select(const double *__restrict__ U,
                                                                const double *__restrict__ eint,
                                                                const double *__restrict__ theta,
                                                                const double *__restrict__ nadFlag,
                                                                const float *__restrict__ P, float thr,
                                                                double *__restrict__ wflag,
                                                                int *__restrict__ list,
                                                                int *__restrict__ counter, Grid g,
                                                                int *__restrict__ pos, int cap) {
  using Bx = RBox<D>;
  using N = Net<D>;
  constexpr auto &ST = StTab<D>::ST;
  constexpr int C = SInfo<D>::C, O = centre_index<D>(), H = Bx::H, NP = Bx::N, NT = Bx::NT;
  constexpr int NIN = N::NIN, KS = (NIN + 31) / 32;
  static_assert(NT % 64 == 0 && N::H == 32, "waves of 64 cells, 32 hidden units");
  __shared__ double sr[NP], sp[NP];
  __shared__ double cr0[NT], cp0[NT], cir[NT], cip[NT];
  __shared__ float cxr[NT], cxp[NT], cz[NT];
  const int nbx = (g.nx + Bx::BX - 1) / Bx::BX, nby = (g.ny + Bx::BY - 1) / Bx::BY;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (b % nbx) * Bx::BX, j0 = ((b / nbx) % nby) * Bx::BY, k0 = (b / (nbx * nby)) * Bx::BZ;
  const int tid = threadIdx.x;
  const double gm1 = g.gamma - 1.0;
  for (int idx = tid; idx < NP; idx += NT) {
    const int a = idx % Bx::LX, r = idx / Bx::LX, bb = r % Bx::LY, cz_ = r / Bx::LY;
    const int gi = i0 + a - H, gj = j0 + bb - H, gk = D == 3 ? k0 + cz_ - H : k0;
    // box cells beyond the grid's far edges by more than the halo are never read
    if (gi < g.nx + H && gj < g.ny + H && (D == 2 || gk < g.nz + H)) {
      const size_t q = map_cell<D, NV>(g, gi, gj, gk).idx;
      sr[idx] = U[q];
      sp[idx] = gm1 * cell_eint<NV>(U, eint, q, g.ncell);
    } else {
      sr[idx] = 0.0;
      sp[idx] = 0.0;
    }
  }
  __syncthreads();
  auto corner_of = [&](int t) {
    const int tx = t % Bx::BX, ty = (t / Bx::BX) % Bx::BY, tz = t / (Bx::BX * Bx::BY);
    return (tz * Bx::LY + ty) * Bx::LX + tx;
  };
  const int tx = tid % Bx::BX, ty = (tid / Bx::BX) % Bx::BY, tz = tid / (Bx::BX * Bx::BY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  const bool live = i < g.nx && j < g.ny && (D == 2 || k < g.nz);
  const int c = live ? (k * g.ny + j) * g.nx + i : 0;
  {  // this thread's cell: centre values, inverse deviations, range features
    double r0 = 0.0, p0 = 0.0, ir = 0.0, ip = 0.0;
    float xr = 0.0f, xp = 0.0f;
    if (live) {
      const int corner = corner_of(tid);
      auto at = [&](int n) {
        return corner + (ST.off[n][0] + H) + Bx::LX * (ST.off[n][1] + H) +
               (D == 3 ? Bx::LX * Bx::LY * (ST.off[n][2] + H) : 0);
      };
      r0 = sr[at(O)];
      p0 = sp[at(O)];
      double smr = 0.0, smp = 0.0;
#pragma unroll
      for (int n = 0; n < C; n++) {
        if (n == O) continue;
        smr = fmax(smr, fabs(sr[at(n)] - r0));
        smp = fmax(smp, fabs(sp[at(n)] - p0));
      }
      ir = 1.0 / (smr + 1e-300);
      ip = 1.0 / (smp + 1e-300);
      xr = float(fmin(smr / fabs(r0), 4.0));
      xp = float(fmin(smp / fabs(p0), 4.0));
    }
    cr0[tid] = r0;
    cp0[tid] = p0;
    cir[tid] = ir;
    cip[tid] = ip;
    cxr[tid] = xr;
    cxp[tid] = xp;
  }
  __syncthreads();
  // ---- network, per wave on its 64 cells
  const int lane = tid & 63, lg = lane >> 4, col = lane & 15, wbase = tid & ~63;
  bf16x8 A1[2][KS], A2[2];
  float b1v[2][4], b2v[2][4], w3v[2][4];
  int foff[KS][8];
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int kk = 32 * ks + 8 * lg + e;
        A1[mt][ks][e] = (__bf16)(kk < NIN ? P[N::oW1 + (16 * mt + col) * NIN + kk] : 0.0f);
      }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int j1 = e < 4 ? 4 * lg + e : 16 + 4 * lg + (e - 4);
      A2[mt][e] = (__bf16)P[N::oW2 + (16 * mt + col) * N::H + j1];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      b1v[mt][e] = P[N::oB1 + 16 * mt + 4 * lg + e];
      b2v[mt][e] = P[N::oB2 + 16 * mt + 4 * lg + e];
      w3v[mt][e] = P[N::oW3 + 16 * mt + 4 * lg + e];
    }
  }
#pragma unroll
  for (int ks = 0; ks < KS; ks++)
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k0f = 32 * ks + e;
      foff[ks][e] = lg == 0 ? feature_lds<D>(k0f)
                            : (lg == 1 ? feature_lds<D>(k0f + 8)
                                       : (lg == 2 ? feature_lds<D>(k0f + 16) : feature_lds<D>(k0f + 24)));
    }
  const float b3 = P[N::oB3];
#pragma unroll 1
  for (int grp = 0; grp < 4; grp++) {
    const int cl = wbase + 16 * grp + col;  // the cell (thread index) of this lane's column
    const int corner = corner_of(cl);
    const double r0 = cr0[cl], p0 = cp0[cl], ir = cir[cl], ip = cip[cl];
    const float xr = cxr[cl], xp = cxp[cl];
    f32x4 acc[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) acc[mt][e] = b1v[mt][e];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      bf16x8 bh, bm, bl;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int kk = 32 * ks + 8 * lg + e;
        const double v = kk < C - 1 ? sr[corner + foff[ks][e]] : sp[corner + foff[ks][e]];
        float x;
        if (kk < C - 1) x = float((v - r0) * ir);
        else if (kk < 2 * C - 2) x = float((v - p0) * ip);
        else if (kk == 2 * C - 2) x = xr;
        else if (kk == 2 * C - 1) x = xp;
        else x = 0.0f;
        __bf16 h, m, l;
        split3(x, h, m, l);
        bh[e] = h;
        bm[e] = m;
        bl[e] = l;
      }
#pragma unroll
      for (int mt = 0; mt < 2; mt++) {
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[mt][ks], bh, acc[mt], 0, 0, 0);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[mt][ks], bm, acc[mt], 0, 0, 0);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[mt][ks], bl, acc[mt], 0, 0, 0);
      }
    }
    // layer 2: K slot 8(l/16)+e holds hidden unit 4(l/16)+e (e < 4) or 16+4(l/16)+e-4
    bf16x8 hh, hm, hl;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const float h1 = fmaxf(e < 4 ? acc[0][e] : acc[1][e - 4], 0.0f);
      __bf16 h, m, l;
      split3(h1, h, m, l);
      hh[e] = h;
      hm[e] = m;
      hl[e] = l;
    }
    f32x4 acc2[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
#pragma unroll
      for (int e = 0; e < 4; e++) acc2[mt][e] = b2v[mt][e];
      acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[mt], hh, acc2[mt], 0, 0, 0);
      acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[mt], hm, acc2[mt], 0, 0, 0);
      acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[mt], hl, acc2[mt], 0, 0, 0);
    }
    float z = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) z = fmaf(w3v[mt][e], fmaxf(acc2[mt][e], 0.0f), z);
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    if (lg == 0) cz[cl] = z + b3;
  }
  __syncthreads();
  bool flagged = false;
  if (live) {
    flagged = cz[tid] > thr || theta[c] < 1.0 || nadFlag[c] > 0.0;
    wflag[c] = flagged ? 1.0 : 0.0;
  }
  list_append(live, flagged, c, list, counter, pos, cap);
}

// The learned selection on the brick the Euler linear pass has staged (its
// five conserved planes in lds, radius-2 halo): the same features, network and
// decisions as k_learned_select, with the density from plane 0 and the
// pressure, formed from the staged planes, written over plane 1 (the momenta
// are no longer needed), and each
// cell's centre values, inverse deviations and range features passed to the
// network lanes by shuffles within the wave instead of through LDS.
template <int D>
__device__ void learned_tail(double *lds, const LearnArgs &la, const Grid &g, int i0, int j0, int k0, int tid) {
  using Bx = RBox<D>;
  using N = Net<D>;
  constexpr auto &ST = StTab<D>::ST;
  constexpr int C = SInfo<D>::C, O = centre_index<D>(), H = Bx::H, NP = Bx::N, NT = Bx::NT;
  constexpr int NIN = N::NIN, KS = (NIN + 31) / 32;
  static_assert(NT % 64 == 0 && N::H == 32, "waves of 64 cells, 32 hidden units");
  const double *sr = lds;
  double *sp = lds + NP;
  const double gm1 = g.gamma - 1.0;
  __syncthreads();  // every linear pass of the block has read plane 1
  for (int idx = tid; idx < NP; idx += NT) {
    const int a = idx % Bx::LX, r = idx / Bx::LX, bb = r % Bx::LY, cz_ = r / Bx::LY;
    const int ex = box_out<Bx::BX, H>(a), ey = box_out<Bx::BY, H>(bb);
    const int ez = D == 3 ? box_out<Bx::BZ, H>(cz_) : 0;
    const int gi = i0 + a - H, gj = j0 + bb - H, gk = D == 3 ? k0 + cz_ - H : k0;
    // the cells the linear pass staged (every stencil cell of a live cell)
    // from the five staged planes: E - |m|^2 / (2 rho) is even in the mirrored
    // momenta, so it equals the limiter's per-cell plane bit for bit without
    // a global load
    if (ex * ex + ey * ey + ez * ez <= H * H && gi < g.nx + H && gj < g.ny + H && (D == 2 || gk < g.nz + H)) {
      double st[5];
#pragma unroll
      for (int w = 0; w < 5; w++) st[w] = lds[w * NP + idx];
      sp[idx] = gm1 * Hydro::eint(st);
    }
  }
  __syncthreads();
  auto corner_of = [&](int t) {
    const int tx = t % Bx::BX, ty = (t / Bx::BX) % Bx::BY, tz = t / (Bx::BX * Bx::BY);
    return (tz * Bx::LY + ty) * Bx::LX + tx;
  };
  const int tx = tid % Bx::BX, ty = (tid / Bx::BX) % Bx::BY, tz = tid / (Bx::BX * Bx::BY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  const bool live = i < g.nx && j < g.ny && (D == 2 || k < g.nz);
  const int c = live ? (k * g.ny + j) * g.nx + i : 0;
  // this thread's cell: centre values, inverse deviations, range features
  double r0 = 0.0, p0 = 0.0, ir = 0.0, ip = 0.0;
  float xr = 0.0f, xp = 0.0f;
  if (live) {
    const int corner = corner_of(tid);
    auto at = [&](int n) {
      return corner + (ST.off[n][0] + H) + Bx::LX * (ST.off[n][1] + H) +
             (D == 3 ? Bx::LX * Bx::LY * (ST.off[n][2] + H) : 0);
    };
    r0 = sr[at(O)];
    p0 = sp[at(O)];
    double smr = 0.0, smp = 0.0;
#pragma unroll
    for (int n = 0; n < C; n++) {
      if (n == O) continue;
      smr = fmax(smr, fabs(sr[at(n)] - r0));
      smp = fmax(smp, fabs(sp[at(n)] - p0));
    }
    ir = 1.0 / (smr + 1e-300);
    ip = 1.0 / (smp + 1e-300);
    xr = float(fmin(smr / fabs(r0), 4.0));
    xp = float(fmin(smp / fabs(p0), 4.0));
  }
  // ---- network, per wave on its 64 cells
  const float *P = la.P;
  const int lane = tid & 63, lg = lane >> 4, col = lane & 15, wbase = tid & ~63;
  bf16x8 A1[2][KS], A2[2];
  float b1v[2][4], b2v[2][4], w3v[2][4];
  int foff[KS][8];
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int kk = 32 * ks + 8 * lg + e;
        A1[mt][ks][e] = (__bf16)(kk < NIN ? P[N::oW1 + (16 * mt + col) * NIN + kk] : 0.0f);
      }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int j1 = e < 4 ? 4 * lg + e : 16 + 4 * lg + (e - 4);
      A2[mt][e] = (__bf16)P[N::oW2 + (16 * mt + col) * N::H + j1];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      b1v[mt][e] = P[N::oB1 + 16 * mt + 4 * lg + e];
      b2v[mt][e] = P[N::oB2 + 16 * mt + 4 * lg + e];
      w3v[mt][e] = P[N::oW3 + 16 * mt + 4 * lg + e];
    }
  }
#pragma unroll
  for (int ks = 0; ks < KS; ks++)
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k0f = 32 * ks + e;
      foff[ks][e] = lg == 0 ? feature_lds<D>(k0f)
                            : (lg == 1 ? feature_lds<D>(k0f + 8)
                                       : (lg == 2 ? feature_lds<D>(k0f + 16) : feature_lds<D>(k0f + 24)));
    }
  const float b3 = P[N::oB3];
  float myz = 0.0f;
#pragma unroll 1
  for (int grp = 0; grp < 4; grp++) {
    const int src = 16 * grp + col;  // the lane (of this wave) owning this lane's column cell
    const int cl = wbase + src;
    const int corner = corner_of(cl);
    const double cr0 = __shfl(r0, src), cp0 = __shfl(p0, src), cir = __shfl(ir, src), cip = __shfl(ip, src);
    const float cxr = __shfl(xr, src), cxp = __shfl(xp, src);
    f32x4 acc[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) acc[mt][e] = b1v[mt][e];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      bf16x8 bh, bm, bl;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int kk = 32 * ks + 8 * lg + e;
        const double v = kk < C - 1 ? sr[corner + foff[ks][e]] : sp[corner + foff[ks][e]];
        float x;
        if (kk < C - 1) x = float((v - cr0) * cir);
        else if (kk < 2 * C - 2) x = float((v - cp0) * cip);
        else if (kk == 2 * C - 2) x = cxr;
        else if (kk == 2 * C - 1) x = cxp;
        else x = 0.0f;
        __bf16 h, m, l;
        split3(x, h, m, l);
        bh[e] = h;
        bm[e] = m;
        bl[e] = l;
      }
#pragma unroll
      for (int mt = 0; mt < 2; mt++) {
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[mt][ks], bh, acc[mt], 0, 0, 0);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[mt][ks], bm, acc[mt], 0, 0, 0);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[mt][ks], bl, acc[mt], 0, 0, 0);
      }
    }
    bf16x8 hh, hm, hl;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const float h1 = fmaxf(e < 4 ? acc[0][e] : acc[1][e - 4], 0.0f);
      __bf16 h, m, l;
      split3(h1, h, m, l);
      hh[e] = h;
      hm[e] = m;
      hl[e] = l;
    }
    f32x4 acc2[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
#pragma unroll
      for (int e = 0; e < 4; e++) acc2[mt][e] = b2v[mt][e];
      acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[mt], hh, acc2[mt], 0, 0, 0);
      acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[mt], hm, acc2[mt], 0, 0, 0);
      acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[mt], hl, acc2[mt], 0, 0, 0);
    }
    float z = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
      for (int e = 0; e < 4; e++) z = fmaf(w3v[mt][e], fmaxf(acc2[mt][e], 0.0f), z);
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    if (lg == grp) myz = z + b3;  // the owner lane of this group's cells
  }
  bool flagged = false;
  if (live) {
    flagged = myz > la.thr || la.theta[c] < 1.0 || la.nadFlag[c] > 0.0;
    la.wflag[c] = flagged ? 1.0 : 0.0;
  }
  list_append(live, flagged, c, la.list, la.counter, la.pos, la.cap);
}

// Features of every cell, feature-major rows: X[c][NIN] (training data).
template <int D, int NV>
__global__ __launch_bounds__(BLK) void k_features(const double *__restrict__ U, const double *__restrict__ eint,
                                                  float *__restrict__ X, Grid g) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int i, j, k;
  cell_ijk<D>(c, g, i, j, k);
  float x[Net<D>::NIN];
  cell_features<D, NV>(U, eint, g, i, j, k, x);
#pragma unroll
  for (int n = 0; n < Net<D>::NIN; n++) X[size_t(c) * Net<D>::NIN + n] = x[n];
}

// Network logits of n feature rows (inference check).
template <int D>
__global__ __launch_bounds__(BLK) void k_logits(const float *__restrict__ X, const float *__restrict__ P,
                                                float *__restrict__ Z, int n) {
  const int s = blockIdx.x * BLK + threadIdx.x;
  if (s >= n) return;
  float x[Net<D>::NIN];
#pragma unroll
  for (int i = 0; i < Net<D>::NIN; i++) x[i] = X[size_t(s) * Net<D>::NIN + i];
  Z[s] = net_logit<D>(P, x);
}

// Labels from the linear (FSL) and full-WENO (FSW) face states of every cell.
template <int D, int NQ, int NV>
__global__ __launch_bounds__(BLK) void k_labels(const double *__restrict__ U, const double *__restrict__ FSL,
                                                const double *__restrict__ FSW, float tau,
                                                float *__restrict__ Y, Grid g) {
  constexpr auto &ST = StTab<D>::ST;
  constexpr int C = SInfo<D>::C, FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q;
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int i, j, k;
  cell_ijk<D>(c, g, i, j, k);
  const size_t nc = g.ncell;
  double lo[2] = {1.7976931348623157e308, 1.7976931348623157e308}, hi[2] = {-1.7976931348623157e308,
                                                                            -1.7976931348623157e308};
#pragma unroll
  for (int n = 0; n < C; n++) {
    const CellRef r = map_cell<D, NV>(g, i + ST.off[n][0], j + ST.off[n][1], k + ST.off[n][2]);
    const double d = ld(U, nc, r, 0), e = ld(U, nc, r, 4);
    lo[0] = fmin(lo[0], d);
    hi[0] = fmax(hi[0], d);
    lo[1] = fmin(lo[1], e);
    hi[1] = fmax(hi[1], e);
  }
  const double sc[2] = {hi[0] - lo[0] + 1e-3 * fabs(U[c]) + 1e-300,
                        hi[1] - lo[1] + 1e-3 * fabs(U[4 * nc + c]) + 1e-300};
  double d = 0.0;
#pragma unroll 1
  for (int fq = 0; fq < FQ; fq++) {
    const size_t fb = fs_base(c, FQ * NV);
    const size_t o0 = fb + (size_t(fq) * NV + 0) * FSB, o4 = fb + (size_t(fq) * NV + 4) * FSB;
    d = fmax(d, fabs(FSW[o0] - FSL[o0]) / sc[0]);
    d = fmax(d, fabs(FSW[o4] - FSL[o4]) / sc[1]);
  }
  Y[c] = d > double(tau) ? 1.0f : 0.0f;
}

// ------------------------------------------------------------------ training
constexpr int TB = 64;  // samples per training block (one wave)

__device__ inline uint32_t mix32(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return uint32_t((x ^ (x >> 31)) >> 32);
}

// Sample index b of a batch: explicit (idx != nullptr) or drawn uniformly from
// [0, n) by a counter hash of (seed, b).
__device__ inline int batch_index(const int *__restrict__ idx, int b, int n, uint64_t seed) {
  return idx ? idx[b] : int(uint64_t(mix32(seed * 0x100000001B3ull + uint64_t(b))) % uint64_t(n));
}

// Per-block partial gradient of the class-weighted logistic loss
//   L_s = w_s * softplus(z_s (1 - 2 y_s)),  w_s = posW if y_s = 1 else 1,
// over TB samples of a batch of nb: part[blk][NP] gradient sums, lossPart[blk].
template <int D>
__global__ __launch_bounds__(TB) void k_train_grad(const float *__restrict__ X, const float *__restrict__ Y,
                                                   const int *__restrict__ idx, int n, int nb,
                                                   uint64_t seed, const float *__restrict__ P,
                                                   float posW, float *__restrict__ part,
                                                   float *__restrict__ lossPart) {
  using N = Net<D>;
  __shared__ float xs[TB][N::NIN + 1];
  __shared__ float h1s[TB][N::H + 1], h2s[TB][N::H + 1], d1s[TB][N::H + 1], d2s[TB][N::H + 1];
  __shared__ float dzs[TB], ls[TB];
  const int t = threadIdx.x;
  const int b = blockIdx.x * TB + t;
  const bool live = b < nb;
  const int s = live ? batch_index(idx, b, n, seed) : 0;
  float x[N::NIN];
#pragma unroll
  for (int i = 0; i < N::NIN; i++) {
    x[i] = live ? X[size_t(s) * N::NIN + i] : 0.0f;
    xs[t][i] = x[i];
  }
  const float y = live ? Y[s] : 0.0f;
  // forward
  float a1[N::H], a2[N::H];
#pragma unroll 4
  for (int j = 0; j < N::H; j++) {
    float a = P[N::oB1 + j];
#pragma unroll
    for (int i = 0; i < N::NIN; i++) a = fmaf(P[N::oW1 + j * N::NIN + i], x[i], a);
    a1[j] = a;
    h1s[t][j] = fmaxf(a, 0.0f);
  }
#pragma unroll 4
  for (int j = 0; j < N::H; j++) {
    float a = P[N::oB2 + j];
#pragma unroll
    for (int i = 0; i < N::H; i++) a = fmaf(P[N::oW2 + j * N::H + i], fmaxf(a1[i], 0.0f), a);
    a2[j] = a;
    h2s[t][j] = fmaxf(a, 0.0f);
  }
  float z = P[N::oB3];
#pragma unroll
  for (int i = 0; i < N::H; i++) z = fmaf(P[N::oW3 + i], fmaxf(a2[i], 0.0f), z);
  // loss and its derivative in z
  const float w = y > 0.5f ? posW : 1.0f;
  const float m = y > 0.5f ? -z : z;  // softplus argument
  const float sp = fmaxf(m, 0.0f) + log1pf(expf(-fabsf(m)));
  const float sig = 1.0f / (1.0f + expf(-z));
  const float dz = live ? w * (sig - y) : 0.0f;
  ls[t] = live ? w * sp : 0.0f;
  dzs[t] = dz;
  // backward through the hidden layers
#pragma unroll
  for (int j = 0; j < N::H; j++) d2s[t][j] = a2[j] > 0.0f ? dz * P[N::oW3 + j] : 0.0f;
#pragma unroll 4
  for (int i = 0; i < N::H; i++) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < N::H; j++) acc = fmaf(d2s[t][j], P[N::oW2 + j * N::H + i], acc);
    d1s[t][i] = a1[i] > 0.0f ? acc : 0.0f;
  }
  __syncthreads();
  // parameter gradients summed over the block's samples (ascending sample order)
  float *out = part + size_t(blockIdx.x) * N::NP;
  for (int p = t; p < N::NP; p += TB) {
    float acc = 0.0f;
    if (p < N::oB1) {
      const int j = p / N::NIN, i = p % N::NIN;
      for (int q = 0; q < TB; q++) acc = fmaf(d1s[q][j], xs[q][i], acc);
    } else if (p < N::oW2) {
      const int j = p - N::oB1;
      for (int q = 0; q < TB; q++) acc += d1s[q][j];
    } else if (p < N::oB2) {
      const int j = (p - N::oW2) / N::H, i = (p - N::oW2) % N::H;
      for (int q = 0; q < TB; q++) acc = fmaf(d2s[q][j], h1s[q][i], acc);
    } else if (p < N::oW3) {
      const int j = p - N::oB2;
      for (int q = 0; q < TB; q++) acc += d2s[q][j];
    } else if (p < N::oB3) {
      const int i = p - N::oW3;
      for (int q = 0; q < TB; q++) acc = fmaf(dzs[q], h2s[q][i], acc);
    } else {
      for (int q = 0; q < TB; q++) acc += dzs[q];
    }
    out[p] = acc;
  }
  if (t == 0) {
    float l = 0.0f;
    for (int q = 0; q < TB; q++) l += ls[q];
    lossPart[blockIdx.x] = l;
  }
}

#ifdef KFVM_SMALL_KERNELS
// grad[p] = sum over blocks of part[blk][p] / nb (fixed order); loss likewise.
__global__ __launch_bounds__(BLK) void k_train_reduce(const float *__restrict__ part,
                                                      const float *__restrict__ lossPart, int nblk,
                                                      int np, int nb, float *__restrict__ grad,
                                                      float *__restrict__ loss) {
  const int p = blockIdx.x * BLK + threadIdx.x;
  if (p < np) {
    float acc = 0.0f;
    for (int b = 0; b < nblk; b++) acc += part[size_t(b) * np + p];
    grad[p] = acc / float(nb);
  }
  if (p == 0) {
    float l = 0.0f;
    for (int b = 0; b < nblk; b++) l += lossPart[b];
    *loss = l / float(nb);
  }
}

// bfloat16 rounding (to nearest, ties to even) of a float, kept as a float.
__device__ inline float bf16_round(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

// Adam step on the FP32 master parameters M; Pbf receives their bfloat16
// values, which the forward pass uses.
__global__ __launch_bounds__(BLK) void k_adam(float *__restrict__ M, float *__restrict__ Pbf,
                                              const float *__restrict__ grad, float *__restrict__ m1,
                                              float *__restrict__ m2, int np, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * BLK + threadIdx.x;
  if (p >= np) return;
  const float gp = grad[p];
  const float a = b1 * m1[p] + (1.0f - b1) * gp;
  const float v = b2 * m2[p] + (1.0f - b2) * gp * gp;
  m1[p] = a;
  m2[p] = v;
  const float th = M[p] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  M[p] = th;
  Pbf[p] = bf16_round(th);
}

#endif  // KFVM_SMALL_KERNELS

// ------------------------------------------------------------- NAD check
// Numerical admissibility of the new cell averages: rho and p of every cell
// within [min - delta, max + delta] of its 3^D neighbourhood at the previous
// step, delta = max(fv max(|min|, |max|), fr (max - min)).  Marks the
// violating cells in nadFlag (1 / 0) and counts them.  One block per brick:
// rho and p of the old state over the brick and a one-cell halo are staged in
// LDS.
template <int D, int NV>
__global__ __launch_bounds__(RBox<D>::NT) void k_nad(const double *__restrict__ Uold,
                                                     const double *__restrict__ Unew, double fr, double fv,
                                                     double *__restrict__ nadFlag, int *__restrict__ count,
                                                     Grid g) {
  using Bx = RBox<D>;
  constexpr int LX = Bx::BX + 2, LY = Bx::BY + 2, LZ = D == 3 ? Bx::BZ + 2 : 1, NP = LX * LY * LZ;
  __shared__ double sr[NP], sp[NP];
  const int nbx = (g.nx + Bx::BX - 1) / Bx::BX, nby = (g.ny + Bx::BY - 1) / Bx::BY;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (b % nbx) * Bx::BX, j0 = ((b / nbx) % nby) * Bx::BY, k0 = (b / (nbx * nby)) * Bx::BZ;
  const int tid = threadIdx.x;
  const size_t nc = g.ncell;
  const double gm1 = g.gamma - 1.0;
  for (int idx = tid; idx < NP; idx += Bx::NT) {
    const int a = idx % LX, r = idx / LX, bb = r % LY, cz = r / LY;
    const int gi = i0 + a - 1, gj = j0 + bb - 1, gk = D == 3 ? k0 + cz - 1 : k0;
    if (gi < g.nx + 1 && gj < g.ny + 1 && (D == 2 || gk < g.nz + 1)) {
      const CellRef cr = map_cell<D, NV>(g, gi, gj, gk);
      double s[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) s[v] = ld(Uold, nc, cr, v);
      sr[idx] = s[0];
      sp[idx] = gm1 * (NV == 9 ? Mhd::eint(s) : Hydro::eint(s));
    }
  }
  __syncthreads();
  const int tx = tid % Bx::BX, ty = (tid / Bx::BX) % Bx::BY, tz = tid / (Bx::BX * Bx::BY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  bool bad = false;
  if (i < g.nx && j < g.ny && (D == 2 || k < g.nz)) {
    const size_t c = (size_t(k) * g.ny + j) * g.nx + i;
    double lo[2] = {1.7976931348623157e308, 1.7976931348623157e308};
    double hi[2] = {-1.7976931348623157e308, -1.7976931348623157e308};
    constexpr int zl = D == 3 ? -1 : 0, zh = D == 3 ? 1 : 0;
#pragma unroll
    for (int xo = -1; xo <= 1; xo++)
#pragma unroll
      for (int yo = -1; yo <= 1; yo++)
#pragma unroll
        for (int zo = zl; zo <= zh; zo++) {
          const int q = (tx + 1 + xo) + LX * ((ty + 1 + yo) + LY * (D == 3 ? tz + 1 + zo : 0));
          lo[0] = fmin(lo[0], sr[q]);
          hi[0] = fmax(hi[0], sr[q]);
          lo[1] = fmin(lo[1], sp[q]);
          hi[1] = fmax(hi[1], sp[q]);
        }
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) s[v] = Unew[size_t(v) * nc + c];
    const double val[2] = {s[0], gm1 * (NV == 9 ? Mhd::eint(s) : Hydro::eint(s))};
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const double del = fmax(fv * fmax(fabs(lo[q]), fabs(hi[q])), fr * (hi[q] - lo[q]));
      bad = bad || !(val[q] >= lo[q] - del && val[q] <= hi[q] + del);
    }
    nadFlag[c] = bad ? 1.0 : 0.0;
  }
  const unsigned long long mask = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && mask) atomicAdd(count, __popcll(mask));
}

}  // namespace dev
}  // namespace kfvm
