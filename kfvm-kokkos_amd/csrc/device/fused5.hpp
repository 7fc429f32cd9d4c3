// Fused reconstruction + Riemann + flux divergence + RK update of the Euler
// (5-variable) equations in sparse mode.
//
//   k_pp_rows     Common_K.H:299-643: the limiter's exact sequence for the
//                 cells whose linear states failed the fast test of the linear
//                 pass (and that the selector did not hand to WENO): their
//                 linear states are recomputed (Stencil_K.H:32-90), limited and
//                 stored in a compact row, like the WENO cells' states.
//   k_flux_fused  Stencil_K.H:32-90 linear face states of every cell that has
//                 no row, recomputed from the brick staged in LDS, then
//                 Common_K.H:180-297 Riemann solves, Numeric_K.H:109-196 flux
//                 divergence and Numeric_K.H:198-290 RK register update and
//                 error norm.  Nothing of the face states reaches HBM except
//                 the brick-boundary planes the linear pass wrote.
//
// Brick geometry and staging are those of k_linear_fused (RBox): a wave holds
// one row block of BX x 4 cells, so x neighbours are adjacent lanes and their
// face states move by lane shuffles; y and z neighbours exchange through LDS.
// Per face point q every thread solves the faces below its cell in x, y (and
// z); the faces above the last cell of the brick along an axis are solved by
// one extra wave per axis (wave a takes axis a), whose threads read the outer
// state from the surface planes or the rows.  A face shared by two bricks is
// solved in both from the same two states, so both see the same flux.
#pragma once

namespace kfvm {
namespace dev {

// Rows of the listed cells: row t holds the F*Q*5 face states of one cell,
// value (f*Q + q)*5 + v.
template <int D, int NQ>
struct RowShape {
  static constexpr int Q = Shape<D, NQ>::Q, F = Shape<D, NQ>::F, N = F * Q * 5;
};

// ------------------------------------------------------------- k_pp_rows
// One group of LPG lanes per listed cell (lane pl = face point pl): the
// point's five linear values from the stencil gathered through the boundary
// index map (weights W0 [F*Q][C], the stencil summed in ascending order as in
// every linear reconstruction here), the limiter with group reductions, the
// limited states into row rowBase + (an atomic count), and pos[c] = row.
template <int D, int NQ>
__global__ __launch_bounds__(BLK) void k_pp_rows(const double *__restrict__ U, double *__restrict__ rows,
                                                 int *__restrict__ pos, double *__restrict__ thetaNew,
                                                 const double *__restrict__ props,
                                                 const double *__restrict__ wflag, Grid g,
                                                 const double *__restrict__ W0, const int *__restrict__ list,
                                                 const int *__restrict__ count, const int *__restrict__ wenoCount,
                                                 int *__restrict__ rowCount) {
  constexpr auto &ST = StTab<D>::ST;
  constexpr int C = SInfo<D>::C, FQ = Shape<D, NQ>::F * Shape<D, NQ>::Q, LPG = PPGroup<FQ>::LPG;
  constexpr int GPW = 64 / LPG;
  const int n = *count;
  const int lane = threadIdx.x & 63, pl = lane % LPG;
  const int gid = (blockIdx.x * (BLK / 64) + (threadIdx.x >> 6)) * GPW + lane / LPG;
  const int ngroups = gridDim.x * (BLK / 64) * GPW;
  const size_t nc = g.ncell;
  const int rowBase = *wenoCount;  // the WENO cells' rows come first
  for (int t = gid; t < n; t += ngroups) {
    const int c = list[t];
    if (wflag[c] != 0.0) continue;  // uniform over the group
    int i, j, k;
    cell_ijk<D>(c, g, i, j, k);
    double ref[5];
#pragma unroll
    for (int w = 0; w < 5; w++) ref[w] = U[size_t(w) * nc + c];
    const PPBounds pb = grp_pp_bounds<D, LPG>(U, props, g, i, j, k, ref, pl);
    const bool own = pl < FQ;
    double st[5];
#pragma unroll
    for (int w = 0; w < 5; w++) st[w] = 0.0;
    if (own) {
      const double *wt = W0 + size_t(pl) * C;
#pragma unroll 3
      for (int m = 0; m < C; m++) {
        const CellRef r = map_cell<D, 5>(g, i + ST.off[m][0], j + ST.off[m][1], k + ST.off[m][2]);
        const double w = wt[m];
#pragma unroll
        for (int v = 0; v < 5; v++) st[v] = fma(w, ld(U, nc, r, v), st[v]);
      }
    } else {
#pragma unroll
      for (int w = 0; w < 5; w++) st[w] = ref[w];
    }
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
    int row = 0;
    if (pl == 0) row = rowBase + atomicAdd(rowCount, 1);
    row = __shfl(row, lane - pl);
    if (own) {
      double *dst = rows + size_t(row) * (FQ * 5) + pl * 5;
#pragma unroll
      for (int w = 0; w < 5; w++) dst[w] = st[w];
    }
    if (pl == 0) {
      thetaNew[c] = fmin(thD, thP);
      pos[c] = row;
    }
  }
}

// --------------------------------------------------------- k_flux_fused
template <int D, int NQ>
struct FusedShape {
  using Bx = RBox<D>;
  static constexpr int Q = Shape<D, NQ>::Q, F = Shape<D, NQ>::F;
  // extra faces per axis: one per transverse position of the brick
  static constexpr int EX = Bx::BY * Bx::BZ, EY = Bx::BX * Bx::BZ, EZ = D == 3 ? Bx::BX * Bx::BY : 0;
  static constexpr int RX = Bx::BX + 2, RY = Bx::BY + 2, RZ = D == 3 ? Bx::BZ + 2 : 1;
};

struct St5 {
  double v[5];
};

template <int DIR>
__device__ inline double hydro_riemann(const Grid &g, const double *uL, const double *uR, double *Fl) {
  if (g.rs == 2) return Hydro::solve_hllc<DIR>(g.gamma, uL, uR, Fl);
  return Hydro::solve_hll<DIR>(g.gamma, g.rs == 0, uL, uR, Fl);
}

template <int D, int NQ>
__global__ __launch_bounds__(RBox<D>::NT, 2) void k_flux_fused(const double *__restrict__ U,
                                                               const double *__restrict__ WQ,
                                                               const double *__restrict__ rows,
                                                               const int *__restrict__ pos,
                                                               const double *__restrict__ SB, Surf sb, Grid g,
                                                               RKArgs a) {
  using Bx = RBox<D>;
  using FS_ = FusedShape<D, NQ>;
  constexpr auto ST = make_stencil_tables<D>();
  constexpr int C = SInfo<D>::C, F = FS_::F, Q = FS_::Q, H = Bx::H, NP = Bx::N, NT = Bx::NT;
  constexpr int RN = RowShape<D, NQ>::N;
  constexpr int BX = Bx::BX, BY = Bx::BY, BZ = Bx::BZ;
  static_assert(BX == 16 && NT == BLK, "wave rows of 16 cells, one brick per block of BLK threads");
  __shared__ double lds[5 * NP];
  __shared__ double Ty[NT * 5];                      // y exchange (states, then fluxes)
  __shared__ double Tz[D == 3 ? NT * 5 : 1];         // z exchange
  __shared__ double Xt[(FS_::EX + FS_::EY + FS_::EZ) * 5];  // extra faces (states, then fluxes)
  __shared__ int rowTab[FS_::RX * FS_::RY * FS_::RZ];

  const int nbx = (g.nx + BX - 1) / BX, nby = (g.ny + BY - 1) / BY;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (b % nbx) * BX, j0 = ((b / nbx) % nby) * BY, k0 = (b / (nbx * nby)) * BZ;
  const int tid = threadIdx.x;
  const size_t nc = g.ncell;
  // stage the brick and its radius-2 halo (as k_linear_fused)
  for (int idx = tid; idx < NP; idx += NT) {
    const int xa = idx % Bx::LX, r = idx / Bx::LX, yb = r % Bx::LY, zc = r / Bx::LY;
    const int ex = box_out<BX, H>(xa), ey = box_out<BY, H>(yb);
    const int ez = D == 3 ? box_out<BZ, H>(zc) : 0;
    const int gi = i0 + xa - H, gj = j0 + yb - H, gk = D == 3 ? k0 + zc - H : k0;
    const bool need = ex * ex + ey * ey + ez * ez <= H * H && gi < g.nx + H && gj < g.ny + H &&
                      (D == 2 || gk < g.nz + H);
    if (need) {
      const CellRef cr = map_cell<D, 5>(g, gi, gj, gk);
#pragma unroll
      for (int w = 0; w < 5; w++) lds[w * NP + idx] = ld(U, nc, cr, w);
    } else {
#pragma unroll
      for (int w = 0; w < 5; w++) lds[w * NP + idx] = 0.0;
    }
  }
  // rows of the brick's cells and their face neighbours (-1: no row; cells
  // beyond a wall are never read, periodic ones wrap)
  for (int e = tid; e < FS_::RX * FS_::RY * FS_::RZ; e += NT) {
    const int x = i0 - 1 + e % FS_::RX, y = j0 - 1 + (e / FS_::RX) % FS_::RY;
    const int z = D == 3 ? k0 - 1 + e / (FS_::RX * FS_::RY) : k0;
    const bool wall = (x < 0 && g.bc[0] != 0) || (x >= g.nx && g.bc[1] != 0) || (y < 0 && g.bc[2] != 0) ||
                      (y >= g.ny && g.bc[3] != 0) ||
                      (D == 3 && ((z < 0 && g.bc[4] != 0) || (z >= g.nz && g.bc[5] != 0)));
    int r = -1;
    if (!wall) {
      const int xw = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
      const int yw = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
      const int zw = D == 3 ? (z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z)) : z;
      if (xw >= 0 && xw < g.nx && yw >= 0 && yw < g.ny && zw >= 0 && (D == 2 || zw < g.nz))
        r = pos[(size_t(zw) * g.ny + yw) * g.nx + xw];
    }
    rowTab[e] = r;
  }
  __syncthreads();

  const int tx = tid % BX, ty = (tid / BX) % BY, tz = tid / (BX * BY);
  const int ci = i0 + tx, cj = j0 + ty, ck = k0 + tz;
  const int lastX = min(BX, g.nx - i0) - 1, lastY = min(BY, g.ny - j0) - 1;
  const int lastZ = D == 3 ? min(BZ, g.nz - k0) - 1 : 0;
  const bool live = tx <= lastX && ty <= lastY && tz <= lastZ;
  const size_t c = live ? (size_t(ck) * g.ny + cj) * g.nx + ci : 0;
  const int corner = (tz * Bx::LY + ty) * Bx::LX + tx;
  auto row_at = [&](int x, int y, int z) {  // brick-local coordinates, -1 .. B
    return rowTab[(x + 1) + FS_::RX * ((y + 1) + FS_::RY * (D == 3 ? z + 1 : 0))];
  };
  const int myRow = live ? row_at(tx, ty, tz) : -1;
  // 2D batches: the brick's slice and its step size
  const int slice = D == 2 ? k0 : 0;
  const double betaDt = a.sliceDt ? a.beta * a.sliceDt[slice] : a.betaDt;
  const double bhatDt = a.sliceDt ? a.bhat * a.sliceDt[slice] : a.bhatDt;
  const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy, rdz = 1.0 / g.dz;

  // State at point q of face fn of the cell at brick-local (x,y,z) outside
  // this brick's cells (or across the domain edge): a row, the surface
  // planes, or the face boundary condition applied to `inner` (this brick's
  // cell's state on the shared face, face f = fn ^ 1 of it).
  auto outer_state = [&](int x, int y, int z, int fn, int q, const St5 &inner, St5 &out) {
    const int ax = fn >> 1;
    int gx = i0 + x, gy = j0 + y, gz = k0 + z;
    const int n_ax = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
    const int co = ax == 0 ? gx : (ax == 1 ? gy : gz);
    if (co < 0 || co >= n_ax) {
      const int f = fn ^ 1;  // the inner cell's face
      if (g.bc[f] != 0) {
        const unsigned m = g.negmask[f];
#pragma unroll
        for (int v = 0; v < 5; v++) out.v[v] = ((m >> v) & 1u) ? -inner.v[v] : inner.v[v];
        return;
      }
      const int w = co < 0 ? co + n_ax : co - n_ax;
      if (ax == 0) gx = w;
      else if (ax == 1) gy = w;
      else gz = w;
    }
    const int r = row_at(x, y, z);
    if (r >= 0) {
      const double *src = rows + size_t(r) * RN + (fn * Q + q) * 5;
#pragma unroll
      for (int v = 0; v < 5; v++) out.v[v] = src[v];
    } else {
      const double *src = SB + surf_index<D, Q>(sb, g, fn, gx, gy, gz, q, 0);
      const size_t T = size_t(sb.T[ax]);
#pragma unroll
      for (int v = 0; v < 5; v++) out.v[v] = src[v * T];
    }
  };

  double rhs[5];
#pragma unroll
  for (int v = 0; v < 5; v++) rhs[v] = 0.0;
  double vmax = 0.0;
  const int wave = tid >> 6, lane = tid & 63;
  const double *wq0 = WQ;

#pragma unroll 1
  for (int q = 0; q < Q; q++) {
    // ---- this cell's states at point q on its 2D faces (W,E,S,N,B,T)
    St5 s[F];
    if (myRow >= 0) {
      const double *src = rows + size_t(myRow) * RN;
#pragma unroll
      for (int f = 0; f < F; f++)
#pragma unroll
        for (int v = 0; v < 5; v++) s[f].v[v] = src[(f * Q + q) * 5 + v];
    } else {
#pragma unroll 1
      for (int v = 0; v < 5; v++) {
        // weights opaque to loop-invariant code motion, so they stay scalar
        // loads next to their FMAs instead of being hoisted out of the
        // variable loop (see k_linear_fused)
        const double *wt0 = wq0 + size_t(q) * C * F;
        asm volatile("" : "+s"(wt0));
        const cdouble *wt = (const cdouble *)wt0;
        const double *p = lds + v * NP + corner;
        double acc[F];
#pragma unroll
        for (int f = 0; f < F; f++) acc[f] = 0.0;
#pragma unroll
        for (int n = 0; n < C; n++) {
          const double x = p[(ST.off[n][0] + H) + Bx::LX * (ST.off[n][1] + H) +
                             (D == 3 ? Bx::LX * Bx::LY * (ST.off[n][2] + H) : 0)];
#pragma unroll
          for (int f = 0; f < F; f++) acc[f] = fma(wt[n * F + f], x, acc[f]);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, F, 0);
        }
        // v is a run-time index: select into the state registers
#pragma unroll
        for (int f = 0; f < F; f++)
#pragma unroll
          for (int w = 0; w < 5; w++) s[f].v[w] = w == v ? acc[f] : s[f].v[w];
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // ---- upper states for the y / z neighbours and the extra faces
#pragma unroll
    for (int v = 0; v < 5; v++) {
      Ty[tid * 5 + v] = s[3].v[v];
      if (D == 3) Tz[tid * 5 + v] = s[5].v[v];
    }
    if (live) {
      if (tx == lastX) {
#pragma unroll
        for (int v = 0; v < 5; v++) Xt[(ty + BY * tz) * 5 + v] = s[1].v[v];
      }
      if (ty == lastY) {
#pragma unroll
        for (int v = 0; v < 5; v++) Xt[(FS_::EX + tx + BX * tz) * 5 + v] = s[3].v[v];
      }
      if (D == 3 && tz == lastZ) {
#pragma unroll
        for (int v = 0; v < 5; v++) Xt[(FS_::EX + FS_::EY + tx + BX * ty) * 5 + v] = s[5].v[v];
      }
    }
    __syncthreads();
    // ---- lower faces of this cell
    double Fw[5], Fs[5], Fb[5];
    {
      St5 uL;
#pragma unroll
      for (int v = 0; v < 5; v++) uL.v[v] = __shfl_up(s[1].v[v], 1, 64);
      if (live && tx == 0) outer_state(-1, ty, tz, 1, q, s[0], uL);
      const double sp = hydro_riemann<0>(g, uL.v, s[0].v, Fw);
      if (live) vmax = fmax(vmax, sp);
    }
    {
      St5 uL;
      if (ty > 0) {
#pragma unroll
        for (int v = 0; v < 5; v++) uL.v[v] = Ty[(tid - BX) * 5 + v];
      } else {
        uL = s[2];
        if (live) outer_state(tx, -1, tz, 3, q, s[2], uL);
      }
      const double sp = hydro_riemann<1>(g, uL.v, s[2].v, Fs);
      if (live) vmax = fmax(vmax, sp);
    }
    if (D == 3) {
      St5 uL;
      if (tz > 0) {
#pragma unroll
        for (int v = 0; v < 5; v++) uL.v[v] = Tz[(tid - BX * BY) * 5 + v];
      } else {
        uL = s[4];
        if (live) outer_state(tx, ty, -1, 5, q, s[4], uL);
      }
      const double sp = hydro_riemann<2>(g, uL.v, s[4].v, Fb);
      if (live) vmax = fmax(vmax, sp);
    }
    // ---- extra faces: wave a solves the faces above the brick's last cells along axis a
    if (wave < D) {
      const int ne = wave == 0 ? FS_::EX : (wave == 1 ? FS_::EY : FS_::EZ);
      const int xoff = wave == 0 ? 0 : (wave == 1 ? FS_::EX : FS_::EX + FS_::EY);
      if (lane < ne) {
        int x, y, z;
        if (wave == 0) {
          x = lastX;
          y = lane % BY;
          z = lane / BY;
        } else if (wave == 1) {
          x = lane % BX;
          y = lastY;
          z = lane / BX;
        } else {
          x = lane % BX;
          y = lane / BX;
          z = lastZ;
        }
        const bool el = x <= lastX && y <= lastY && z <= lastZ;
        if (el) {
          St5 uL, uR;
          double Fl[5];
#pragma unroll
          for (int v = 0; v < 5; v++) uL.v[v] = Xt[(xoff + lane) * 5 + v];
          double sp;
          if (wave == 0) {
            outer_state(x + 1, y, z, 0, q, uL, uR);
            sp = hydro_riemann<0>(g, uL.v, uR.v, Fl);
          } else if (wave == 1) {
            outer_state(x, y + 1, z, 2, q, uL, uR);
            sp = hydro_riemann<1>(g, uL.v, uR.v, Fl);
          } else {
            outer_state(x, y, z + 1, 4, q, uL, uR);
            sp = hydro_riemann<2>(g, uL.v, uR.v, Fl);
          }
          vmax = fmax(vmax, sp);
#pragma unroll
          for (int v = 0; v < 5; v++) Xt[(xoff + lane) * 5 + v] = Fl[v];
        }
      }
    }
    __syncthreads();
    // ---- exchange the lower-face fluxes, then the divergence
#pragma unroll
    for (int v = 0; v < 5; v++) {
      Ty[tid * 5 + v] = Fs[v];
      if (D == 3) Tz[tid * 5 + v] = Fb[v];
    }
    __syncthreads();
    double Fe[5], Fn[5], Ft[5];
#pragma unroll
    for (int v = 0; v < 5; v++) Fe[v] = __shfl_down(Fw[v], 1, 64);
    if (live) {
      if (tx == lastX) {
#pragma unroll
        for (int v = 0; v < 5; v++) Fe[v] = Xt[(ty + BY * tz) * 5 + v];
      }
#pragma unroll
      for (int v = 0; v < 5; v++)
        Fn[v] = ty == lastY ? Xt[(FS_::EX + tx + BX * tz) * 5 + v] : Ty[(tid + BX) * 5 + v];
      if (D == 3) {
#pragma unroll
        for (int v = 0; v < 5; v++)
          Ft[v] = tz == lastZ ? Xt[(FS_::EX + FS_::EY + tx + BX * ty) * 5 + v] : Tz[(tid + BX * BY) * 5 + v];
      }
      // the reference divides each term by the spacing; multiplying by its
      // reciprocal differs by rounding only (as in k_flux_rk)
      if (D == 2) {
        const double w = g.wt[q];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          rhs[v] += w * (Fw[v] - Fe[v]) * rdx;
          rhs[v] += w * (Fs[v] - Fn[v]) * rdy;
        }
      } else {
        const double w = g.wt[q / NQ] * g.wt[q % NQ];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          rhs[v] += w * (Fw[v] - Fe[v]) * rdx;
          rhs[v] += w * (Fs[v] - Fn[v]) * rdy;
          rhs[v] += w * (Fb[v] - Ft[v]) * rdz;
        }
      }
    }
    __syncthreads();
  }

  // ---- RK register update and error norm (the epilogue of k_flux_rk)
  double errLoc = 0.0, posLoc = 1.0;
  const bool write = !a.sliceWrite || a.sliceWrite[slice] != 0;
  if (live) {
    if (a.mode == RK_PRIME) {
      if (write) {
#pragma unroll
        for (int v = 0; v < 5; v++) a.K[size_t(v) * nc + c] = rhs[v];
      }
    } else if (a.mode == RK_STAGE) {
#pragma unroll
      for (int v = 0; v < 5; v++) {
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
    } else if (write) {
      bool hasNan = false;
      double st[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const size_t o = size_t(v) * nc + c;
        const double u = U[o];
        const double uh = a.Uhat[o] + bhatDt * rhs[v];
        a.Uhat[o] = uh;
        a.K[o] = rhs[v];
        st[v] = u;
        hasNan = hasNan || isnan(u);
        const double e = (u - uh) / (a.atol + a.rtol * fmax(fabs(u), fabs(uh)));
        errLoc += e * e;
      }
      posLoc = hasNan ? -1.0 : Hydro::positivity(st);
    }
  }
  // reductions over the block (Ty is free again): same combination order as k_flux_rk
  double *sh = Ty;
  const double bmax = block_max(vmax, sh, tid);
  if (tid == 0 && bmax > 0.0) atomicMax(a.vmaxBits + slice, (unsigned long long)__double_as_longlong(bmax));
  if (a.mode == RK_FSAL) {
    const double es = block_sum(errLoc, sh, tid);
    const double pm = block_min(posLoc, sh, tid);
    if (tid == 0) {
      a.errPartial[b] = es;
      a.posPartial[b] = pm;
    }
  }
}

}  // namespace dev
}  // namespace kfvm
