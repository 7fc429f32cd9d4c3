// Definitions of the per-(dimension, face-rule) launchers declared in
// launch_decl.hpp; included by the instantiation units inst_*.hip only.
#pragma once

#include <cstdlib>
#include <type_traits>

#include "launch_decl.hpp"

#include "stage_kernels.hpp"

namespace kfvm {

template <int D, int NQ>
unsigned Launch<D, NQ>::bricks(const Grid &g) {
  using Bx = dev::RBox<D>;
  return unsigned(((g.nx + Bx::BX - 1) / Bx::BX) * ((g.ny + Bx::BY - 1) / Bx::BY) *
                  ((g.nz + Bx::BZ - 1) / Bx::BZ));
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::cell_props(const double *U, double *props, const Grid &g, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_cell_props<D>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, props, g);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::linear(int nv, const double *U, double *FS, double *WI, double *thetaNew,
                                 const double *props, const Grid &g, const double *W0, const double *W0T,
                                 int *ppList, int *ppCount, hipStream_t s, const LearnArgs *la) {
  if (nv == 5 && la)
    hipLaunchKernelGGL((dev::k_linear_fused<D, NQ, false, true>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s,
                       U, FS, WI, thetaNew, props, g, W0T, ppList, ppCount, nullptr, Surf{}, *la);
  else if (nv == 5)
    hipLaunchKernelGGL((dev::k_linear_fused<D, NQ>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, U, FS, WI,
                       thetaNew, props, g, W0T, ppList, ppCount);
  else
    hipLaunchKernelGGL((dev::k_linear<D, NQ, 1>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FS, WI,
                       thetaNew, g, W0);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::select(const double *U, const double *WI, const double *theta, double *wflag,
                                 int *list, int *counter, double wThresh, const Grid &g, int *pos, int cap,
                                 hipStream_t s) {
  hipLaunchKernelGGL((dev::k_select<D, NQ>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, WI, theta, wflag,
                     list, counter, wThresh, g, pos, cap);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::weno(int nv, const double *U, double *FS, double *thetaNew, const Grid &g,
                               const double *FW, const double *DW, const int *list, const int *counter,
                               double *WS, int cap, const double *WC, hipStream_t s) {
  if (nv == 5) {
    // persistent: every workgroup stages the weights once, then strides over the list
    if (!WC) return hipErrorInvalidValue;
    if constexpr (dev::WenoMfma<D, NQ>::USE) {
      // KFVM_WENO_SCALAR=1 (read at every launch) runs the scalar kernel
      // instead: the two are compared in the tests
      const char *ev = std::getenv("KFVM_WENO_SCALAR");
      if (!(ev && ev[0] == '1')) {
        constexpr int cellsPerBlock = 16 * dev::WenoMfma<D, NQ>::WPB;
        const unsigned nb = std::min<unsigned>(unsigned((g.ncell + cellsPerBlock - 1) / cellsPerBlock), 2048u);
        hipLaunchKernelGGL((dev::k_wenorec5_mfma<D, NQ>), dim3(nb), dim3(dev::BLK), 0, s, U, FS, g, WC, list,
                           counter, WS, WS ? cap : 0);
        return hipGetLastError();
      }
    }
    const unsigned nb = std::min<unsigned>(unsigned((g.ncell + dev::CPB - 1) / dev::CPB), 2048u);
    hipLaunchKernelGGL((dev::k_wenorec5<D, NQ>), dim3(nb), dim3(dev::BLK), 0, s, U, FS, g, WC, list, counter, WS,
                       WS ? cap : 0);
  } else {
    hipLaunchKernelGGL((dev::k_weno<D, NQ, 1>), dim3(std::min<unsigned>(stage_blocks(g.ncell), 1u << 16)),
                       dim3(dev::BLK), 0, s, U, FS, thetaNew, g, FW, DW, list, counter);
  }
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::linear_surf(const double *U, double *WI, double *thetaNew, const double *props,
                                      const Grid &g, const double *W0T, int *ppList, int *ppCount, double *SB,
                                      const Surf &sb, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_linear_fused<D, NQ, true>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, U,
                     nullptr, WI, thetaNew, props, g, W0T, ppList, ppCount, SB, sb);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::pp_rows(const double *U, double *rows, int *pos, double *thetaNew, const double *props,
                                  const double *wflag, const Grid &g, const double *W0, const int *list,
                                  const int *count, const int *wenoCount, int *rowCount, hipStream_t s) {
  const unsigned np = std::min<unsigned>(stage_blocks(g.ncell), 4096u);
  hipLaunchKernelGGL((dev::k_pp_rows<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, rows, pos, thetaNew, props,
                     wflag, g, W0, list, count, wenoCount, rowCount);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::flux_fused(const double *U, const double *WQ, const double *rows, const int *pos,
                                     const double *SB, const Surf &sb, const Grid &g, const RKArgs &a, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_flux_fused<D, NQ>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, U, WQ, rows, pos,
                     SB, sb, g, a);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::pp_lanes(const double *U, double *FS, double *thetaNew, const double *props,
                                   const double *wflag, const Grid &g, const int *list, const int *count,
                                   hipStream_t s) {
  const unsigned np = std::min<unsigned>(stage_blocks(g.ncell), 4096u);
  hipLaunchKernelGGL((dev::k_pp_lanes<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, FS, thetaNew, props, wflag, g,
                     list, count);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::weno_post(const double *U, double *FS, double *thetaNew, const double *props,
                                    const Grid &g, const int *list, const int *counter, double *WS, int cap,
                                    hipStream_t s) {
  const unsigned np = std::min<unsigned>(stage_blocks(g.ncell), 1u << 16);
  if (list)
    hipLaunchKernelGGL((dev::k_weno_post<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, FS, thetaNew, props, g, list,
                       counter, WS, WS ? cap : 0);
  else
    hipLaunchKernelGGL((dev::k_weno_post_cells<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, FS, thetaNew, props, g);
  return hipGetLastError();
}

template <int D, int NQ>
unsigned Launch<D, NQ>::mhd_bricks(const Grid &g) {
  using Bx = dev::MBox<D>;
  return unsigned(((g.nx + Bx::BX - 1) / Bx::BX) * ((g.ny + Bx::BY - 1) / Bx::BY) *
                  ((g.nz + Bx::BZ - 1) / Bx::BZ));
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::mhd_props(const double *U, double *props, const Grid &g, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_mhd_props<D>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, props, g);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::mhd_linear(const double *U, double *FS, double *SRC, double *WI, double *thetaNew,
                                     double *spd, const double *props, const Grid &g, const double *WT,
                                     int *ppList, int *ppCount, hipStream_t s) {
  // WI null: the learned selection does not read the KXRCF indicator
  hipLaunchKernelGGL((dev::k_mhd_linear<D, NQ>), dim3(mhd_bricks(g)), dim3(dev::MBox<D>::NT), 0, s, U, FS, SRC,
                     WI, thetaNew, spd, props, g, WT, ppList, ppCount);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::mhd_select(const double *U, const double *WI, const double *theta, double *wflag,
                                     int *list, int *counter, double wThresh, const Grid &g, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_select<D, NQ, dev::Mhd>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, WI,
                     theta, wflag, list, counter, wThresh, g, nullptr, 0);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::mhd_weno(const double *U, double *FS, double *SRC, const Grid &g, const double *FT,
                                   const double *DT, const double *CT, const int *list, const int *counter,
                                   double *thetaNew, double *spd, const double *props, const double *WC,
                                   hipStream_t s) {
  // weights in LDS: persistent (every workgroup stages the table once, then strides over the list)
  constexpr bool lds = dev::MWenoLds<D, NQ>::USE;
  if (lds && !WC) return hipErrorInvalidValue;
  const unsigned nb = std::min<unsigned>(unsigned((g.ncell + dev::MCPB - 1) / dev::MCPB), lds ? 2048u : 1u << 16);
  hipLaunchKernelGGL((dev::k_mhd_wenorec<D, NQ>), dim3(nb), dim3(dev::BLK), 0, s, U, FS, SRC, g, FT, DT, CT,
                     list, counter, WC);
  const unsigned np = std::min<unsigned>(stage_blocks(g.ncell), 1u << 16);
  if (list)
    hipLaunchKernelGGL((dev::k_mhd_weno_post<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, FS, SRC, g, list,
                       counter, thetaNew, spd, props);
  else
    hipLaunchKernelGGL((dev::k_mhd_weno_post_cells<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, FS, SRC, g,
                       thetaNew, spd, props);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::mhd_pp(const double *U, double *FS, double *SRC, double *thetaNew, double *spd,
                                 const double *props, const double *wflag, const Grid &g, const int *list,
                                 const int *count, hipStream_t s) {
  const unsigned np = std::min<unsigned>(stage_blocks(g.ncell), 1024u);
  hipLaunchKernelGGL((dev::k_mhd_pp<D, NQ>), dim3(np), dim3(dev::BLK), 0, s, U, FS, SRC, thetaNew, spd, props,
                     wflag, g, list, count);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::learned_select(int nv, const double *U, const double *eint, const double *theta,
                                         const double *nadFlag, const float *P, float thr, double *wflag, int *list,
                                         int *counter, const Grid &g, int *pos, int cap, hipStream_t s) {
  if (nv == 9)
    hipLaunchKernelGGL((dev::k_learned_select<D, 9>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, U, eint,
                       theta, nadFlag, P, thr, wflag, list, counter, g, pos, cap);
  else
    hipLaunchKernelGGL((dev::k_learned_select<D, 5>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, U, eint,
                       theta, nadFlag, P, thr, wflag, list, counter, g, pos, cap);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::features(int nv, const double *U, const double *eint, float *X, const Grid &g,
                                   hipStream_t s) {
  if (nv == 9)
    hipLaunchKernelGGL((dev::k_features<D, 9>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, eint, X, g);
  else
    hipLaunchKernelGGL((dev::k_features<D, 5>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, eint, X, g);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::labels(int nv, const double *U, const double *FSL, const double *FSW, float tau, float *Y,
                                 const Grid &g, hipStream_t s) {
  if (nv == 9)
    hipLaunchKernelGGL((dev::k_labels<D, NQ, 9>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FSL, FSW,
                       tau, Y, g);
  else
    hipLaunchKernelGGL((dev::k_labels<D, NQ, 5>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FSL, FSW,
                       tau, Y, g);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::logits(const float *X, const float *P, float *Z, int n, hipStream_t s) {
  hipLaunchKernelGGL((dev::k_logits<D>), dim3(stage_blocks(n)), dim3(dev::BLK), 0, s, X, P, Z, n);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::train_grad(const float *X, const float *Y, const int *idx, int n, int nb, uint64_t seed,
                                     const float *P, float posW, float *part, float *lossPart, hipStream_t s) {
  const unsigned nblk = unsigned((nb + dev::TB - 1) / dev::TB);
  hipLaunchKernelGGL((dev::k_train_grad<D>), dim3(nblk), dim3(dev::TB), 0, s, X, Y, idx, n, nb, seed, P, posW,
                     part, lossPart);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::nad(int nv, const double *Uold, const double *Unew, double fr, double fv, double *nadFlag,
                              int *count, const Grid &g, hipStream_t s) {
  if (nv == 9)
    hipLaunchKernelGGL((dev::k_nad<D, 9>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, Uold, Unew, fr, fv,
                       nadFlag, count, g);
  else
    hipLaunchKernelGGL((dev::k_nad<D, 5>), dim3(bricks(g)), dim3(dev::RBox<D>::NT), 0, s, Uold, Unew, fr, fv,
                       nadFlag, count, g);
  return hipGetLastError();
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::flow_stats(int nv, const double *U, const Grid &g, double ifac, double *out,
                                     hipStream_t s) {
  if (nv == 9)
    hipLaunchKernelGGL((dev::k_flow_stats<D, 9>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, g, ifac, out);
  else
    hipLaunchKernelGGL((dev::k_flow_stats<D, 5>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, g, ifac, out);
  return hipGetLastError();
}

template <int D, int NQ>
dim3 Launch<D, NQ>::flux_grid(const Grid &g, int shape) {
  auto count = [&](auto bk) {
    using Bk = decltype(bk);
    return dim3(unsigned((g.nx + Bk::X - 1) / Bk::X) * unsigned((g.ny + Bk::Y - 1) / Bk::Y) *
                unsigned((g.nz + Bk::Z - 1) / Bk::Z));
  };
  if (D == 2 && shape == 1) return count(dev::Brick<D, 1>{});
  if (D == 2 && shape == 2) return count(dev::Brick<D, 2>{});
  return count(dev::Brick<D, 0>{});
}

template <int D, int NQ>
hipError_t Launch<D, NQ>::flux_rk(int nv, const double *U, const double *FS, const Grid &g, const RKArgs &a,
                                  hipStream_t s, int shape) {
  auto run = [&](auto shp) {
    constexpr int S = decltype(shp)::value;
    using Bk = dev::Brick<D, S>;
    const dim3 blk(Bk::X, Bk::Y, Bk::Z);
    if (nv == 5)
      hipLaunchKernelGGL((dev::k_flux_rk<D, NQ, 5, S>), flux_grid(g, S), blk, 0, s, U, FS, g, a);
    else if (nv == 9)
      hipLaunchKernelGGL((dev::k_flux_rk<D, NQ, 9, S>), flux_grid(g, S), blk, 0, s, U, FS, g, a);
    else
      hipLaunchKernelGGL((dev::k_flux_rk<D, NQ, 1, S>), flux_grid(g, S), blk, 0, s, U, FS, g, a);
    return hipGetLastError();
  };
  if constexpr (D == 2) {
    if (shape == 1) return run(std::integral_constant<int, 1>{});
    if (shape == 2) return run(std::integral_constant<int, 2>{});
  }
  if (shape != 0 && D == 3) return hipErrorInvalidValue;
  return run(std::integral_constant<int, 0>{});
}

}  // namespace kfvm
