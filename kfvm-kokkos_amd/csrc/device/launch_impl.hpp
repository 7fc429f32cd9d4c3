// Per-(dimension, face-rule) launchers of the stage kernels.  Each
// instantiation lives in its own translation unit (inst_*.hip) so the kernel
// variants compile in parallel; dispatch.hip selects one at run time.
#pragma once

#include "stage_kernels.hpp"

namespace kfvm {

inline unsigned stage_blocks(int64_t n) { return unsigned((n + dev::BLK - 1) / dev::BLK); }

template <int D, int NQ>
struct Launch {
  static hipError_t linear(int nv, const double *U, double *FS, double *WI, const Grid &g,
                           const double *W0, hipStream_t s) {
    if (nv == 5)
      hipLaunchKernelGGL((dev::k_linear<D, NQ, 5>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FS, WI, g, W0);
    else
      hipLaunchKernelGGL((dev::k_linear<D, NQ, 1>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FS, WI, g, W0);
    return hipGetLastError();
  }
  static hipError_t select(const double *U, const double *WI, const double *theta, double *wflag,
                           int *list, int *counter, double wThresh, const Grid &g, hipStream_t s) {
    hipLaunchKernelGGL((dev::k_select<D, NQ>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, WI, theta, wflag,
                       list, counter, wThresh, g);
    return hipGetLastError();
  }
  static hipError_t weno(int nv, const double *U, double *FS, const Grid &g, const double *FW,
                         const double *DW, const int *list, const int *counter, hipStream_t s) {
    const unsigned nb = std::min<unsigned>(stage_blocks(g.ncell), 1u << 16);
    if (nv == 5)
      hipLaunchKernelGGL((dev::k_weno<D, NQ, 5>), dim3(nb), dim3(dev::BLK), 0, s, U, FS, g, FW, DW, list, counter);
    else
      hipLaunchKernelGGL((dev::k_weno<D, NQ, 1>), dim3(nb), dim3(dev::BLK), 0, s, U, FS, g, FW, DW, list, counter);
    return hipGetLastError();
  }
  static hipError_t pp(const double *U, double *FS, double *theta, const Grid &g, hipStream_t s) {
    hipLaunchKernelGGL((dev::k_pp<D, NQ>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FS, theta, g);
    return hipGetLastError();
  }
  static hipError_t flux_rk(int nv, const double *U, const double *FS, const Grid &g,
                            const RKArgs &a, hipStream_t s) {
    if (nv == 5)
      hipLaunchKernelGGL((dev::k_flux_rk<D, NQ, 5>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FS, g, a);
    else
      hipLaunchKernelGGL((dev::k_flux_rk<D, NQ, 1>), dim3(stage_blocks(g.ncell)), dim3(dev::BLK), 0, s, U, FS, g, a);
    return hipGetLastError();
  }
};

extern template struct Launch<2, 2>;
extern template struct Launch<2, 3>;
extern template struct Launch<3, 2>;
extern template struct Launch<3, 3>;

}  // namespace kfvm
