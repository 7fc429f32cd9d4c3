// Per-(dimension, face-rule) launchers of the stage kernels: the class and
// its member declarations.  The members are defined (launch_impl.hpp) and
// explicitly instantiated one (D, NQ) per translation unit (inst_*.hip), so
// the kernel variants compile in parallel; dispatch.hip selects one at run
// time and, seeing only these declarations, compiles none of them again.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_api.hpp"

namespace kfvm {

template <int D, int NQ>
struct Launch {
  static unsigned bricks(const Grid &g);
  static hipError_t cell_props(const double *U, double *props, const Grid &g, hipStream_t s);
  static hipError_t linear(int nv, const double *U, double *FS, double *WI, double *thetaNew,
                           const double *props, const Grid &g, const double *W0, const double *W0T,
                           int *ppList, int *ppCount, hipStream_t s, const LearnArgs *la);
  static hipError_t select(const double *U, const double *WI, const double *theta, double *wflag,
                           int *list, int *counter, double wThresh, const Grid &g, int *pos, int cap,
                           hipStream_t s);
  static hipError_t weno(int nv, const double *U, double *FS, double *thetaNew, const Grid &g,
                         const double *FW, const double *DW, const int *list, const int *counter,
                         double *WS, int cap, const double *WC, hipStream_t s);
  // ---- fused path (Euler, sparse WENO)
  static hipError_t linear_surf(const double *U, double *WI, double *thetaNew, const double *props,
                                const Grid &g, const double *W0T, int *ppList, int *ppCount, double *SB,
                                const Surf &sb, hipStream_t s);
  static hipError_t pp_rows(const double *U, double *rows, int *pos, double *thetaNew, const double *props,
                            const double *wflag, const Grid &g, const double *W0, const int *list,
                            const int *count, const int *wenoCount, int *rowCount, hipStream_t s);
  static hipError_t flux_fused(const double *U, const double *WQ, const double *rows, const int *pos,
                               const double *SB, const Surf &sb, const Grid &g, const RKArgs &a, hipStream_t s);
  static hipError_t pp_lanes(const double *U, double *FS, double *thetaNew, const double *props,
                             const double *wflag, const Grid &g, const int *list, const int *count,
                             hipStream_t s);
  static hipError_t weno_post(const double *U, double *FS, double *thetaNew, const double *props,
                              const Grid &g, const int *list, const int *counter, double *WS, int cap,
                              hipStream_t s);
  // ---- GLM-MHD
  static unsigned mhd_bricks(const Grid &g);
  static hipError_t mhd_props(const double *U, double *props, const Grid &g, hipStream_t s);
  static hipError_t mhd_linear(const double *U, double *FS, double *SRC, double *WI, double *thetaNew, double *spd,
                               const double *props, const Grid &g, const double *WT, int *ppList, int *ppCount,
                               hipStream_t s);
  static hipError_t mhd_select(const double *U, const double *WI, const double *theta, double *wflag,
                               int *list, int *counter, double wThresh, const Grid &g, hipStream_t s);
  static hipError_t mhd_weno(const double *U, double *FS, double *SRC, const Grid &g, const double *FT,
                             const double *DT, const double *CT, const int *list, const int *counter,
                             double *thetaNew, double *spd, const double *props, const double *WC, hipStream_t s);
  static hipError_t mhd_pp(const double *U, double *FS, double *SRC, double *thetaNew, double *spd,
                           const double *props, const double *wflag, const Grid &g, const int *list,
                           const int *count, hipStream_t s);
  // ---- learned troubled-cell indicator
  // nv 5 (Euler; eint = the limiter's internal-energy plane) or 9 (GLM-MHD;
  // the internal energy comes from the conserved variables)
  static hipError_t learned_select(int nv, const double *U, const double *eint, const double *theta,
                                   const double *nadFlag, const float *P, float thr, double *wflag, int *list,
                                   int *counter, const Grid &g, int *pos, int cap, hipStream_t s);
  static hipError_t features(int nv, const double *U, const double *eint, float *X, const Grid &g, hipStream_t s);
  static hipError_t labels(int nv, const double *U, const double *FSL, const double *FSW, float tau, float *Y,
                           const Grid &g, hipStream_t s);
  static hipError_t logits(const float *X, const float *P, float *Z, int n, hipStream_t s);
  static hipError_t train_grad(const float *X, const float *Y, const int *idx, int n, int nb, uint64_t seed,
                               const float *P, float posW, float *part, float *lossPart, hipStream_t s);
  static hipError_t nad(int nv, const double *Uold, const double *Unew, double fr, double fv, double *nadFlag,
                        int *count, const Grid &g, hipStream_t s);
  static hipError_t flow_stats(int nv, const double *U, const Grid &g, double ifac, double *out, hipStream_t s);
  // one-dimensional: the kernel renumbers the bricks; shape: the brick shape (2D: 0-2, 3D: 0)
  static dim3 flux_grid(const Grid &g, int shape);
  static hipError_t flux_rk(int nv, const double *U, const double *FS, const Grid &g, const RKArgs &a,
                            hipStream_t s, int shape);
};

extern template struct Launch<2, 2>;
extern template struct Launch<2, 3>;
extern template struct Launch<3, 2>;
extern template struct Launch<3, 3>;

}  // namespace kfvm
