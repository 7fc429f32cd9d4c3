// Run-time selection of the kernel instantiation and the shape-independent
// small kernels (RK stage start, reductions, fills).
#define KFVM_SMALL_KERNELS
#include "launch_decl.hpp"
#include "stage_kernels.hpp"

namespace kfvm {

#define KFVM_SELECT(dim, nq, CALL)                          \
  do {                                                      \
    if (dim == 3 && nq == 2) { using L = Launch<3, 2>; return CALL; } \
    if (dim == 3 && nq == 3) { using L = Launch<3, 3>; return CALL; } \
    if (dim == 2 && nq == 2) { using L = Launch<2, 2>; return CALL; } \
    if (dim == 2 && nq == 3) { using L = Launch<2, 3>; return CALL; } \
    return hipErrorInvalidValue;                            \
  } while (0)

hipError_t launch_cell_props(int dim, const double *U, double *props, const Grid &g, hipStream_t s) {
  KFVM_SELECT(dim, 2, L::cell_props(U, props, g, s));
}

hipError_t launch_linear(int dim, int nq, int nv, const double *U, double *FS, double *WI,
                         double *thetaNew, const double *props, const Grid &g, const double *W0,
                         const double *W0T, int *ppList, int *ppCount, hipStream_t s, const LearnArgs *la) {
  if (nv != 5 && nv != 1) return hipErrorInvalidValue;
  if (la && nv != 5) return hipErrorInvalidValue;
  KFVM_SELECT(dim, nq, L::linear(nv, U, FS, WI, thetaNew, props, g, W0, W0T, ppList, ppCount, s, la));
}

hipError_t launch_linear_surf(int dim, int nq, const double *U, double *WI, double *thetaNew,
                              const double *props, const Grid &g, const double *W0T, int *ppList, int *ppCount,
                              double *SB, const Surf &sb, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::linear_surf(U, WI, thetaNew, props, g, W0T, ppList, ppCount, SB, sb, s));
}
hipError_t launch_pp_rows(int dim, int nq, const double *U, double *rows, int *pos, double *thetaNew,
                          const double *props, const double *wflag, const Grid &g, const double *W0,
                          const int *list, const int *count, const int *wenoCount, int *rowCount, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::pp_rows(U, rows, pos, thetaNew, props, wflag, g, W0, list, count, wenoCount, rowCount, s));
}
hipError_t launch_flux_fused(int dim, int nq, const double *U, const double *WQ, const double *rows, const int *pos,
                             const double *SB, const Surf &sb, const Grid &g, const RKArgs &a, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::flux_fused(U, WQ, rows, pos, SB, sb, g, a, s));
}
size_t surf_elems(int dim, int nq, const Grid &g, Surf &sb) {
  const int q = dim == 3 ? nq * nq : nq;
  return dim == 3 ? surf_layout(3, g, dev::RBox<3>::BX, dev::RBox<3>::BY, dev::RBox<3>::BZ, q, sb)
                  : surf_layout(2, g, dev::RBox<2>::BX, dev::RBox<2>::BY, dev::RBox<2>::BZ, q, sb);
}

hipError_t launch_pp_lanes(int dim, int nq, const double *U, double *FS, double *thetaNew,
                           const double *props, const double *wflag, const Grid &g,
                           const int *list, const int *count, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::pp_lanes(U, FS, thetaNew, props, wflag, g, list, count, s));
}

hipError_t launch_select(int dim, int nq, const double *U, const double *WI, const double *theta,
                         double *wflag, int *list, int *counter, double wThresh, const Grid &g,
                         int *pos, int cap, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::select(U, WI, theta, wflag, list, counter, wThresh, g, pos, cap, s));
}

void compact_weno_weights(int dim, int fq, int nI, const double *face, const double *deriv, double *out,
                          int nCell, const double *cell) {
  const int C = dim == 3 ? SInfo<3>::C : SInfo<2>::C, Cc = dim == 3 ? SInfo<3>::Cc : SInfo<2>::Cc;
  const int Cb = dim == 3 ? SInfo<3>::Cb : SInfo<2>::Cb, nSub = dim == 3 ? SInfo<3>::nSub : SInfo<2>::nSub;
  const int SUB = C + Cc + (nSub - 2) * Cb;
  int row = 0;
  for (int sb = 0; sb < nSub; sb++) {
    const int ns = sb == 0 ? C : (sb == 1 ? Cc : Cb);
    for (int m = 0; m < ns; m++, row++) {
      for (int p = 0; p < fq; p++) out[size_t(row) * fq + p] = face[(size_t(sb) * fq + p) * C + m];
      for (int i = 0; i < nI; i++) out[size_t(SUB) * fq + size_t(row) * nI + i] = deriv[(size_t(sb) * nI + i) * C + m];
      for (int p = 0; p < nCell; p++)
        out[size_t(SUB) * (fq + nI) + size_t(row) * nCell + p] = cell[(size_t(sb) * nCell + p) * C + m];
    }
  }
}

hipError_t launch_weno(int dim, int nq, int nv, const double *U, double *FS, double *thetaNew,
                       const Grid &g, const double *FW, const double *DW, const double *FWS,
                       const double *DWS, const int *list, const int *counter, double *WS, int cap,
                       hipStream_t s) {
  if (nv != 5 && nv != 1) return hipErrorInvalidValue;
  if (WS && cap < 0) return hipErrorInvalidValue;
  (void)DWS;
  KFVM_SELECT(dim, nq, L::weno(nv, U, FS, thetaNew, g, FW, DW, list, counter, WS, cap, nv == 5 ? FWS : nullptr, s));
}

hipError_t launch_weno_post(int dim, int nq, const double *U, double *FS, double *thetaNew,
                            const double *props, const Grid &g, const int *list, const int *counter,
                            double *WS, int cap, hipStream_t s) {
  if (WS && cap < 0) return hipErrorInvalidValue;
  KFVM_SELECT(dim, nq, L::weno_post(U, FS, thetaNew, props, g, list, counter, WS, cap, s));
}

hipError_t launch_mhd_props(int dim, const double *U, double *props, const Grid &g, hipStream_t s) {
  KFVM_SELECT(dim, 2, L::mhd_props(U, props, g, s));
}
hipError_t launch_mhd_linear(int dim, int nq, const double *U, double *FS, double *SRC, double *WI,
                             double *thetaNew, double *spd, const double *props, const Grid &g, const double *WT,
                             int *ppList, int *ppCount, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::mhd_linear(U, FS, SRC, WI, thetaNew, spd, props, g, WT, ppList, ppCount, s));
}
hipError_t launch_mhd_select(int dim, int nq, const double *U, const double *WI, const double *theta,
                             double *wflag, int *list, int *counter, double wThresh, const Grid &g,
                             hipStream_t s) {
  KFVM_SELECT(dim, nq, L::mhd_select(U, WI, theta, wflag, list, counter, wThresh, g, s));
}
hipError_t launch_mhd_weno(int dim, int nq, const double *U, double *FS, double *SRC, const Grid &g,
                           const double *FT, const double *DT, const double *CT, const int *list,
                           const int *counter, double *thetaNew, double *spd, const double *props,
                           const double *WC, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::mhd_weno(U, FS, SRC, g, FT, DT, CT, list, counter, thetaNew, spd, props, WC, s));
}
hipError_t launch_mhd_pp(int dim, int nq, const double *U, double *FS, double *SRC, double *thetaNew,
                         double *spd, const double *props, const double *wflag, const Grid &g, const int *list,
                         const int *count, hipStream_t s) {
  KFVM_SELECT(dim, nq, L::mhd_pp(U, FS, SRC, thetaNew, spd, props, wflag, g, list, count, s));
}
hipError_t launch_flow_stats(int dim, int nv, const double *U, const Grid &g, double ifac, double *out,
                            hipStream_t s) {
  if (nv != 5 && nv != 9) return hipErrorInvalidValue;
  KFVM_SELECT(dim, 2, L::flow_stats(nv, U, g, ifac, out, s));
}
hipError_t launch_mhd_chmax(int dim, const double *spd, const Grid &g, unsigned long long *chBits, hipStream_t s) {
  (void)dim;
  const unsigned nb = std::min<unsigned>(stage_blocks(g.ncell), 2048u);
  hipLaunchKernelGGL(dev::k_mhd_chmax, dim3(nb), dim3(dev::BLK), 0, s, spd, int64_t(g.ncell), chBits);
  return hipGetLastError();
}

hipError_t launch_flux_rk(int dim, int nq, int nv, const double *U, const double *FS,
                          const Grid &g, const RKArgs &a, hipStream_t s, int shape) {
  if (nv != 5 && nv != 1 && nv != 9) return hipErrorInvalidValue;
  if (shape < 0 || shape > (dim == 2 ? 2 : 0)) return hipErrorInvalidValue;
  KFVM_SELECT(dim, nq, L::flux_rk(nv, U, FS, g, a, s, shape));
}

int flux_rk_partials(int dim, const Grid &g, int shape) {
  const dim3 gr = dim == 3 ? Launch<3, 2>::flux_grid(g, 0) : Launch<2, 2>::flux_grid(g, shape);
  return int(gr.x * gr.y * gr.z);
}

hipError_t launch_stage_pre(int nv, const double *Uprev, const double *K, double *U, double *Uhat,
                            double *Utmp, double betaDt, double bhatDt, int ncell, hipStream_t s) {
  if (nv == 5)
    hipLaunchKernelGGL((dev::k_stage_pre<5>), dim3(stage_blocks(ncell)), dim3(dev::BLK), 0, s, Uprev,
                       K, U, Uhat, Utmp, betaDt, bhatDt, ncell);
  else if (nv == 1)
    hipLaunchKernelGGL((dev::k_stage_pre<1>), dim3(stage_blocks(ncell)), dim3(dev::BLK), 0, s, Uprev,
                       K, U, Uhat, Utmp, betaDt, bhatDt, ncell);
  else if (nv == 9)
    hipLaunchKernelGGL((dev::k_stage_pre<9>), dim3(stage_blocks(ncell)), dim3(dev::BLK), 0, s, Uprev,
                       K, U, Uhat, Utmp, betaDt, bhatDt, ncell);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_stage_pre_slices(int nv, const double *Uprev, const double *K, double *U, double *Uhat,
                                   double *Utmp, const double *sliceDt, double beta, double bhat, int ncell,
                                   int cellsPerSlice, hipStream_t s) {
  if (nv == 5)
    hipLaunchKernelGGL((dev::k_stage_pre_slices<5>), dim3(stage_blocks(ncell)), dim3(dev::BLK), 0, s, Uprev, K, U,
                       Uhat, Utmp, sliceDt, beta, bhat, ncell, cellsPerSlice);
  else if (nv == 1)
    hipLaunchKernelGGL((dev::k_stage_pre_slices<1>), dim3(stage_blocks(ncell)), dim3(dev::BLK), 0, s, Uprev, K, U,
                       Uhat, Utmp, sliceDt, beta, bhat, ncell, cellsPerSlice);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_commit_slices(const double *U, double *Uprev, const int *mask, int nv, int ncell,
                                int cellsPerSlice, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_commit_slices, dim3(stage_blocks(ncell)), dim3(dev::BLK), 0, s, U, Uprev, mask, nv,
                     ncell, cellsPerSlice);
  return hipGetLastError();
}

hipError_t launch_finalize_slices(const double *errPartial, const double *posPartial, int blocksPerSlice,
                                  int nslice, double *out, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_finalize_slices, dim3(nslice), dim3(dev::BLK), 0, s, errPartial, posPartial,
                     blocksPerSlice, out);
  return hipGetLastError();
}

hipError_t launch_finalize(const double *errPartial, const double *posPartial, int nPart,
                           double *out, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_finalize, dim3(1), dim3(1024), 0, s, errPartial, posPartial, nPart, out);
  return hipGetLastError();
}

int net_inputs(int dim) { return dim == 3 ? dev::Net<3>::NIN : dev::Net<2>::NIN; }
int net_params(int dim) { return dim == 3 ? dev::Net<3>::NP : dev::Net<2>::NP; }
int net_train_block() { return dev::TB; }

hipError_t launch_learned_select(int dim, int nv, const double *U, const double *eint, const double *theta,
                                 const double *nadFlag, const float *P, float thr, double *wflag, int *list,
                                 int *counter, const Grid &g, int *pos, int cap, hipStream_t s) {
  if (nv != 5 && nv != 9) return hipErrorInvalidValue;
  KFVM_SELECT(dim, 2, L::learned_select(nv, U, eint, theta, nadFlag, P, thr, wflag, list, counter, g, pos, cap, s));
}
hipError_t launch_features(int dim, int nv, const double *U, const double *eint, float *X, const Grid &g,
                           hipStream_t s) {
  if (nv != 5 && nv != 9) return hipErrorInvalidValue;
  KFVM_SELECT(dim, 2, L::features(nv, U, eint, X, g, s));
}
hipError_t launch_labels(int dim, int nq, int nv, const double *U, const double *FSL, const double *FSW, float tau,
                         float *Y, const Grid &g, hipStream_t s) {
  if (nv != 5 && nv != 9) return hipErrorInvalidValue;
  KFVM_SELECT(dim, nq, L::labels(nv, U, FSL, FSW, tau, Y, g, s));
}
hipError_t launch_logits(int dim, const float *X, const float *P, float *Z, int n, hipStream_t s) {
  KFVM_SELECT(dim, 2, L::logits(X, P, Z, n, s));
}
hipError_t launch_train_grad(int dim, const float *X, const float *Y, const int *idx, int n, int nb,
                             uint64_t seed, const float *P, float posW, float *part, float *lossPart,
                             hipStream_t s) {
  KFVM_SELECT(dim, 2, L::train_grad(X, Y, idx, n, nb, seed, P, posW, part, lossPart, s));
}
hipError_t launch_train_reduce(const float *part, const float *lossPart, int nblk, int np, int nb, float *grad,
                               float *loss, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_train_reduce, dim3(stage_blocks(np)), dim3(dev::BLK), 0, s, part, lossPart, nblk, np,
                     nb, grad, loss);
  return hipGetLastError();
}
hipError_t launch_adam(float *M, float *Pbf, const float *grad, float *m1, float *m2, int np, float lr,
                       float b1, float b2, float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_adam, dim3(stage_blocks(np)), dim3(dev::BLK), 0, s, M, Pbf, grad, m1, m2, np, lr, b1,
                     b2, eps, c1, c2);
  return hipGetLastError();
}
hipError_t launch_nad(int dim, int nv, const double *Uold, const double *Unew, double fr, double fv,
                      double *nadFlag, int *count, const Grid &g, hipStream_t s) {
  if (nv != 5 && nv != 9) return hipErrorInvalidValue;
  KFVM_SELECT(dim, 2, L::nad(nv, Uold, Unew, fr, fv, nadFlag, count, g, s));
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s) {
  if (bytes % 4) return hipErrorInvalidValue;
  const size_t n = bytes / 4;
  const unsigned nb = unsigned(std::min<size_t>((n + dev::BLK - 1) / dev::BLK, 1024));
  hipLaunchKernelGGL(dev::k_zero32, dim3(nb ? nb : 1), dim3(dev::BLK), 0, s, static_cast<unsigned *>(p), n);
  return hipGetLastError();
}

hipError_t launch_fill(double *p, double v, size_t n, hipStream_t s) {
  const unsigned nb = unsigned(std::min<size_t>((n + dev::BLK - 1) / dev::BLK, 4096));
  hipLaunchKernelGGL(dev::k_fill, dim3(nb ? nb : 1), dim3(dev::BLK), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace kfvm
