// Host-callable launchers of the stage kernels and the plain-data structs they
// take by value.  No torch or framework types: raw device pointers, sizes and a
// HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace kfvm {

// Grid and physics constants passed by value to every stage kernel.
struct Grid {
  int nx, ny, nz;    // nz = 1 in 2D
  int ncell;
  int bc[6];         // per face W,E,S,N,B,T: 0 periodic, 1 outflow, 2 reflecting, 3 noslip
  unsigned negmask[6];  // variables negated by the mirror at that face (reflecting / noslip)
  double dx, dy, dz;
  double gamma;
  double wt[3];      // 1D Gauss weights on (-1/2, 1/2)
  double vel[3];     // linear advection velocity
  int haveVel;       // 0: velocity (cos gamma, sin gamma, 0) with no z flux
  int rs;            // 0 LLF, 1 HLL, 2 HLLC, 3 exact upwind, 4 KEPES (GLM-MHD)
};

// Face-state storage FS: cells in blocks of FSB consecutive cells, each block
// holding all np planes of its cells contiguously, so that the states of a
// brick of cells sit in a few pages (address translation stays cheap):
// plane p of cell c is FS[fs_base(c, np) + p * FSB].
constexpr int FSB = 64;
__host__ __device__ inline size_t fs_base(size_t c, int np) {
  return (c / FSB) * size_t(np) * FSB + (c % FSB);
}
inline size_t fs_elems(size_t ncell, int np) { return (ncell + FSB - 1) / FSB * FSB * size_t(np); }

// Surface planes of the fused flux path: the reconstructed states of every
// brick's outer faces, one plane family per axis a, planes at the brick
// boundaries (plane p lies below brick p along a; p = 0 .. nb_a).  Side 0 holds
// the state of the cell below the plane (on its upper face), side 1 that of the
// cell above it (on its lower face).  Value (point q, variable v) of the cell
// with transverse index t sits at
//   SB[off[a] + ((p * 2 + side) * (Q * 5) + q * 5 + v) * T[a] + t]
// with t = k ny + j (a = x), k nx + i (a = y), j nx + i (a = z), so that one
// store or load of a wave row covers consecutive cells.
struct Surf {
  long long off[3];
  int T[3];
  int B[3];  // brick extent along each axis
};
// Surface planes for a grid of bricks bx x by x bz holding q face points;
// returns the number of doubles and fills sb.
inline size_t surf_layout(int dim, const Grid &g, int bx, int by, int bz, int q, Surf &sb) {
  const int n[3] = {g.nx, g.ny, g.nz}, b[3] = {bx, by, bz};
  const int T[3] = {g.ny * g.nz, g.nx * g.nz, g.nx * g.ny};
  size_t off = 0;
  for (int a = 0; a < 3; a++) {
    sb.off[a] = (long long)off;
    sb.T[a] = T[a];
    sb.B[a] = b[a];
    if (a < dim) off += size_t((n[a] + b[a] - 1) / b[a] + 1) * 2 * size_t(q) * 5 * size_t(T[a]);
  }
  return off;
}
// Index of (face f of cell (i,j,k), point q, variable v) in SB.
template <int D, int Q>
__host__ __device__ inline size_t surf_index(const Surf &sb, const Grid &g, int f, int i, int j, int k, int q,
                                             int v) {
  const int a = f >> 1, hi = f & 1;
  const int c = a == 0 ? i : (a == 1 ? j : k);
  const int p = c / sb.B[a] + hi;
  const int t = a == 0 ? k * g.ny + j : (a == 1 ? k * g.nx + i : j * g.nx + i);
  return size_t(sb.off[a]) + (size_t(p * 2 + (hi ? 0 : 1)) * (Q * 5) + q * 5 + v) * size_t(sb.T[a]) + t;
}

// Scratch memory of the entry point that allocates nothing: a file-scope
// device array of this many bytes per GPU (noalloc.hip).  noalloc_scratch()
// returns its device address on the current device, taken with
// hipGetSymbolAddress on the first call per device and kept.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
void *noalloc_scratch();

// Learned WENO selection run by the Euler linear pass on its staged brick:
// network parameters and the selection's inputs and outputs (those of
// launch_learned_select).
struct LearnArgs {
  const double *eint = nullptr;     // per-cell internal energy (the limiter's plane)
  const double *theta = nullptr;    // positivity theta of the previous RHS
  const double *nadFlag = nullptr;  // cells that failed the last admissibility check
  const float *P = nullptr;         // network parameters (bfloat16 values)
  float thr = 0.0f;                 // logit threshold
  double *wflag = nullptr;
  int *list = nullptr, *counter = nullptr, *pos = nullptr;
  int cap = 0;
};

enum RKMode : int { RK_PRIME = 0, RK_STAGE = 1, RK_FSAL = 2 };

// Register-update epilogue of the flux kernel.
struct RKArgs {
  int mode;
  double g1, g2, g3, delta, betaDt, bhatDt;
  double atol, rtol;
  double *K;            // RHS register (written in PRIME and FSAL modes)
  const double *Uprev;  // step start state
  double *Uhat;         // embedded solution
  double *Utmp;         // 3S* register or nullptr
  double *Unext;        // next stage state (STAGE mode)
  double *errPartial;   // per-block error sums (FSAL mode)
  double *posPartial;   // per-block positivity minima (FSAL mode)
  unsigned long long *vmaxBits;  // running max wave speed (bit pattern of a non-negative double)
  const unsigned long long *chBits;  // GLM-MHD: {max(c_f + |u|), max |u|} of this evaluation
  // GLM-MHD: limited states at the interior points [(p * 9 + v)][ncell], this
  // evaluation's theta and the differentiation matrix [NQ][NQ + 2]: the flux
  // kernel forms and integrates the source terms
  const double *src, *theta, *Dm;
  // 2D batch of independent problems (slices): per-slice step sizes; the
  // epilogue then uses beta * sliceDt[s] and bhat * sliceDt[s] and keeps one
  // running max wave speed per slice (vmaxBits[s])
  const double *sliceDt;
  double beta, bhat;
  // PRIME / FSAL modes of a batch: write K and Uhat only for slices with
  // sliceWrite[s] != 0 (the others keep the values of their accepted step)
  const int *sliceWrite;
  // Euler sparse WENO: per-cell row in the compact WENO staging (-1: the
  // cell's states are in FS) and that staging, [t][F*Q*5]
  const int *pos;
  const double *ws;
};

// Euler: per-cell internal energy and sound speed, props [2][ncell], read by
// the limiter bounds of launch_linear and launch_weno_post.
hipError_t launch_cell_props(int dim, const double *U, double *props, const Grid &g, hipStream_t s);
// Linear reconstruction, KXRCF indicator and positivity limiter (all cells).
// W0: full-stencil face weights [F*Q][C]; W0T: the same regrouped per pass
// of linear_pass_points(dim, nq) points as [F*Q/P][C][P].  Euler: cells
// whose states violate a limiter bound are appended to ppList (count in
// *ppCount, zeroed by the caller) and limited by launch_pp_lanes.
// la (Euler, may be null): the learned WENO selection runs in the same kernel
// on the staged brick (outputs those of launch_learned_select).
hipError_t launch_linear(int dim, int nq, int nv, const double *U, double *FS, double *WI,
                         double *thetaNew, const double *props, const Grid &g, const double *W0,
                         const double *W0T, int *ppList, int *ppCount, hipStream_t s,
                         const LearnArgs *la = nullptr);
// Euler: the limiter's exact sequence on the cells of launch_linear whose
// states failed the fast test (list, size *count); those wflag marks for
// WENO are skipped.
hipError_t launch_pp_lanes(int dim, int nq, const double *U, double *FS, double *thetaNew,
                           const double *props, const double *wflag, const Grid &g,
                           const int *list, const int *count, hipStream_t s);
// ---- fused path of the Euler equations in sparse mode.  Rows: the listed
// cells' limited face states, row t = [F*Q*5] (WENO cells at their list
// positions, then the cells of the limiter's exact sequence), pos[cell] = the
// cell's row or -1.
// Surface planes for the grid (doubles; fills sb).
size_t surf_elems(int dim, int nq, const Grid &g, Surf &sb);
// Linear reconstruction with the fast limiter test (slow cells to ppList, the
// others get theta 1) and the KXRCF indicator (WI may be null), storing only
// the states on brick-boundary faces, into SB.
hipError_t launch_linear_surf(int dim, int nq, const double *U, double *WI, double *thetaNew,
                              const double *props, const Grid &g, const double *W0T, int *ppList, int *ppCount,
                              double *SB, const Surf &sb, hipStream_t s);
// The limiter's exact sequence on the slow cells of launch_linear_surf not
// marked for WENO: recomputed linear states, limited, into rows
// *wenoCount + 0, 1, ... (counted in *rowCount, zeroed by the caller).
hipError_t launch_pp_rows(int dim, int nq, const double *U, double *rows, int *pos, double *thetaNew,
                          const double *props, const double *wflag, const Grid &g, const double *W0,
                          const int *list, const int *count, const int *wenoCount, int *rowCount, hipStream_t s);
// Linear states of the cells without a row (recomputed, weights WQ
// [Q][C][F]), rows of the others, SB across brick boundaries; Riemann solves,
// divergence and the RK epilogue of launch_flux_rk.
hipError_t launch_flux_fused(int dim, int nq, const double *U, const double *WQ, const double *rows, const int *pos,
                             const double *SB, const Surf &sb, const Grid &g, const RKArgs &a, hipStream_t s);
// pos (may be null): per-cell list position below cap, -1 otherwise
hipError_t launch_select(int dim, int nq, const double *U, const double *WI, const double *theta,
                         double *wflag, int *list, int *counter, double wThresh, const Grid &g,
                         int *pos, int cap, hipStream_t s);
// KFVM-WENO reconstruction of the listed cells (list == nullptr: all).  Euler
// leaves the face states in linearised primitive variables for
// launch_weno_post; linear advection completes here.
// FW, DW: reference weight layouts (linear advection).  Euler: FWS is the
// compact table of the listed substencil rows only, [SUB][F*Q] face weights
// then [SUB][I] derivative weights (SUB = C + Cc + B*Cb, substencils back to
// back; compact_weno_weights), DWS is not read.  WS (Euler, may be null): compact staging
// [(F*Q*5)][cap] for the first cap listed cells, shared with launch_weno_post.
// Compact WENO weight table (see launch_weno) from the reference layouts
// face [nSub][F*Q][C] and deriv [nSub][I][C]; out has SUB*(F*Q + I) doubles,
// then, with nCell > 0 (GLM-MHD), SUB*nCell rows of the interior-point
// weights cell [nSub][nCell][C].
void compact_weno_weights(int dim, int fq, int nI, const double *face, const double *deriv, double *out,
                          int nCell = 0, const double *cell = nullptr);
hipError_t launch_weno(int dim, int nq, int nv, const double *U, double *FS, double *thetaNew,
                       const Grid &g, const double *FW, const double *DW, const double *FWS,
                       const double *DWS, const int *list, const int *counter, double *WS, int cap,
                       hipStream_t s);
// Euler: back to conserved variables and the positivity limiter on the
// listed cells (in WS for the first cap of them, then copied to FS).
hipError_t launch_weno_post(int dim, int nq, const double *U, double *FS, double *thetaNew,
                            const double *props, const Grid &g, const int *list, const int *counter,
                            double *WS, int cap, hipStream_t s);
// GLM-MHD (9 variables).  SRC [(p * 9 + v)][ncell]: states, then source
// terms, at the NQ^D interior points; props [2][ncell]: per-cell internal
// energy and sound speed (launch_mhd_props, before the linear pass or WENO);
// spd [2][ncell]: per-cell maxima of c_f + |u| and |u| over the cell's limited
// face states, which launch_mhd_chmax folds into chBits for the flux kernel's
// cleaning speed.
hipError_t launch_mhd_props(int dim, const double *U, double *props, const Grid &g, hipStream_t s);
// WT: full-stencil weights per pass, the face passes [F*Q/P][C][P]
// (P = linear_pass_points) then the interior passes [NQ^D/PI][C][PI]
// (PI = mhd_interior_pass_points).  Cells whose states violate a limiter
// bound are appended to ppList (count in *ppCount, zeroed by the caller) for
// launch_mhd_pp; the others get theta = 1.  WI may be null (learned selection).
hipError_t launch_mhd_linear(int dim, int nq, const double *U, double *FS, double *SRC, double *WI,
                             double *thetaNew, double *spd, const double *props, const Grid &g, const double *WT,
                             int *ppList, int *ppCount, hipStream_t s);
hipError_t launch_mhd_select(int dim, int nq, const double *U, const double *WI, const double *theta,
                             double *wflag, int *list, int *counter, double wThresh, const Grid &g,
                             hipStream_t s);
// FT, DT, CT: stencil-major face [nSub][C][F*Q], smoothness [nSub][C][I] and
// interior [nSub][C][NQ^D] weights; WC: the compact table of the same weights
// (compact_weno_weights with the interior rows), read instead of them where it
// fits in LDS.  The listed cells (list null: every cell) are converted back
// and limited (theta, spd).
hipError_t launch_mhd_weno(int dim, int nq, const double *U, double *FS, double *SRC, const Grid &g,
                           const double *FT, const double *DT, const double *CT, const int *list,
                           const int *counter, double *thetaNew, double *spd, const double *props,
                           const double *WC, hipStream_t s);
// The limiter's exact sequence on the listed cells whose WENO flag is 0.
hipError_t launch_mhd_pp(int dim, int nq, const double *U, double *FS, double *SRC, double *thetaNew,
                         double *spd, const double *props, const double *wflag, const Grid &g, const int *list,
                         const int *count, hipStream_t s);
// Flow statistics of U (Euler nv 5, GLM-MHD nv 9): per-block partial sums
// out[3 b + s] (s: kinetic energy, enstrophy, magnetic helicity) over
// stage_blocks(ncell) blocks, each cell weighted by ifac = dvol / volume.
hipError_t launch_flow_stats(int dim, int nv, const double *U, const Grid &g, double ifac, double *out,
                             hipStream_t s);
// Cleaning speed: the per-cell maxima spd folded into chBits (zeroed by the
// caller).  The source terms are formed inside launch_flux_rk (RKArgs src,
// theta, Dm).
hipError_t launch_mhd_chmax(int dim, const double *spd, const Grid &g, unsigned long long *chBits, hipStream_t s);
// shape: the flux kernel's brick shape (2D: 0 = 16x16, 1 = 32x8, 2 = 64x4;
// 3D: 0 = 16x4x4); the partial count follows it.
hipError_t launch_flux_rk(int dim, int nq, int nv, const double *U, const double *FS,
                          const Grid &g, const RKArgs &a, hipStream_t s, int shape = 0);
int flux_rk_partials(int dim, const Grid &g, int shape = 0);
// Batch of slices: per-slice finalize (out[2 s], out[2 s + 1]) over the flux
// kernel's blocks, which are slice-major (blocksPerSlice each), and the copy of
// the accepted slices' states (mask[s] != 0) from U to Uprev.
hipError_t launch_finalize_slices(const double *errPartial, const double *posPartial, int blocksPerSlice,
                                  int nslice, double *out, hipStream_t s);
hipError_t launch_commit_slices(const double *U, double *Uprev, const int *mask, int nv, int ncell,
                                int cellsPerSlice, hipStream_t s);
hipError_t launch_stage_pre_slices(int nv, const double *Uprev, const double *K, double *U, double *Uhat,
                                   double *Utmp, const double *sliceDt, double beta, double bhat, int ncell,
                                   int cellsPerSlice, hipStream_t s);
hipError_t launch_stage_pre(int nv, const double *Uprev, const double *K, double *U, double *Uhat,
                            double *Utmp, double betaDt, double bhatDt, int ncell, hipStream_t s);
hipError_t launch_finalize(const double *errPartial, const double *posPartial, int nPart,
                           double *out, hipStream_t s);
hipError_t launch_fill(double *p, double v, size_t n, hipStream_t s);
// Zero `bytes` (a multiple of 4) with a kernel, not a memset (graph capture).
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);

// ---- learned troubled-cell indicator (Euler nv 5, GLM-MHD nv 9); P = network parameters
int net_inputs(int dim);  // features per cell
int net_params(int dim);  // floats in the parameter vector
int net_train_block();    // samples per training block
// eint: per-cell internal energy (the first plane of the limiter's cell
// properties) for nv 5; unused for nv 9
hipError_t launch_learned_select(int dim, int nv, const double *U, const double *eint, const double *theta,
                                 const double *nadFlag, const float *P, float thr, double *wflag, int *list,
                                 int *counter, const Grid &g, int *pos, int cap, hipStream_t s);
hipError_t launch_features(int dim, int nv, const double *U, const double *eint, float *X, const Grid &g,
                           hipStream_t s);
hipError_t launch_labels(int dim, int nq, int nv, const double *U, const double *FSL, const double *FSW, float tau,
                         float *Y, const Grid &g, hipStream_t s);
hipError_t launch_logits(int dim, const float *X, const float *P, float *Z, int n, hipStream_t s);
hipError_t launch_train_grad(int dim, const float *X, const float *Y, const int *idx, int n, int nb,
                             uint64_t seed, const float *P, float posW, float *part, float *lossPart,
                             hipStream_t s);
hipError_t launch_train_reduce(const float *part, const float *lossPart, int nblk, int np, int nb, float *grad,
                               float *loss, hipStream_t s);
hipError_t launch_adam(float *M, float *Pbf, const float *grad, float *m1, float *m2, int np, float lr,
                       float b1, float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t launch_nad(int dim, int nv, const double *Uold, const double *Unew, double fr, double fv,
                      double *nadFlag, int *count, const Grid &g, hipStream_t s);

}  // namespace kfvm
