/* kfvm.h -- C ABI of the MI355X KFVM-WENO stage library (libkfvm_amd.so).
 *
 * The reference exposes this path through the C++ class KFVM::Solver
 * (src/Solver.H:54-160), constructed from a ProblemSetup parsed from an init
 * file (src/ProblemSetup.H:142-205) and driven by src/main.cpp:23-55.  It has
 * no FFI of its own; each entry point below replaces one member of that
 * surface (cited per function).  Conventions:
 *   - every function returns 0 on success and a negative code on failure;
 *     kfvm_last_error() then describes the failure (thread-local string);
 *   - integers are int64_t, reals are double; arrays are host memory in the
 *     reference's plot/checkpoint order [NV][nz][ny][nx] over interior cells;
 *   - one solver owns one HIP stream on the current device; calls on one
 *     solver must be serialised, different solvers may run concurrently.
 */
#ifndef KFVM_H
#define KFVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kfvm_solver kfvm_solver;

/* Step outcome codes (TSStatus, src/Types.H:43, plus the two terminal
 * conditions Solver::TakeStep reports, src/Solver.cpp:332-340). */
enum {
  KFVM_ACCEPTED = 0,
  KFVM_TOLERANCE = 1,
  KFVM_UNPHYSICAL = 2,
  KFVM_REJECT_LIMIT = 3,
  KFVM_STAGNATED = 4
};

typedef struct kfvm_info {
  int64_t dim, nq, nv, nx, ny, nz, ncell;
  int64_t nsub, nfaces, nqface, ncells_stencil, nindic; /* weight-table shape */
  int64_t rk_stages;
  double dx, dy, dz, gamma;
} kfvm_info;

/* Solver::Solver(ProblemSetup&) after ProblemSetup::setFromFile
 * (src/Solver.cpp:36-81, src/ProblemSetup.H:142-205): parse the init file
 * (reals with std::stof semantics), build the GP weight tables, fill the
 * initial condition of the registry problem named by problem:name, prime the
 * FSAL register and choose dt0 = 0.01 cfl dmin / vmax.  The reference
 * compiles its problem in, so its init files carry no problem:name: add that
 * line (a registry problem of a reference exec, e.g. riem_3d_1c, also sets
 * that exec's dimension, quadrature, equation, Riemann solver, RK table and
 * boundary conditions).  An init file without it fails with a message that
 * lists the registry.  Skipped lines are reported on stderr. */
int kfvm_create(const char *init_path, kfvm_solver **out);
/* Same, from the text of an init file. */
int kfvm_create_from_text(const char *init_text, kfvm_solver **out);
int kfvm_destroy(kfvm_solver *s);
/* A batch of n independent 2D problems of one configuration (same mesh and
 * scheme), problem k with seed seeds[k], advanced together: every kernel of a
 * stage covers all of them in one launch, each problem keeps its own step size
 * and controller (Solver::TakeStep per problem).  State arrays are
 * [NV][n][ny][nx].  No counterpart in the reference (one problem per run). */
int kfvm_create_batch(const char *init_text, const uint64_t *seeds, int64_t n, kfvm_solver **out);
/* Time and next step size of every problem of a batch (n = its size). */
int kfvm_get_slice_times(kfvm_solver *s, double *t, double *dt, int64_t n);

/* Solver::Solve (src/Solver.cpp:84-140): step until final time / maxsteps. */
int kfvm_solve(kfvm_solver *s, int64_t *steps_taken);
/* Solver::TakeStep (src/Solver.cpp:188-343); *status gets a KFVM_* code. */
int kfvm_take_step(kfvm_solver *s, int *status);
/* Stage-granular fixed-dt advance: exactly nstages RK stages (RHS evaluation
 * plus register update, Solver.cpp:225-258) of steps of size dt, queued on
 * the solver's stream without host synchronisation; every completed step is
 * accepted.  No reference counterpart: it is the benchmark driver of the
 * per-stage update and follows TakeStep's stage sequence. */
int kfvm_advance_stages(kfvm_solver *s, int64_t nstages, double dt);
/* Solver::evalRHS(ConsDataView, Real t) (src/Solver.H:129,
 * src/Solver.cpp:410-530) on the current solution (which = 0) or the
 * step-start solution (which = 1) at time t; the result becomes the FSAL
 * register.  *max_vel gets the largest wave speed over all faces.  t is the
 * time the reference hands to time-dependent user boundary conditions and
 * sources; the compiled-in problems have none, so it is only checked to be
 * finite. */
int kfvm_eval_rhs(kfvm_solver *s, int which, double t, double *max_vel);
/* kfvm_eval_rhs without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Evaluates the RHS of the device state u_dev ([NV][nz][ny][nx] doubles, the
 * layout of kfvm_get_state) into rhs_dev (same layout, device memory) at time
 * t, with the solver's configuration, weights, previous positivity theta and
 * WENO mode, and leaves the solver unchanged.  Every per-cell scratch buffer
 * of the evaluation lives in one file-scope device array of 256 MiB per GPU
 * (held until the process ends); grids whose scratch exceeds it are rejected
 * with an error.  Runs on the solver's stream and returns after it; inputs
 * must be complete before the call.  All calls on one GPU share that array:
 * calls that overlap on one GPU race. */
int kfvm_eval_rhs_noalloc(kfvm_solver *s, const double *u_dev, double *rhs_dev, double t, double *max_vel);

/* Solution / RHS / WENO-flag access (U_halo interior, RHS, wenoFlagView). */
int kfvm_get_state(kfvm_solver *s, double *host, int64_t n);
int kfvm_set_state(kfvm_solver *s, const double *host, int64_t n);
int kfvm_get_rhs(kfvm_solver *s, double *host, int64_t n);
/* flags: [2][nz][ny][nx], slot 0 WENO flag, slot 1 positivity theta */
int kfvm_get_flags(kfvm_solver *s, double *host, int64_t n);
int kfvm_set_flags(kfvm_solver *s, const double *host, int64_t n);

/* The solver's time, next step size and step number (Solver::time, dt, nTS;
 * src/Solver.H:63-66); set_time sets time and next step size, on a batch
 * those of every problem. */
int kfvm_get_time(kfvm_solver *s, double *t, double *dt, int64_t *step);
int kfvm_set_time(kfvm_solver *s, double t, double dt);
int kfvm_get_info(kfvm_solver *s, kfvm_info *info);
/* Counters printed at the end of Solver::Solve (src/Solver.cpp:129-137). */
int kfvm_get_counters(kfvm_solver *s, int64_t *rhs_evals, int64_t *rejects_tol,
                      int64_t *rejects_unphys, int64_t *last_nweno);
/* Force the next RHS evaluation to run KFVM-WENO on every cell, as the
 * reference does after an unphysical rejection (src/Solver.cpp:298, 437-443). */
int kfvm_force_full_weno(kfvm_solver *s);

/* Stencil::faceWeights [nsub][2*dim][Q][C] and Stencil::derivWeights
 * [nsub][I][C] (src/stencil/Stencil.H:95-103).  set_weights replaces the
 * tables the stage kernels read (e.g. with a committed table). */
int kfvm_get_weights(kfvm_solver *s, double *face, int64_t nface, double *deriv, int64_t nderiv);
int kfvm_set_weights(kfvm_solver *s, const double *face, int64_t nface, const double *deriv,
                     int64_t nderiv);

/* Plot / checkpoint output (WriterPDI::writePlot / writeCkpt,
 * src/io/WriterPDI.cpp:91-160, datasets of user/pdi_*_xdmf.yml) of the current
 * state under io:datadir / io:basename; path receives the HDF5 file's name.
 * kfvm_solve writes them by itself when io:plotfreq / io:ckptfreq are set, and
 * io:restartfile restarts from a checkpoint (Solver.cpp:60-66).  The HDF5
 * library (1.10 or newer) is loaded at run time; KFVM_HDF5_LIB names it. */
int kfvm_write_plot(kfvm_solver *s, char *path, int64_t len);
int kfvm_write_checkpoint(kfvm_solver *s, char *path, int64_t len);

/* Flow statistics of the current state (Solver::evalFlowStats,
 * src/Solver.cpp:154-186, with FlowStats<Hydro> Hydro_K.H:103-173 and
 * FlowStats<MHD_GLM> MHD_GLM_K.H:108-207): domain averages of the kinetic
 * energy and the enstrophy (and, GLM-MHD, B . curl B), with central
 * differences through the boundary conditions; linear advection has one
 * statistic, always 0.  stats receives up to 3 values, *n their number.
 * kfvm_solve writes them as rows (time, statistics) of
 * <datadir>/R2_NX.._NY..[_NZ..]/<basename>_stats.h5, dataset stats_data
 * [time:maxsteps][1 + n], every io:statsfreq steps, before the first step and
 * after the last (WriterPDI::writeFlowStats, src/io/WriterPDI.cpp:69-86).
 * Single problems only (not batches). */
int kfvm_flow_stats(kfvm_solver *s, double *stats, int *n);

/* Per-kernel device timing with HIP events on the solver's stream.  While
 * enabled, every launch is bracketed by events; kfvm_profile_get returns the
 * accumulated milliseconds and launch count of one kernel family
 * ("linear", "select", "weno", "pp", "flux_rk", "stage_pre"). */
int kfvm_profile_enable(kfvm_solver *s, int on);
int kfvm_profile_get(kfvm_solver *s, const char *name, double *ms, int64_t *count);
int kfvm_profile_reset(kfvm_solver *s);
int kfvm_sync(kfvm_solver *s);

/* ---- learned troubled-cell indicator (Euler and GLM-MHD) ----
 * None of these has a counterpart in the reference: they add a small network
 * (features -> 32 -> 32 -> logit, bfloat16 weights, FP32 sums) that replaces
 * the KXRCF selection (WenoSelect_K, src/numeric/Numeric_K.H:292-398) of which
 * cells get WENO, trained on the GPU on labels the solver generates.  The
 * features are the stencil's density and gas pressure (for GLM-MHD the
 * pressure without the magnetic and cleaning energies). */
typedef struct kfvm_indicator kfvm_indicator;

/* Switch the solver to the learned selection with n bfloat16 parameters
 * (kfvm_indicator_sizes) and logit threshold; n = 0 returns to KXRCF.  While
 * on, each accepted step is checked for numerical admissibility: rho and p of
 * every cell inside its neighbourhood's previous range widened by a tolerance
 * (kfvm_set_nad).  Cells that fail get WENO in every stage of the next step;
 * a step in which more than max_frac of the cells fail is redone with KXRCF. */
int kfvm_set_indicator(kfvm_solver *s, const uint16_t *weights, int64_t n, double threshold);
/* Admissibility check of learned steps: on/off, tolerance
 * delta = max(value_frac |bound|, range_frac (max - min)) and the fraction of
 * failing cells that makes a step be redone; defaults on, 0.05, 1e-3, 0.02. */
int kfvm_set_nad(kfvm_solver *s, int on, double range_frac, double value_frac, double max_frac);
/* Steps redone with KXRCF, violating cells summed over the checked steps. */
int kfvm_get_indicator_counters(kfvm_solver *s, int64_t *nad_fallbacks, int64_t *nad_cells, int *learned);

int kfvm_indicator_create(int dim, kfvm_indicator **out);
int kfvm_indicator_destroy(kfvm_indicator *p);
/* Append the samples of the solver's current state: features of every cell and
 * label 1 where the full-WENO face states (after the limiter) differ from the
 * linear ones by more than tau times the stencil's range of rho or E (a range
 * below 1e-3 of the cell's value counts as that floor). */
int kfvm_indicator_collect(kfvm_indicator *p, kfvm_solver *s, double tau, int64_t *nsamples,
                           int64_t *npositive);
int kfvm_indicator_samples(kfvm_indicator *p, float *X, float *Y, int64_t offset, int64_t n);
int kfvm_indicator_init(kfvm_indicator *p, uint64_t seed);
/* iters Adam steps (lr, betas 0.9/0.999) on batches drawn from the samples;
 * class weight pos_weight on label-1 samples; loss = last batch's mean loss. */
int kfvm_indicator_train(kfvm_indicator *p, int iters, int batch, double lr, double pos_weight, uint64_t seed,
                         double *loss);
int kfvm_indicator_get_params(kfvm_indicator *p, float *params, int64_t n);
int kfvm_indicator_set_params(kfvm_indicator *p, const float *params, int64_t n);
int kfvm_indicator_get_weights(kfvm_indicator *p, uint16_t *weights, int64_t n);
/* Mean loss and its gradient over n explicit samples at params (as given). */
int kfvm_indicator_gradient(kfvm_indicator *p, const float *params, const float *X, const float *Y, int64_t n,
                            double pos_weight, float *grad, double *loss);
int kfvm_indicator_logits(kfvm_indicator *p, const float *params, const float *X, int64_t n, float *z);

/* ---- host-only entry points (no device work; usable without a GPU) ---- */

/* The writer without a solver: a plot (checkpoint = 0) or checkpoint file of
 * host arrays U [NV][nz][ny][nx] (n values) and flags [2][nz][ny][nx] for the
 * mesh and io: keys of init_text; and the reader of a checkpoint. */
int kfvm_write_output_file(const char *init_text, int checkpoint, const double *U, const double *flags,
                           int64_t n, int step, double time, double dt, double err_est, char *path,
                           int64_t len);
int kfvm_read_checkpoint_file(const char *init_text, const char *path, double *U, int64_t n, int *step,
                              double *time, double *dt, double *err_est);
/* IntegrateIC_K (src/numeric/Numeric_K.H:31-107) without a device: the
 * NQ^D-point Gauss cell averages of the problem the init text names, written
 * to U [NV][nz][ny][nx] (n values) exactly as kfvm_create fills the solver. */
int kfvm_initial_condition(const char *init_text, double *U, int64_t n);
/* Feature and parameter counts of the learned indicator's network. */
int kfvm_indicator_sizes(int dim, int64_t *ninputs, int64_t *nparams);

/* Shapes of the weight tables for (dim, nq): nface = nsub*2*dim*Q*C,
 * nderiv = nsub*I*C. */
int kfvm_weight_sizes(int dim, int nq, int64_t *nface, int64_t *nderiv);
/* Stencil::findWeights (src/stencil/Stencil.cpp:281-846): the GP
 * reconstruction weights for length scale lfac and cell sizes dx, dy, dz. */
int kfvm_build_weights(int dim, int nq, double lfac, double dx, double dy, double dz,
                       double *face, int64_t nface, double *deriv, int64_t nderiv);
/* Stencil::cellWeights [nsub][nq^dim][C] (src/stencil/Stencil.H:99,
 * filled at src/stencil/Stencil.cpp:287-844): GP point values at the cell's
 * interior Gauss points, point index (a*nq + b)*nq + c over (x, y, z).
 * ncell must equal nsub * nq^dim * C. */
int kfvm_build_cell_weights(int dim, int nq, double lfac, double dx, double dy, double dz,
                            double *cell, int64_t ncell);
/* Numeric::DiffMat (src/numeric/Numeric.H:59-106): the [nq][nq+2]
 * differentiation matrix of the Lagrange basis on {-1/2, Gauss points, +1/2},
 * row = Gauss point, column 0 = -1/2 face, nq+1 = +1/2 face.  n = nq*(nq+2). */
int kfvm_diff_matrix(int nq, double *dm, int64_t n);
/* ProblemSetup::setFromFile on text (src/ProblemSetup.H:142-205): writes the
 * parsed configuration as a JSON object into json[len]. */
int kfvm_parse_init(const char *init_text, char *json, int64_t len);

const char *kfvm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
