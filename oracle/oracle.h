/* CPU oracle for the KFVM-WENO per-stage update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++17/OpenMP restatement of the
 * reference algorithm in the reference's own unfused dataflow (cell BCs ->
 * linear recon -> face BCs -> KXRCF selector -> sparse WENO -> positivity
 * limiter -> face BCs -> Riemann sweeps -> IntegrateRHS -> RK registers).  It
 * is used by tests/ as the checker of the HIP path and by bench.py as the
 * timed CPU baseline.  The product library never links or calls it.
 */
#ifndef KFVM_ORACLE_H
#define KFVM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_config {
  int dim;        /* 2 or 3 */
  int nq;         /* face Gauss points per direction: 2 or 3 */
  int eq;         /* 0 hydro, 1 linadv, 2 GLM-MHD */
  int rs;         /* 0 LLF, 1 HLL, 2 HLLC, 3 exact (linadv), 4 KEPES (MHD) */
  int rk;         /* 0 ThreeStarP, 1 FourStarP, 2 SSP4_3_2, 3 SSP10_4_3 */
  int bc[6];      /* 0 periodic, 1 outflow, 2 reflecting, 3 noslip */
  int64_t n[3];   /* cells per direction (n[2] = 1 in 2D) */
  double lo[3], hi[3];
  double gamma, wenoThresh;
  int haveAdvVel;
  double advVel[3];
  double cfl, finalTime, minDt, atol, rtol, rejectionThresh, fixedDt;
  int rejectionLimit, maxSteps;
  int nThreads; /* OpenMP threads, <= 0: runtime default */
} orc_config;

typedef struct orc_solver orc_solver;

/* faceW: [nSub][2*dim][Q][C], derivW: [nSub][I][C] (the same table the GPU uses) */
orc_solver *orc_create(const orc_config *cfg, const double *faceW, const double *derivW);
void orc_destroy(orc_solver *s);

/* Cell averages in [NV][nz][ny][nx] order. */
void orc_set_state(orc_solver *s, const double *U);
void orc_get_state(const orc_solver *s, double *U);
/* Fill U with the NQ^D-point Gauss average of a registry problem.  Returns 0 on success. */
int orc_set_ic(orc_solver *s, const char *problem, const double *user, uint64_t seed);

/* One RHS evaluation of U (which=0) or Uprev (which=1); returns max wave speed. */
double orc_eval_rhs(orc_solver *s, int which);
void orc_get_rhs(const orc_solver *s, double *rhs);
/* Flags [2][nz][ny][nx]: slot 0 WENO flag, slot 1 positivity theta. */
void orc_get_flags(const orc_solver *s, double *flags);
void orc_set_flags(orc_solver *s, const double *flags);

/* Solver::Solver tail: FSAL prime and dt0 = 0.01 cfl dmin / vmax. */
void orc_prime(orc_solver *s);
/* Solver::TakeStep.  Returns the final status of the step (0 accepted). */
int orc_take_step(orc_solver *s);
void orc_get_time(const orc_solver *s, double *t, double *dt, int64_t *step, double *errEst);
void orc_set_time(orc_solver *s, double t, double dt);
int64_t orc_rhs_evals(const orc_solver *s);
/* Force the next RHS to full WENO (as after an unphysical rejection). */
void orc_force_full_weno(orc_solver *s);
int64_t orc_num_weno(const orc_solver *s);
int orc_last_step(const orc_solver *s);
/* GLM-MHD tables: interior-point weights [nSub][NQ^D][C] and the
 * differentiation matrix [NQ][NQ+2]; required before the first RHS. */
void orc_set_mhd_tables(orc_solver *s, const double *cellW, const double *Dm);
/* Own face states of every cell from the current state: linear (which = 0) or
 * full WENO + positivity limiter (which = 1); out[((f*Q + q)*NV + v)*ncell + cell]. */
void orc_face_states(orc_solver *s, int which, double *out);
/* which = 2 of orc_face_states: linear reconstruction followed by the
 * positivity limiter (the limiter's theta is not kept). */
/* Nonlinear WENO-AO weights [nSub][NV] of interior cell (i, j, k) of the
 * current state (KernelWenoRecon_K::createOmegas, Stencil_K.H:556-580). */
void orc_weno_omegas(orc_solver *s, int64_t i, int64_t j, int64_t k, double *out);
/* RKFSAL_ErrEst_K (Numeric_K.H:262-290) over ncell cells of interleaved
 * states U[cell][nv], Uhat[cell][nv]: the sum of squared scaled differences
 * and the minimum positivity (-1 for a NaN).  eq as in orc_config. */
void orc_error_norm(int nv, int eq, int64_t ncell, const double *U, const double *Uhat, double atol,
                    double rtol, double *errSum, double *posFlag);
/* One face-point Riemann solve in direction dir (0 x, 1 y, 2 z): eq 0 hydro
 * (rs 0 LLF, 1 HLL, 2 HLLC) or 2 GLM-MHD (rs 0 LLF, 1 HLL, 4 KEPES, cleaning
 * speed ch).  Writes the flux to F and returns the maximum wave speed. */
double orc_riemann_point(int eq, int rs, double gamma, double ch, int dir, const double *uL,
                         const double *uR, double *F);

#ifdef __cplusplus
}
#endif

#endif
