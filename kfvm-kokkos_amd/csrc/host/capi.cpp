// extern "C" boundary of the stage library (declared in include/kfvm.h).
#include "../../../include/kfvm.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "indicator.hpp"
#include "problems.hpp"
#include "solver.hpp"

struct kfvm_solver {
  kfvm::Solver *s;
};
struct kfvm_indicator {
  kfvm::Indicator *t;
};

namespace {
thread_local std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return -1;
}

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const std::exception &e) {
    return fail(e.what());
  } catch (...) {
    return fail("unknown error");
  }
}

int create(const kfvm::Config &cfg, const std::string &warnings, kfvm_solver **out) {
  if (!out) return fail("null output pointer");
  *out = nullptr;
  return guarded([&] {
    // ProblemSetup::setFromFile reports every skipped line (ProblemSetup.H:190-195)
    if (!warnings.empty()) std::fputs(warnings.c_str(), stderr);
    kfvm::Solver *s = new kfvm::Solver(cfg);
    *out = new kfvm_solver{s};
    return 0;
  });
}

#define CHECK_SOLVER(s) \
  if (!(s) || !(s)->s) return fail("null solver")
#define CHECK_IND(p) \
  if (!(p) || !(p)->t) return fail("null indicator")

int64_t state_size(const kfvm::Solver &s) { return s.ncell() * s.nv(); }
}  // namespace

extern "C" {

const char *kfvm_last_error(void) { return g_err.c_str(); }

int kfvm_create(const char *init_path, kfvm_solver **out) {
  if (!init_path) return fail("null init path");
  kfvm::Config cfg;
  std::string warn;
  if (!kfvm::parse_init_file(init_path, cfg, &warn))
    return fail(std::string("could not open input file: ") + init_path);
  return create(cfg, warn, out);
}

int kfvm_create_from_text(const char *init_text, kfvm_solver **out) {
  if (!init_text) return fail("null init text");
  kfvm::Config cfg;
  std::string warn;
  kfvm::parse_init_text(init_text, cfg, &warn);
  return create(cfg, warn, out);
}

int kfvm_create_batch(const char *init_text, const uint64_t *seeds, int64_t n, kfvm_solver **out) {
  if (!init_text || !seeds || n < 1) return fail("null init text or no seeds");
  kfvm::Config cfg;
  std::string warn;
  kfvm::parse_init_text(init_text, cfg, &warn);
  cfg.slices = int(n);
  cfg.sliceSeeds.assign(seeds, seeds + n);
  return create(cfg, warn, out);
}

int kfvm_get_slice_times(kfvm_solver *s, double *t, double *dt, int64_t n) {
  CHECK_SOLVER(s);
  if (n != s->s->slices()) return fail("one value per slice");
  s->s->sliceTimes(t, dt);
  return 0;
}

int kfvm_destroy(kfvm_solver *s) {
  if (!s) return 0;
  return guarded([&] {
    delete s->s;
    delete s;
    return 0;
  });
}

int kfvm_solve(kfvm_solver *s, int64_t *steps_taken) {
  CHECK_SOLVER(s);
  return guarded([&] {
    const int64_t n = s->s->solve();
    if (steps_taken) *steps_taken = n;
    return 0;
  });
}

int kfvm_take_step(kfvm_solver *s, int *status) {
  CHECK_SOLVER(s);
  return guarded([&] {
    const int st = s->s->takeStep();
    if (status) *status = st;
    return 0;
  });
}

int kfvm_advance_stages(kfvm_solver *s, int64_t nstages, double dt) {
  CHECK_SOLVER(s);
  if (nstages < 0 || !(dt > 0.0)) return fail("need nstages >= 0 and dt > 0");
  return guarded([&] {
    s->s->advanceStages(nstages, dt);
    return 0;
  });
}

int kfvm_eval_rhs(kfvm_solver *s, int which, double t, double *max_vel) {
  CHECK_SOLVER(s);
  if (which != 0 && which != 1) return fail("which must be 0 (current) or 1 (step start)");
  if (!std::isfinite(t)) return fail("time must be finite");
  return guarded([&] {
    const double v = s->s->evalRHS(which);
    if (max_vel) *max_vel = v;
    return 0;
  });
}

namespace {
// true if p lies in device memory with at least `bytes` bytes from p to the
// end of its allocation (clears the error a host pointer leaves)
bool device_span(const void *p, size_t bytes) {
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const char *b = static_cast<const char *>(base), *q = static_cast<const char *>(p);
  return q >= b && size_t(q - b) + bytes <= size;
}
}  // namespace

int kfvm_eval_rhs_noalloc(kfvm_solver *s, const double *u_dev, double *rhs_dev, double t, double *max_vel) {
  CHECK_SOLVER(s);
  if (!std::isfinite(t)) return fail("time must be finite");
  return guarded([&] {
    const size_t bytes = size_t(state_size(*s->s)) * sizeof(double);
    if (!u_dev || !device_span(u_dev, bytes))
      return fail("u must point to device memory holding NV*ncell doubles");
    if (!rhs_dev || !device_span(rhs_dev, bytes))
      return fail("rhs must point to device memory holding NV*ncell doubles");
    const size_t need = s->s->noAllocBytes();
    if (need > kfvm::kNoAllocScratchBytes)
      return fail("this grid needs " + std::to_string(need) + " bytes of scratch; the no-allocation entry point "
                  "has " + std::to_string(kfvm::kNoAllocScratchBytes) + " (use kfvm_eval_rhs)");
    const double v = s->s->evalRhsNoAlloc(u_dev, rhs_dev);
    if (max_vel) *max_vel = v;
    return 0;
  });
}

int kfvm_get_state(kfvm_solver *s, double *host, int64_t n) {
  CHECK_SOLVER(s);
  if (!host || n != state_size(*s->s)) return fail("state buffer must hold NV*ncell doubles");
  return guarded([&] {
    s->s->getState(host);
    return 0;
  });
}

int kfvm_set_state(kfvm_solver *s, const double *host, int64_t n) {
  CHECK_SOLVER(s);
  if (!host || n != state_size(*s->s)) return fail("state buffer must hold NV*ncell doubles");
  return guarded([&] {
    s->s->setState(host);
    return 0;
  });
}

int kfvm_get_rhs(kfvm_solver *s, double *host, int64_t n) {
  CHECK_SOLVER(s);
  if (!host || n != state_size(*s->s)) return fail("rhs buffer must hold NV*ncell doubles");
  return guarded([&] {
    s->s->getRHS(host);
    return 0;
  });
}

int kfvm_get_flags(kfvm_solver *s, double *host, int64_t n) {
  CHECK_SOLVER(s);
  if (!host || n != 2 * s->s->ncell()) return fail("flag buffer must hold 2*ncell doubles");
  return guarded([&] {
    s->s->getFlags(host);
    return 0;
  });
}

int kfvm_set_flags(kfvm_solver *s, const double *host, int64_t n) {
  CHECK_SOLVER(s);
  if (!host || n != 2 * s->s->ncell()) return fail("flag buffer must hold 2*ncell doubles");
  return guarded([&] {
    s->s->setFlags(host);
    return 0;
  });
}

int kfvm_get_time(kfvm_solver *s, double *t, double *dt, int64_t *step) {
  CHECK_SOLVER(s);
  if (t) *t = s->s->time;
  if (dt) *dt = s->s->dt;
  if (step) *step = s->s->nTS;
  return 0;
}

int kfvm_set_time(kfvm_solver *s, double t, double dt) {
  CHECK_SOLVER(s);
  return guarded([&] {
    s->s->setTime(t, dt);
    return 0;
  });
}

int kfvm_get_info(kfvm_solver *s, kfvm_info *info) {
  CHECK_SOLVER(s);
  if (!info) return fail("null info");
  const kfvm::Config &c = s->s->config();
  const kfvm::WeightTable &w = s->s->weights();
  info->dim = c.dim;
  info->nq = c.nQuad;
  info->nv = s->s->nv();
  info->nx = c.nX;
  info->ny = c.nY;
  info->nz = c.dim == 3 ? c.nZ : s->s->slices();
  info->ncell = s->s->ncell();
  info->nsub = w.nSub;
  info->nfaces = w.F;
  info->nqface = w.Q;
  info->ncells_stencil = w.C;
  info->nindic = w.I;
  info->rk_stages = s->s->rkStages();
  info->dx = c.dx();
  info->dy = c.dy();
  info->dz = c.dz();
  info->gamma = c.gamma;
  return 0;
}

int kfvm_get_counters(kfvm_solver *s, int64_t *rhs_evals, int64_t *rejects_tol,
                      int64_t *rejects_unphys, int64_t *last_nweno) {
  CHECK_SOLVER(s);
  if (rhs_evals) *rhs_evals = s->s->nRhsEval;
  if (rejects_tol) *rejects_tol = s->s->nRejectThresh;
  if (rejects_unphys) *rejects_unphys = s->s->nRejectUnphys;
  if (last_nweno) *last_nweno = s->s->lastNWeno;
  return 0;
}

int kfvm_set_indicator(kfvm_solver *s, const uint16_t *weights, int64_t n, double threshold) {
  CHECK_SOLVER(s);
  if (n < 0 || (n > 0 && !weights)) return fail("null indicator weights");
  return guarded([&] {
    std::vector<float> P(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; i++) {
      const uint32_t u = uint32_t(weights[i]) << 16;
      std::memcpy(&P[size_t(i)], &u, 4);
    }
    s->s->setIndicator(P.data(), int(n), threshold);
    return 0;
  });
}

int kfvm_set_nad(kfvm_solver *s, int on, double range_frac, double value_frac, double max_frac) {
  CHECK_SOLVER(s);
  if (!(range_frac >= 0.0) || !(value_frac >= 0.0) || !(max_frac >= 0.0))
    return fail("NAD tolerances must be >= 0");
  s->s->setNad(on != 0, range_frac, value_frac, max_frac);
  return 0;
}

int kfvm_get_indicator_counters(kfvm_solver *s, int64_t *nad_fallbacks, int64_t *nad_cells, int *learned) {
  CHECK_SOLVER(s);
  if (nad_fallbacks) *nad_fallbacks = s->s->nNadFallback;
  if (nad_cells) *nad_cells = s->s->nNadCells;
  if (learned) *learned = s->s->learned() ? 1 : 0;
  return 0;
}

int kfvm_indicator_create(int dim, kfvm_indicator **out) {
  if (!out) return fail("null output pointer");
  *out = nullptr;
  return guarded([&] {
    kfvm::Indicator *t = new kfvm::Indicator(dim);
    *out = new kfvm_indicator{t};
    return 0;
  });
}

int kfvm_indicator_destroy(kfvm_indicator *p) {
  if (!p) return 0;
  return guarded([&] {
    delete p->t;
    delete p;
    return 0;
  });
}

int kfvm_indicator_collect(kfvm_indicator *p, kfvm_solver *s, double tau, int64_t *nsamples,
                           int64_t *npositive) {
  CHECK_IND(p);
  CHECK_SOLVER(s);
  return guarded([&] {
    p->t->collect(*s->s, tau);
    if (nsamples) *nsamples = p->t->samples();
    if (npositive) *npositive = p->t->positives();
    return 0;
  });
}

int kfvm_indicator_samples(kfvm_indicator *p, float *X, float *Y, int64_t offset, int64_t n) {
  CHECK_IND(p);
  return guarded([&] {
    p->t->getSamples(X, Y, offset, n);
    return 0;
  });
}

int kfvm_indicator_init(kfvm_indicator *p, uint64_t seed) {
  CHECK_IND(p);
  return guarded([&] {
    p->t->init(seed);
    return 0;
  });
}

int kfvm_indicator_train(kfvm_indicator *p, int iters, int batch, double lr, double pos_weight, uint64_t seed,
                         double *loss) {
  CHECK_IND(p);
  return guarded([&] {
    const double l = p->t->train(iters, batch, lr, pos_weight, seed);
    if (loss) *loss = l;
    return 0;
  });
}

int kfvm_indicator_get_params(kfvm_indicator *p, float *params, int64_t n) {
  CHECK_IND(p);
  if (!params || n != p->t->params()) return fail("parameter buffer has the wrong size");
  return guarded([&] {
    p->t->getParams(params);
    return 0;
  });
}

int kfvm_indicator_set_params(kfvm_indicator *p, const float *params, int64_t n) {
  CHECK_IND(p);
  if (!params || n != p->t->params()) return fail("parameter buffer has the wrong size");
  return guarded([&] {
    p->t->setParams(params);
    return 0;
  });
}

int kfvm_indicator_get_weights(kfvm_indicator *p, uint16_t *weights, int64_t n) {
  CHECK_IND(p);
  if (!weights || n != p->t->params()) return fail("weight buffer has the wrong size");
  return guarded([&] {
    p->t->getBf16(weights);
    return 0;
  });
}

int kfvm_indicator_gradient(kfvm_indicator *p, const float *params, const float *X, const float *Y, int64_t n,
                            double pos_weight, float *grad, double *loss) {
  CHECK_IND(p);
  if (!params || !X || !Y || !grad || n <= 0 || n > (int64_t(1) << 24)) return fail("bad gradient arguments");
  return guarded([&] {
    const double l = p->t->gradient(params, X, Y, n, pos_weight, grad);
    if (loss) *loss = l;
    return 0;
  });
}

int kfvm_indicator_logits(kfvm_indicator *p, const float *params, const float *X, int64_t n, float *z) {
  CHECK_IND(p);
  if (!params || !X || !z || n <= 0 || n > (int64_t(1) << 28)) return fail("bad logits arguments");
  return guarded([&] {
    p->t->logits(params, X, n, z);
    return 0;
  });
}

namespace {
int copy_path(const std::string &p, char *out, int64_t len) {
  if (out && len > 0) {
    const size_t n = std::min<size_t>(p.size(), size_t(len - 1));
    std::memcpy(out, p.data(), n);
    out[n] = '\0';
  }
  return 0;
}
}  // namespace

int kfvm_write_plot(kfvm_solver *s, char *path, int64_t len) {
  CHECK_SOLVER(s);
  return guarded([&] {
    const std::string p = s->s->writePlot();
    s->s->plotNum++;
    return copy_path(p, path, len);
  });
}

int kfvm_write_checkpoint(kfvm_solver *s, char *path, int64_t len) {
  CHECK_SOLVER(s);
  return guarded([&] { return copy_path(s->s->writeCkpt(), path, len); });
}

int kfvm_flow_stats(kfvm_solver *s, double *stats, int *n) {
  CHECK_SOLVER(s);
  if (!stats || !n) return fail("kfvm_flow_stats: null output");
  return guarded([&] {
    *n = s->s->flowStats(stats);
    return 0;
  });
}

int kfvm_force_full_weno(kfvm_solver *s) {
  CHECK_SOLVER(s);
  s->s->useSparseWeno = false;
  return 0;
}

int kfvm_get_weights(kfvm_solver *s, double *face, int64_t nface, double *deriv, int64_t nderiv) {
  CHECK_SOLVER(s);
  const kfvm::WeightTable &w = s->s->weights();
  if (nface != int64_t(w.face.size()) || nderiv != int64_t(w.deriv.size()))
    return fail("weight buffers have the wrong size");
  if (face) std::memcpy(face, w.face.data(), w.face.size() * sizeof(double));
  if (deriv) std::memcpy(deriv, w.deriv.data(), w.deriv.size() * sizeof(double));
  return 0;
}

int kfvm_set_weights(kfvm_solver *s, const double *face, int64_t nface, const double *deriv,
                     int64_t nderiv) {
  CHECK_SOLVER(s);
  const kfvm::WeightTable &w = s->s->weights();
  if (!face || !deriv || nface != int64_t(w.face.size()) || nderiv != int64_t(w.deriv.size()))
    return fail("weight buffers have the wrong size");
  return guarded([&] {
    s->s->setWeights(face, deriv);
    return 0;
  });
}

int kfvm_profile_enable(kfvm_solver *s, int on) {
  CHECK_SOLVER(s);
  return guarded([&] {
    s->s->profileEnable(on != 0);
    return 0;
  });
}

int kfvm_profile_get(kfvm_solver *s, const char *name, double *ms, int64_t *count) {
  CHECK_SOLVER(s);
  if (!name) return fail("null kernel name");
  return guarded([&] {
    double m = 0.0;
    int64_t c = 0;
    s->s->profileGet(name, m, c);
    if (ms) *ms = m;
    if (count) *count = c;
    return 0;
  });
}

int kfvm_profile_reset(kfvm_solver *s) {
  CHECK_SOLVER(s);
  return guarded([&] {
    s->s->profileReset();
    return 0;
  });
}

int kfvm_sync(kfvm_solver *s) {
  CHECK_SOLVER(s);
  return guarded([&] {
    s->s->sync();
    return 0;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- host only
#include <sstream>

#include "weights.hpp"

extern "C" {

int kfvm_weight_sizes(int dim, int nq, int64_t *nface, int64_t *nderiv) {
  if ((dim != 2 && dim != 3) || (nq != 2 && nq != 3)) return fail("dim must be 2|3 and nq 2|3");
  const int nSub = dim == 3 ? 8 : 6, C = dim == 3 ? 33 : 13, I = dim == 3 ? 19 : 5;
  const int Q = dim == 3 ? nq * nq : nq;
  if (nface) *nface = int64_t(nSub) * 2 * dim * Q * C;
  if (nderiv) *nderiv = int64_t(nSub) * I * C;
  return 0;
}

int kfvm_build_weights(int dim, int nq, double lfac, double dx, double dy, double dz, double *face,
                       int64_t nface, double *deriv, int64_t nderiv) {
  int64_t nf = 0, nd = 0;
  if (kfvm_weight_sizes(dim, nq, &nf, &nd) != 0) return -1;
  if (!face || !deriv || nface != nf || nderiv != nd) return fail("weight buffers have the wrong size");
  return guarded([&] {
    kfvm::WeightTable w = kfvm::build_weights(dim, nq, lfac, dx, dy, dz);
    std::memcpy(face, w.face.data(), w.face.size() * sizeof(double));
    std::memcpy(deriv, w.deriv.data(), w.deriv.size() * sizeof(double));
    return 0;
  });
}

int kfvm_build_cell_weights(int dim, int nq, double lfac, double dx, double dy, double dz,
                            double *cell, int64_t ncell) {
  if ((dim != 2 && dim != 3) || (nq != 2 && nq != 3)) return fail("dim must be 2|3 and nq 2|3");
  const int nSub = dim == 3 ? 8 : 6, C = dim == 3 ? 33 : 13;
  const int64_t n = int64_t(nSub) * (dim == 3 ? nq * nq * nq : nq * nq) * C;
  if (!cell || ncell != n) return fail("cell weight buffer has the wrong size");
  return guarded([&] {
    kfvm::WeightTable w = kfvm::build_weights(dim, nq, lfac, dx, dy, dz);
    std::memcpy(cell, w.cell.data(), w.cell.size() * sizeof(double));
    return 0;
  });
}

int kfvm_diff_matrix(int nq, double *dm, int64_t n) {
  if (nq != 2 && nq != 3) return fail("nq must be 2|3");
  if (!dm || n != int64_t(nq) * (nq + 2)) return fail("differentiation matrix buffer has the wrong size");
  return guarded([&] {
    const std::vector<double> D = kfvm::diff_matrix(nq);
    std::memcpy(dm, D.data(), D.size() * sizeof(double));
    return 0;
  });
}

int kfvm_write_output_file(const char *init_text, int checkpoint, const double *U, const double *flags,
                           int64_t n, int step, double time, double dt, double err_est, char *path,
                           int64_t len) {
  if (!init_text || !U || !flags) return fail("null argument");
  return guarded([&] {
    kfvm::Config c;
    kfvm::parse_init_text(init_text, c, nullptr);
    if (c.dim == 2) c.nZ = 1;
    if (n != c.nCells() * kfvm::num_vars(c.eq)) return fail("state buffer has the wrong size");
    kfvm::Writer w(c);
    kfvm::OutputFields f;
    f.U = U;
    f.flags = flags;
    const std::string p = checkpoint ? w.writeCkpt(f, step, time, dt, err_est) : w.writePlot(f, step, time);
    return copy_path(p, path, len);
  });
}

int kfvm_read_checkpoint_file(const char *init_text, const char *path, double *U, int64_t n, int *step,
                              double *time, double *dt, double *err_est) {
  if (!init_text || !path || !U) return fail("null argument");
  return guarded([&] {
    kfvm::Config c;
    kfvm::parse_init_text(init_text, c, nullptr);
    if (c.dim == 2) c.nZ = 1;
    if (n != c.nCells() * kfvm::num_vars(c.eq)) return fail("state buffer has the wrong size");
    const kfvm::CkptData d = kfvm::Writer(c).readCkpt(path);
    std::memcpy(U, d.U.data(), d.U.size() * sizeof(double));
    if (step) *step = d.step;
    if (time) *time = d.time;
    if (dt) *dt = d.dt;
    if (err_est) *err_est = d.errEst;
    return 0;
  });
}

int kfvm_indicator_sizes(int dim, int64_t *ninputs, int64_t *nparams) {
  if (dim != 2 && dim != 3) return fail("dim must be 2 or 3");
  const int C = dim == 3 ? 33 : 13, H = 32, nin = 2 * C;
  if (ninputs) *ninputs = nin;
  if (nparams) *nparams = H * nin + H + H * H + H + H + 1;
  return 0;
}

int kfvm_parse_init(const char *init_text, char *json, int64_t len) {
  if (!init_text || !json || len <= 0) return fail("null argument");
  return guarded([&] {
    kfvm::Config c;
    std::string warn;
    kfvm::parse_init_text(init_text, c, &warn);
    std::ostringstream o;
    o.precision(17);
    const char *bcn[] = {"periodic", "outflow", "reflecting", "noslip"};
    o << "{\"dim\":" << c.dim << ",\"n\":[" << c.nX << "," << c.nY << "," << c.nZ << "]"
      << ",\"lo\":[" << c.xLo << "," << c.yLo << "," << c.zLo << "],\"hi\":[" << c.xUp << "," << c.yUp
      << "," << c.zUp << "],\"nq\":" << c.nQuad << ",\"lfac\":" << c.lfac << ",\"eq\":" << int(c.eq)
      << ",\"rs\":" << int(c.rs) << ",\"rk\":" << int(c.rk) << ",\"bc\":[";
    for (int f = 0; f < 6; f++) o << (f ? "," : "") << "\"" << bcn[int(c.bc[f])] << "\"";
    o << "],\"problem\":\"" << c.problem << "\",\"final\":" << c.finalTime << ",\"mindt\":" << c.minDt
      << ",\"cfl\":" << c.cfl << ",\"atol\":" << c.atol << ",\"rtol\":" << c.rtol
      << ",\"rejectionthresh\":" << c.rejectionThresh << ",\"rejectionlimit\":" << c.rejectionLimit
      << ",\"maxsteps\":" << c.maxSteps << ",\"fixeddt\":" << c.fixedDt << ",\"gamma\":" << c.gamma
      << ",\"wenothresh\":" << c.wenoThresh << ",\"advvel\":" << (c.haveAdvVel ? "[" : "null");
    if (c.haveAdvVel) o << c.advVel[0] << "," << c.advVel[1] << "," << c.advVel[2] << "]";
    o << ",\"user\":[";
    for (int i = 0; i < 8; i++) o << (i ? "," : "") << c.user[i];
    o << "],\"seed\":" << c.seed << ",\"problem_set\":" << (c.problemSet ? "true" : "false") << ",\"warnings\":" << std::count(warn.begin(), warn.end(), '\n') << "}";
    const std::string s = o.str();
    if (int64_t(s.size()) + 1 > len) return fail("json buffer too small");
    std::memcpy(json, s.c_str(), s.size() + 1);
    return 0;
  });
}

int kfvm_initial_condition(const char *init_text, double *U, int64_t n) {
  if (!init_text || !U) return fail("null argument");
  return guarded([&] {
    kfvm::Config c;
    std::string warn;
    kfvm::parse_init_text(init_text, c, &warn);
    if (!c.problemSet) return fail("the init text names no problem (problem:name)");
    if (!kfvm::problem_known(c.problem, c.eq)) return fail("unknown problem '" + c.problem + "' for this equation");
    if (c.dim == 2) c.nZ = 1;
    std::vector<double> V;
    kfvm::fill_initial_condition(c, V);
    if (n != int64_t(V.size())) return fail("U must hold NV*ncell doubles");
    std::copy(V.begin(), V.end(), U);
    return 0;
  });
}

}  // extern "C"
