// Host driver of one KFVM problem on one GPU.  See solver.hpp.
//
// Per RK stage the device runs (Solver::evalRHS, src/Solver.cpp:410-530):
//   sparse mode:  linear recon (+KXRCF indicator) -> select -> WENO on the
//                 compacted list -> positivity limiter -> flux/RK kernel
//   full mode:    WENO on every cell -> positivity limiter -> flux/RK kernel
// The RK register update (src/numeric/Numeric_K.H:198-260) and the error norm
// (Numeric_K.H:262-290) are the flux kernel's epilogue, U ping-pongs between
// two buffers, and the host reads three scalars once per step attempt.
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "problems.hpp"
#include "stencil_geom.hpp"

namespace kfvm {

namespace {
// A named host range for rocprofv3 --marker-trace, one per region of the
// reference's Kokkos profiling hooks (Solver.cpp:85 Solver::Solve, 192
// Solver::TakeStep, 411 Solver::evalRHS, 1949 Solver::setIC).
struct Region {
  explicit Region(const char *name) { roctxRangePushA(name); }
  ~Region() { roctxRangePop(); }
  Region(const Region &) = delete;
  Region &operator=(const Region &) = delete;
};

void check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
template <class T>
T *dalloc(size_t n, const char *what) {
  void *p = nullptr;
  check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), what);
  return static_cast<T *>(p);
}
}  // namespace

Solver::Solver(const Config &cfg) : cfg_(cfg) {
  // a failed set-up (bad restart file, device out of memory, ...) releases
  // what it allocated before the exception leaves the constructor
  try {
    setup();
  } catch (...) {
    release();
    throw;
  }
}

void Solver::setup() {
  if (!cfg_.problemSet)
    throw std::runtime_error(
        "the init file names no problem (problem:name): the reference compiles its problem in "
        "(Definitions.H, InitialCondition.H), this library selects it from its registry: " +
        problem_names());
  if (cfg_.dim != 2 && cfg_.dim != 3) throw std::runtime_error("problem:dim must be 2 or 3");
  if (cfg_.rad != 2) throw std::runtime_error("only stencil radius 2 is built");
  if (cfg_.nQuad != 2 && cfg_.nQuad != 3) throw std::runtime_error("problem:nquad must be 2 or 3");
  if (cfg_.fluxBrick < 0 || cfg_.fluxBrick > (cfg_.dim == 2 ? 2 : 0))
    throw std::runtime_error("tune:fluxbrick must be 0, 1 or 2 in 2D and 0 in 3D");
  if (cfg_.dim == 2) cfg_.nZ = 1;
  const int64_t nmin = cfg_.dim == 3 ? std::min({cfg_.nX, cfg_.nY, cfg_.nZ}) : std::min(cfg_.nX, cfg_.nY);
  if (nmin < 2) throw std::runtime_error("every direction needs at least 2 cells");
  for (int a = 0; a < cfg_.dim; a++)
    if ((cfg_.bc[2 * a] == BC::Periodic) != (cfg_.bc[2 * a + 1] == BC::Periodic))
      throw std::runtime_error("periodic boundaries must be set on both sides of an axis");
  if (cfg_.eq == Eq::LinAdv && cfg_.rs != Riemann::Exact)
    throw std::runtime_error("linear advection uses problem:riemann exact");
  if (cfg_.eq == Eq::Hydro && (cfg_.rs == Riemann::Exact || cfg_.rs == Riemann::KEPES))
    throw std::runtime_error("hydro needs problem:riemann hllc, hll or llf");
  if (cfg_.eq == Eq::MhdGlm && (cfg_.rs == Riemann::Exact || cfg_.rs == Riemann::HLLC))
    throw std::runtime_error("mhd_glm needs problem:riemann llf, hll or kepes");
  if (!problem_known(cfg_.problem, cfg_.eq))
    throw std::runtime_error("unknown problem '" + cfg_.problem + "' for this equation");
  if (cfg_.cfl <= 0.0) cfg_.cfl = cfg_.cflValue();

  nv_ = num_vars(cfg_.eq);
  slices_ = cfg_.slices;
  if (slices_ < 1) throw std::runtime_error("problem:slices must be >= 1");
  if (slices_ == 1 && !cfg_.sliceSeeds.empty()) cfg_.seed = cfg_.sliceSeeds[0];  // a batch of one
  if (slices_ > 1) {
    if (cfg_.dim != 2 || cfg_.eq == Eq::MhdGlm)
      throw std::runtime_error("batches of slices are built for 2D hydro and linear advection");
    if (cfg_.plotFreq > 0.0 || cfg_.ckptFreq > 0 || !cfg_.restartFile.empty())
      throw std::runtime_error("batches of slices have no plot / checkpoint output");
    if (!cfg_.sliceSeeds.empty() && int(cfg_.sliceSeeds.size()) != slices_)
      throw std::runtime_error("one seed per slice required");
  }
  ncell_ = cfg_.nCells() * (cfg_.dim == 2 ? slices_ : 1);
  if (ncell_ > (int64_t(1) << 31) - 1) throw std::runtime_error("grid too large for one block");
  rk_ = rk_table(cfg_.rk);
  wts_ = build_weights(cfg_.dim, cfg_.nQuad, cfg_.lfac, cfg_.dx(), cfg_.dy(), cfg_.dz());

  grid_.nx = int(cfg_.nX);
  grid_.ny = int(cfg_.nY);
  grid_.nz = int(cfg_.dim == 3 ? cfg_.nZ : slices_);  // 2D: one z-plane per slice
  grid_.ncell = int(ncell_);
  for (int f = 0; f < 6; f++) {
    grid_.bc[f] = int(cfg_.bc[f]);
    const int ax = f / 2;
    // mirrored walls negate the normal momentum (reflecting) or all momenta
    // (noslip), and for GLM-MHD the normal field (Hydro_BCCoeffs.H:15-21,
    // MHD_GLM_BCCoeffs.H:16-22)
    const unsigned bn = cfg_.eq == Eq::MhdGlm ? (1u << (5 + ax)) : 0u;
    grid_.negmask[f] = cfg_.bc[f] == BC::Reflecting ? ((1u << (1 + ax)) | bn)
                       : cfg_.bc[f] == BC::Noslip   ? (0xEu | bn)
                                                    : 0u;
  }
  grid_.dx = cfg_.dx();
  grid_.dy = cfg_.dy();
  grid_.dz = cfg_.dz();
  grid_.gamma = cfg_.gamma;
  const QuadRule qr = quad_rule(cfg_.nQuad);
  for (int q = 0; q < 3; q++) grid_.wt[q] = q < qr.n ? qr.wt[q] : 0.0;
  grid_.haveVel = cfg_.haveAdvVel ? 1 : 0;
  if (cfg_.haveAdvVel) {
    for (int d = 0; d < 3; d++) grid_.vel[d] = cfg_.advVel[d];
  } else {
    grid_.vel[0] = std::cos(cfg_.gamma);
    grid_.vel[1] = std::sin(cfg_.gamma);
    grid_.vel[2] = 0.0;
  }
  grid_.rs = int(cfg_.rs);

  check(hipStreamCreate(&stream_), "hipStreamCreate");
  const size_t nc = size_t(ncell_);
  const size_t FQ = size_t(wts_.F) * wts_.Q;
  U_[0] = dalloc<double>(nc * nv_, "U0");
  U_[1] = dalloc<double>(nc * nv_, "U1");
  Uprev_ = dalloc<double>(nc * nv_, "Uprev");
  Uhat_ = dalloc<double>(nc * nv_, "Uhat");
  K_ = dalloc<double>(nc * nv_, "K");
  if (rk_.needsTmp()) Utmp_ = dalloc<double>(nc * nv_, "Utmp");
  FS_ = dalloc<double>(fs_elems(nc, int(FQ) * nv_), "face states");
  if (nv_ == 5) {
    // The linear pass stores every face state and a compact WENO staging
    // holds the WENO states of up to a quarter of the cells, unless
    // KFVM_FUSED=1 selects the fused reconstruction + flux path (the flux
    // kernel recomputes the linear states from its brick: at 256^3 it costs
    // 12.6 ms against 7.2 ms for the unfused flux kernel while the linear
    // pass stays at 5.3 ms, since both kernels are bound by FP64 VALU work,
    // not by the face-state traffic)
    const char *ev = std::getenv("KFVM_FUSED");
    fused_ = ev && std::string(ev) == "1";
    pos_ = dalloc<int>(nc, "row of each cell");
    if (fused_) {
      wsCap_ = int(nc);  // every listed cell has a row, in FS_
      SB_ = dalloc<double>(surf_elems(cfg_.dim, cfg_.nQuad, grid_, sb_), "surface planes");
      faceWQ_ = dalloc<double>(size_t(wts_.F) * wts_.Q * wts_.C, "face weights (point-major)");
    } else {
      wsCap_ = int(std::max<size_t>(1, (nc + 3) / 4));
      WS_ = dalloc<double>(size_t(wsCap_) * FQ * nv_, "compact WENO states");
    }
  }
  WI_ = dalloc<double>(fs_elems(nc, int(FQ)), "indicator");
  props_ = dalloc<double>(nc * 2, "cell properties");
  if (cfg_.eq == Eq::MhdGlm) {
    const QuadRule qr2 = quad_rule(cfg_.nQuad);
    const int NI = wts_.nqCell;
    SRC_ = dalloc<double>(nc * NI * nv_, "interior point states");
    cellW_ = dalloc<double>(wts_.cell.size(), "interior weights");
    cellWS_ = dalloc<double>(wts_.cell.size(), "interior weights (stencil-major)");
    diffM_ = dalloc<double>(size_t(qr2.n) * (qr2.n + 2), "differentiation matrix");
    chBits_ = dalloc<unsigned long long>(2, "cleaning speed");
    spd_ = dalloc<double>(nc * 2, "cleaning-speed maxima per cell");
    mhdWT_ = dalloc<double>(size_t(FQ + NI) * wts_.C, "linear weights (pass-major)");
    const std::vector<double> Dm = diff_matrix(cfg_.nQuad);
    check(hipMemcpy(diffM_, Dm.data(), Dm.size() * sizeof(double), hipMemcpyHostToDevice), "copy diff matrix");
  }
  theta_[0] = dalloc<double>(nc, "theta");
  theta_[1] = dalloc<double>(nc, "theta");
  wflag_ = dalloc<double>(nc, "wflag");
  list_ = dalloc<int>(nc, "weno list");
  counter_ = dalloc<int>(4, "weno / limiter counts");
  ppList_ = dalloc<int>(nc, "limiter list");
  faceW_ = dalloc<double>(wts_.face.size(), "face weights");
  derivW_ = dalloc<double>(wts_.deriv.size(), "deriv weights");
  faceWT_ = dalloc<double>(size_t(wts_.F) * wts_.Q * wts_.C, "face weights (pass-major)");
  faceWS_ = dalloc<double>(wts_.face.size(), "face weights (stencil-major)");
  derivWS_ = dalloc<double>(wts_.deriv.size(), "deriv weights (stencil-major)");
  if (nv_ == 5 || nv_ == 9) {
    const int SUB = cfg_.dim == 3 ? SInfo<3>::C + SInfo<3>::Cc + SInfo<3>::B * SInfo<3>::Cb
                                  : SInfo<2>::C + SInfo<2>::Cc + SInfo<2>::B * SInfo<2>::Cb;
    const size_t cellRows = nv_ == 9 ? size_t(wts_.nqCell) : 0;
    faceWC_ = dalloc<double>(size_t(SUB) * (FQ + wts_.I + cellRows), "WENO weights (compact)");
  }
  nPart_ = flux_rk_partials(cfg_.dim, grid_, cfg_.fluxBrick);
  errPart_ = dalloc<double>(nPart_, "err partials");
  posPart_ = dalloc<double>(nPart_, "pos partials");
  redOut_ = dalloc<double>(2, "reduction");
  vmax_ = dalloc<unsigned long long>(size_t(slices_), "vmax");
  if (slices_ > 1) {
    sDtDev_ = dalloc<double>(size_t(slices_), "slice dt");
    sRed_ = dalloc<double>(2 * size_t(slices_), "slice reductions");
    sMask_ = dalloc<int>(size_t(slices_), "slice mask");
    const char *eg = std::getenv("KFVM_GRAPH");
    graphsOn_ = !(eg && std::string(eg) == "0");
    thBak_ = dalloc<double>(nc, "accepted slices' theta");
    wfBak_ = dalloc<double>(nc, "accepted slices' WENO flags");
    check(hipHostMalloc((void **)&sHost_, 4 * size_t(slices_) * sizeof(double), hipHostMallocDefault),
          "hipHostMalloc");
    // three host mask regions (attempt's pending slices, attempt's accepted
    // slices, step's accepted slices), so that no region is rewritten while
    // its asynchronous copy may still be queued
    check(hipHostMalloc((void **)&sMaskHost_, 3 * size_t(slices_) * sizeof(int), hipHostMallocDefault),
          "hipHostMalloc");
  }
  check(hipHostMalloc((void **)&hostRed_, 4 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");

  uploadWeights();
  // device address of the no-allocation entry point's scratch array, taken
  // once per device at set-up; a failure is reported by that entry point
  try {
    (void)noalloc_scratch();
  } catch (const std::exception &) {
  }
  check(launch_fill(theta_[0], 1.0, nc, stream_), "fill theta");
  check(launch_fill(theta_[1], 1.0, nc, stream_), "fill theta");
  check(launch_fill(wflag_, 0.0, nc, stream_), "fill wflag");
  check(launch_zero(counter_, sizeof(int), stream_), "clear counter");

  const bool restart = !cfg_.restartFile.empty();
  if (cfg_.plotFreq > 0.0 || cfg_.ckptFreq > 0 || restart) writer_.reset(new Writer(cfg_));
  {
  Region icRegion("Solver::setIC");
  if (!restart && slices_ > 1) {
    // one initial condition per slice, each with its own seed
    std::vector<double> U0(size_t(ncell_) * nv_), one;
    const size_t cps = size_t(ncell_ / slices_);
    for (int sl = 0; sl < slices_; sl++) {
      Config c1 = cfg_;
      c1.slices = 1;
      c1.seed = cfg_.sliceSeeds.empty() ? cfg_.seed + uint64_t(sl) : cfg_.sliceSeeds[size_t(sl)];
      fill_initial_condition(c1, one);
      for (int v = 0; v < nv_; v++)
        std::copy(one.begin() + v * cps, one.begin() + (v + 1) * cps, U0.begin() + v * ncell_ + sl * cps);
    }
    setState(U0.data());
  } else if (!restart) {
    std::vector<double> U0;
    fill_initial_condition(cfg_, U0);
    setState(U0.data());
  } else {
    // checkpoint restart (Solver.cpp:60-66); errEst and the limiter's theta
    // are restored too when the file has them
    const CkptData d = writer_->readCkpt(cfg_.restartFile);
    setState(d.U.data());
    time = d.time;
    nTS = d.step + 1;
    dt = d.dt;
    if (d.haveErrEst) errEst = d.errEst;
    if (!d.theta.empty())
      check(hipMemcpy(theta_[th_], d.theta.data(), nc * sizeof(double), hipMemcpyHostToDevice), "restore theta");
    plotNum = cfg_.plotFreq > 0.0 ? int(std::ceil(time / cfg_.plotFreq)) : 0;
  }
  }
  useSparseWeno = cfg_.wenoThresh > 0.0;
  if (!restart) {
    // initial outputs (Solver.cpp:50-59)
    if (cfg_.plotFreq > 0.0) {
      writePlot();
      plotNum++;
    }
    if (cfg_.ckptFreq > 0) {
      const int64_t keep = nTS;
      nTS = 0;
      writeCkpt();
      nTS = keep;
    }
  }

  // FSAL prime and the small first step (Solver.cpp:73-80)
  const double maxVel = evalRHS(0);
  if (!restart) dt = 0.01 * cfg_.cfl * cfg_.dmin() / maxVel;
  if (slices_ > 1) {
    std::vector<unsigned long long> vm(static_cast<size_t>(slices_));
    check(hipMemcpy(vm.data(), vmax_, vm.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost), "read vmax");
    sT_.assign(size_t(slices_), 0.0);
    sErr_.assign(size_t(slices_), 1.0);
    sLast_.assign(size_t(slices_), 0);
    sDone_.assign(size_t(slices_), 0);
    sDt_.resize(size_t(slices_));
    for (int sl = 0; sl < slices_; sl++) {
      double v;
      std::memcpy(&v, &vm[size_t(sl)], sizeof(double));
      sDt_[size_t(sl)] = 0.01 * cfg_.cfl * cfg_.dmin() / v;
    }
    dt = *std::min_element(sDt_.begin(), sDt_.end());
  }
}

void Solver::setTime(double t, double dtNext) {
  time = t;
  dt = dtNext;
  for (int sl = 0; sl < slices_ && slices_ > 1; sl++) {
    sT_[size_t(sl)] = t;
    sDt_[size_t(sl)] = dtNext;
    sDone_[size_t(sl)] = t >= cfg_.finalTime;
    sLast_[size_t(sl)] = 0;
  }
}

void Solver::sliceTimes(double *t, double *dtOut) const {
  for (int sl = 0; sl < slices_; sl++) {
    if (t) t[sl] = slices_ > 1 ? sT_[size_t(sl)] : time;
    if (dtOut) dtOut[sl] = slices_ > 1 ? sDt_[size_t(sl)] : dt;
  }
}

// One step attempt of a batch as a HIP graph: the launches of an attempt
// depend only on the buffer parities (cur_, th_), on whether the attempt
// re-primes the FSAL register, on the WENO mode and on the selection (KXRCF
// or the learned network), so one graph is captured
// per such key on its first use and replayed afterwards, with the host-side
// state (parities, mode, counters) advanced as the captured run advanced it.
// KFVM_GRAPH=0, event timing or a failed capture run the attempt eagerly.
template <class F>
void Solver::runAttempt(bool prime, F &&body) {
  if (!graphsOn_ || profOn_) {
    body();
    return;
  }
  const int key = (cur_ & 1) | ((th_ & 1) << 1) | (prime ? 4 : 0) | (useSparseWeno ? 8 : 0) |
                  (learned_ && !forcePlain_ ? 16 : 0);
  AttemptGraph &gr = graphs_[key];
  if (!gr.exec) {
    const int cur0 = cur_, th0 = th_;
    const bool sparse0 = useSparseWeno, full0 = lastRhsFull_;
    const int64_t nRhs0 = nRhsEval;
    hipGraph_t g = nullptr;
    bool ok = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      try {
        body();
      } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      ok = hipStreamEndCapture(stream_, &g) == hipSuccess && g;
    }
    hipGraphExec_t ex = nullptr;
    if (ok) ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    // the captured run advanced the host state without running anything
    gr.cur = cur_;
    gr.th = th_;
    gr.sparse = useSparseWeno;
    gr.full = lastRhsFull_;
    gr.dRhs = nRhsEval - nRhs0;
    cur_ = cur0;
    th_ = th0;
    useSparseWeno = sparse0;
    lastRhsFull_ = full0;
    nRhsEval = nRhs0;
    if (!ok) {  // run eagerly from now on
      std::fprintf(stderr, "warning: HIP graph capture of a batch step failed; running the steps eagerly\n");
      graphsOn_ = false;
      body();
      return;
    }
    nGraphs++;
    gr.exec = ex;
  }
  check(hipGraphLaunch(gr.exec, stream_), "hipGraphLaunch");
  cur_ = gr.cur;
  th_ = gr.th;
  useSparseWeno = gr.sparse;
  lastRhsFull_ = gr.full;
  nRhsEval += gr.dRhs;
}

// One lockstep step of every unfinished slice, each with its own step size
// and controller (the reference's TakeStep, Solver.cpp:188-343, per slice):
// the stages of all slices run in the same launches; slices whose attempt is
// rejected retry in further attempts while the accepted ones are committed.
int Solver::takeStepSlices() {
  const int S = slices_;
  const int cps = int(ncell_ / S);
  const bool fixed = cfg_.fixedDt > 0.0;
  const double dmin = cfg_.dmin(), nDofs = double(cps) * nv_;
  std::vector<char> pending(static_cast<size_t>(S)), firstUnphys(static_cast<size_t>(S), 1);
  for (int sl = 0; sl < S; sl++) {
    pending[size_t(sl)] = !sDone_[size_t(sl)];
    if (!pending[size_t(sl)]) continue;
    double &d = sDt_[size_t(sl)];
    if (fixed) d = cfg_.fixedDt;
    if (sT_[size_t(sl)] + d > cfg_.finalTime) {
      d = cfg_.finalTime - sT_[size_t(sl)];
      sLast_[size_t(sl)] = 1;
    }
  }
  int status = ACCEPTED;
  stateInCur_ = false;
  stagePos_ = 0;
  const int blocksPerSlice = nPart_ / S;
  const bool hydro = nv_ == 5;
  // slices accepted in this step, and slices whose next evaluation runs WENO
  // on every cell (after an unphysical attempt)
  std::vector<char> acceptedNow(static_cast<size_t>(S), 0), forceFull(static_cast<size_t>(S), 0);
  const std::vector<char> doneBefore(sDone_);
  double *hDt = sHost_, *hRed = sHost_ + S, *hVm = sHost_ + 3 * S;
  for (int nT = 0; nT < cfg_.rejectionLimit; nT++) {
    if (std::none_of(pending.begin(), pending.end(), [](char p) { return p != 0; })) break;
    for (int sl = 0; sl < S; sl++) {
      hDt[sl] = pending[size_t(sl)] ? sDt_[size_t(sl)] : 0.0;
      sMaskHost_[sl] = pending[size_t(sl)] ? 1 : 0;
    }
    check(hipMemcpyAsync(sDtDev_, hDt, size_t(S) * sizeof(double), hipMemcpyHostToDevice, stream_), "slice dt");
    check(hipMemcpyAsync(sMask_, sMaskHost_, size_t(S) * sizeof(int), hipMemcpyHostToDevice, stream_), "mask");
    if (nT > 0) {
      // Euler: a slice whose last attempt was unphysical gets WENO on every
      // cell in this attempt's first evaluation (Solver.cpp:298, 437-443) by
      // a previous theta of 0 on all its cells, which the selectors flag; the
      // other slices keep their own selection
      for (int sl = 0; sl < S; sl++)
        if (forceFull[size_t(sl)]) {
          check(launch_fill(theta_[th_] + size_t(sl) * cps, 0.0, size_t(cps), stream_), "force full WENO");
          forceFull[size_t(sl)] = 0;
        }
    }
    // the attempt's device work, from clearing the speed maxima to the host
    // copies of the per-slice reductions
    const bool prime = nT > 0;
    auto attempt = [&] {
      check(launch_zero(vmax_, size_t(S) * sizeof(unsigned long long), stream_), "clear vmax");
      if (prime) {  // the retrying slices' FSAL register, from their step start state
        RKArgs a{};
        a.mode = RK_PRIME;
        a.K = K_;
        a.vmaxBits = vmax_;
        a.sliceWrite = sMask_;
        pipeline(Uprev_, a);
      }
      timed("stage_pre", [&] {
        return launch_stage_pre_slices(nv_, Uprev_, K_, U_[cur_], Uhat_, Utmp_, sDtDev_, rk_.beta[0], rk_.bhat[0],
                                       int(ncell_), cps, stream_);
      });
      for (int st = 1; st < rk_.nStages; st++) {
        RKArgs a{};
        a.mode = RK_STAGE;
        a.g1 = rk_.gam1[st];
        a.g2 = rk_.gam2[st];
        a.g3 = rk_.gam3[st];
        a.delta = rk_.delta[st];
        a.sliceDt = sDtDev_;
        a.beta = rk_.beta[st];
        a.bhat = rk_.bhat[st];
        a.Uprev = Uprev_;
        a.Uhat = Uhat_;
        a.Utmp = Utmp_;
        a.Unext = U_[cur_ ^ 1];
        a.vmaxBits = vmax_;
        pipeline(U_[cur_], a);
        cur_ ^= 1;
      }
      {
        RKArgs a{};
        a.mode = RK_FSAL;
        a.sliceDt = sDtDev_;
        a.bhat = rk_.bhatfsal;
        a.atol = cfg_.atol;
        a.rtol = cfg_.rtol;
        a.K = K_;
        a.Uhat = Uhat_;
        a.errPartial = errPart_;
        a.posPartial = posPart_;
        a.vmaxBits = vmax_;
        a.sliceWrite = sMask_;
        pipeline(U_[cur_], a);
      }
      check(launch_finalize_slices(errPart_, posPart_, blocksPerSlice, S, sRed_, stream_), "finalize");
      check(hipMemcpyAsync(hRed, sRed_, 2 * size_t(S) * sizeof(double), hipMemcpyDeviceToHost, stream_), "read");
      check(hipMemcpyAsync(hVm, vmax_, size_t(S) * sizeof(double), hipMemcpyDeviceToHost, stream_), "read vmax");
      check(hipMemcpyAsync(&hostRed_[3], counter_, sizeof(int), hipMemcpyDeviceToHost, stream_), "read count");
    };
    runAttempt(prime, attempt);
    check(hipStreamSynchronize(stream_), "sync");
    int cnt;
    std::memcpy(&cnt, &hostRed_[3], sizeof(int));
    lastNWeno = lastRhsFull_ ? ncell_ : cnt;
    for (int sl = 0; sl < S; sl++) {
      sMaskHost_[S + sl] = 0;
      if (!pending[size_t(sl)]) continue;
      const double errSum = hRed[2 * sl], posFlag = hRed[2 * sl + 1], maxVel = hVm[sl];
      double &d = sDt_[size_t(sl)];
      const double errNew = 1.0 / std::sqrt(errSum / nDofs);
      double dtfac = std::pow(errNew, rk_.ep1) * std::pow(sErr_[size_t(sl)], rk_.ep2);
      dtfac = std::isnan(dtfac) ? 1.2 : 1.0 + std::atan(dtfac - 1.0);
      if (posFlag < 0.0) {
        d = firstUnphys[size_t(sl)] ? std::fmin(0.25 * cfg_.cfl * dmin / maxVel, d / 4.0) : d / 2.0;
        firstUnphys[size_t(sl)] = 0;
        if (hydro)
          forceFull[size_t(sl)] = 1;
        else
          useSparseWeno = false;  // linear advection: every slice's next evaluation
        nRejectUnphys++;
        sLast_[size_t(sl)] = 0;
      } else if (dtfac >= cfg_.rejectionThresh || fixed) {
        sT_[size_t(sl)] += d;
        d = fixed ? cfg_.fixedDt : std::fmin(d * dtfac, cfg_.cfl * dmin / maxVel);
        sErr_[size_t(sl)] = errNew;
        sMaskHost_[S + sl] = 1;
        acceptedNow[size_t(sl)] = 1;
        pending[size_t(sl)] = 0;
        if (sLast_[size_t(sl)]) sDone_[size_t(sl)] = 1;
      } else {
        d = d * dtfac;
        nRejectThresh++;
        sLast_[size_t(sl)] = 0;
      }
    }
    check(hipMemcpyAsync(sMask_, sMaskHost_ + S, size_t(S) * sizeof(int), hipMemcpyHostToDevice, stream_),
          "mask");
    check(launch_commit_slices(U_[cur_], Uprev_, sMask_, nv_, int(ncell_), cps, stream_), "commit");
    // the accepted slices' theta and WENO flags of their last evaluation: the
    // retries of other slices re-evaluate every slice and would overwrite them
    check(launch_commit_slices(theta_[th_], thBak_, sMask_, 1, int(ncell_), cps, stream_), "keep theta");
    check(launch_commit_slices(wflag_, wfBak_, sMask_, 1, int(ncell_), cps, stream_), "keep flags");
  }
  // slices still rejected after the limit stop with their last attempt's
  // state, theta and flags, as a single run does (stateInCur_)
  if (std::any_of(pending.begin(), pending.end(), [](char p) { return p != 0; })) {
    for (int sl = 0; sl < S; sl++) sMaskHost_[S + sl] = pending[size_t(sl)];
    check(hipMemcpyAsync(sMask_, sMaskHost_ + S, size_t(S) * sizeof(int), hipMemcpyHostToDevice, stream_),
          "mask");
    check(launch_commit_slices(U_[cur_], Uprev_, sMask_, nv_, int(ncell_), cps, stream_), "commit");
    check(launch_commit_slices(theta_[th_], thBak_, sMask_, 1, int(ncell_), cps, stream_), "keep theta");
    check(launch_commit_slices(wflag_, wfBak_, sMask_, 1, int(ncell_), cps, stream_), "keep flags");
  }
  // every slice accepted in this step, stopped at the rejection limit or
  // finished before it keeps the theta and flags of its own last evaluation
  // (the finished ones are re-evaluated with a zero step while the others go on)
  for (int sl = 0; sl < S; sl++)
    acceptedNow[size_t(sl)] = acceptedNow[size_t(sl)] || doneBefore[size_t(sl)] || pending[size_t(sl)];
  if (std::any_of(acceptedNow.begin(), acceptedNow.end(), [](char a) { return a != 0; })) {
    for (int sl = 0; sl < S; sl++) sMaskHost_[2 * S + sl] = acceptedNow[size_t(sl)];
    check(hipMemcpyAsync(sMask_, sMaskHost_ + 2 * S, size_t(S) * sizeof(int), hipMemcpyHostToDevice, stream_),
          "mask");
    check(launch_commit_slices(thBak_, theta_[th_], sMask_, 1, int(ncell_), cps, stream_), "restore theta");
    check(launch_commit_slices(wfBak_, wflag_, sMask_, 1, int(ncell_), cps, stream_), "restore flags");
  }
  // slices still rejected after the limit stop where they are
  for (int sl = 0; sl < S; sl++)
    if (pending[size_t(sl)]) {
      sDone_[size_t(sl)] = 1;
      status = REJECT_LIMIT;
    }
  nTS++;
  lastTimeStep = std::all_of(sDone_.begin(), sDone_.end(), [](char x) { return x != 0; });
  time = 0.0;
  dt = 0.0;
  bool first = true;
  for (int sl = 0; sl < S; sl++)
    if (!sDone_[size_t(sl)] || first) {
      time = first ? sT_[size_t(sl)] : std::fmin(time, sT_[size_t(sl)]);
      dt = first ? sDt_[size_t(sl)] : std::fmin(dt, sDt_[size_t(sl)]);
      first = false;
    }
  if (lastTimeStep) time = *std::min_element(sT_.begin(), sT_.end());
  check(hipStreamSynchronize(stream_), "sync");
  return status;
}

std::string Solver::writePlot() {
  if (!writer_) writer_.reset(new Writer(cfg_));
  std::vector<double> U(static_cast<size_t>(ncell_) * nv_), F(2 * static_cast<size_t>(ncell_));
  getState(U.data());
  getFlags(F.data());
  OutputFields f;
  f.U = U.data();
  f.flags = F.data();
  return writer_->writePlot(f, plotNum, time);
}

std::string Solver::writeCkpt() {
  if (!writer_) writer_.reset(new Writer(cfg_));
  std::vector<double> U(static_cast<size_t>(ncell_) * nv_), F(2 * static_cast<size_t>(ncell_));
  getState(U.data());
  getFlags(F.data());
  OutputFields f;
  f.U = U.data();
  f.flags = F.data();
  return writer_->writeCkpt(f, int(nTS), time, dt, errEst);
}

Solver::~Solver() { release(); }

void Solver::release() noexcept {
  try {
    profileEnable(false);
  } catch (...) {
  }
  void *bufs[] = {SRC_, cellW_, cellWS_, diffM_, chBits_, U_[0], U_[1], Uprev_, Uhat_, K_, Utmp_, FS_, WI_, props_, theta_[0], theta_[1], wflag_, list_, counter_, ppList_,
                  faceW_, derivW_, faceWT_, faceWS_, derivWS_, errPart_, posPart_, redOut_, vmax_, netP_, nadCount_, FSlin_, nadFlag_, sDtDev_, sRed_, sMask_, WS_, pos_, SB_, faceWQ_, thBak_, wfBak_, faceWC_, spd_, mhdWT_, statsPart_};
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto &kv : graphs_)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  graphs_.clear();
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (hostRed_) (void)hipHostFree(hostRed_);
  if (sHost_) (void)hipHostFree(sHost_);
  if (sMaskHost_) (void)hipHostFree(sMaskHost_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

template <class F>
void Solver::timed(const char *name, F &&launch) {
  if (!profOn_) {
    check(launch(), name);
    return;
  }
  hipEvent_t a, b;
  check(hipEventCreate(&a), "hipEventCreate");
  check(hipEventCreate(&b), "hipEventCreate");
  check(hipEventRecord(a, stream_), "hipEventRecord");
  check(launch(), name);
  check(hipEventRecord(b, stream_), "hipEventRecord");
  prof_[name].pending.emplace_back(a, b);
}

void Solver::harvest() {
  check(hipStreamSynchronize(stream_), "sync");
  for (auto &kv : prof_) {
    for (auto &ev : kv.second.pending) {
      float ms = 0.f;
      check(hipEventElapsedTime(&ms, ev.first, ev.second), "hipEventElapsedTime");
      kv.second.ms += ms;
      kv.second.count++;
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    kv.second.pending.clear();
  }
}

void Solver::profileEnable(bool on) {
  if (!on && profOn_) harvest();
  profOn_ = on;
}
void Solver::profileReset() {
  harvest();
  prof_.clear();
}
bool Solver::profileGet(const std::string &name, double &ms, int64_t &count) {
  harvest();
  auto it = prof_.find(name);
  if (it == prof_.end()) {
    ms = 0.0;
    count = 0;
    return false;
  }
  ms = it->second.ms;
  count = it->second.count;
  return true;
}

// GLM-MHD evaluation (Solver.cpp:410-530 with sources): reconstruction of
// face and interior states with the limiter's fast test, selection, WENO
// (converted back and limited), the limiter's exact sequence for the other
// cells that failed the test, the cleaning speed, then the Riemann /
// divergence / source-term / RK kernel.
void Solver::pipelineMhd(const double *Uin, const RKArgs &a0) {
  const int dim = cfg_.dim, nq = cfg_.nQuad;
  const double *thPrev = theta_[th_];
  double *thNew = theta_[th_ ^ 1];
  lastRhsFull_ = !useSparseWeno;
  // counter_: [0] WENO cells, [1] cells of the linear pass for the limiter's exact sequence
  check(launch_zero(counter_, 2 * sizeof(int), stream_), "clear counters");
  timed("cell_props", [&] { return launch_mhd_props(dim, Uin, props_, grid_, stream_); });
  if (useSparseWeno) {
    const bool lrn = learned_ && !forcePlain_;
    timed("linear", [&] {
      return launch_mhd_linear(dim, nq, Uin, FS_, SRC_, lrn ? nullptr : WI_, thNew, spd_, props_, grid_, mhdWT_,
                               ppList_, counter_ + 1, stream_);
    });
    if (lrn) {
      timed("select", [&] {
        return launch_learned_select(dim, 9, Uin, nullptr, thPrev, nadFlag_, netP_, netThr_, wflag_, list_,
                                     counter_, grid_, nullptr, 0, stream_);
      });
    } else {
      const double wThresh = cfg_.wenoThresh * std::pow(cfg_.dmin(), 1.5);
      timed("select", [&] {
        return launch_mhd_select(dim, nq, Uin, WI_, thPrev, wflag_, list_, counter_, wThresh, grid_, stream_);
      });
    }
    timed("weno", [&] {
      return launch_mhd_weno(dim, nq, Uin, FS_, SRC_, grid_, faceWS_, derivWS_, cellWS_, list_, counter_, thNew,
                             spd_, props_, faceWC_, stream_);
    });
    timed("pp", [&] {
      return launch_mhd_pp(dim, nq, Uin, FS_, SRC_, thNew, spd_, props_, wflag_, grid_, ppList_, counter_ + 1,
                           stream_);
    });
  } else {
    timed("weno", [&] {
      return launch_mhd_weno(dim, nq, Uin, FS_, SRC_, grid_, faceWS_, derivWS_, cellWS_, nullptr, nullptr, thNew,
                             spd_, props_, faceWC_, stream_);
    });
    useSparseWeno = cfg_.wenoThresh > 0.0;
  }
  th_ ^= 1;
  check(launch_zero(chBits_, 2 * sizeof(unsigned long long), stream_), "clear speeds");
  timed("chmax", [&] { return launch_mhd_chmax(dim, spd_, grid_, chBits_, stream_); });
  RKArgs a = a0;
  a.chBits = chBits_;
  a.src = SRC_;
  a.theta = thNew;
  a.Dm = diffM_;
  timed("flux_rk", [&] { return launch_flux_rk(dim, nq, nv_, Uin, FS_, grid_, a, stream_, cfg_.fluxBrick); });
  nRhsEval++;
}

void Solver::pipeline(const double *Uin, const RKArgs &a) {
  Region region("Solver::evalRHS");
  if (nv_ == 9) return pipelineMhd(Uin, a);
  const int dim = cfg_.dim, nq = cfg_.nQuad;
  const bool hydro = nv_ == 5;
  // theta of the previous RHS evaluation feeds the selector; this one writes the other buffer
  const double *thPrev = theta_[th_];
  double *thNew = theta_[th_ ^ 1];
  // counter_: [0] WENO cells, [1] cells of the linear pass for the limiter's
  // exact sequence (ppList_)
  if (hydro) {
    check(launch_zero(counter_, 2 * sizeof(int), stream_), "clear counters");
    timed("cell_props", [&] { return launch_cell_props(dim, Uin, props_, grid_, stream_); });
  }
  lastRhsFull_ = !useSparseWeno;
  if (useSparseWeno && fused_) {
    // linear pass (fast limiter test, KXRCF indicator, brick-boundary states)
    // -> selection -> WENO and the limiter's exact sequence into rows ->
    // fused reconstruction + Riemann + divergence + RK update
    const bool lrn = learned_ && !forcePlain_;
    check(launch_zero(counter_ + 2, sizeof(int), stream_), "clear row count");
    timed("linear", [&] {
      return launch_linear_surf(dim, nq, Uin, lrn ? nullptr : WI_, thNew, props_, grid_, faceWT_, ppList_,
                                counter_ + 1, SB_, sb_, stream_);
    });
    if (lrn) {
      timed("select", [&] {
        return launch_learned_select(dim, 5, Uin, props_, thPrev, nadFlag_, netP_, netThr_, wflag_, list_, counter_,
                                     grid_, pos_, wsCap_, stream_);
      });
    } else {
      const double wThresh = cfg_.wenoThresh * std::pow(cfg_.dmin(), 1.5);
      timed("select", [&] {
        return launch_select(dim, nq, Uin, WI_, thPrev, wflag_, list_, counter_, wThresh, grid_, pos_, wsCap_,
                             stream_);
      });
    }
    timed("weno", [&] {
      return launch_weno(dim, nq, nv_, Uin, FS_, thNew, grid_, faceW_, derivW_, faceWC_, derivWS_, list_, counter_,
                         FS_, wsCap_, stream_);
    });
    timed("weno_post", [&] {
      return launch_weno_post(dim, nq, Uin, FS_, thNew, props_, grid_, list_, counter_, FS_, wsCap_, stream_);
    });
    timed("pp", [&] {
      return launch_pp_rows(dim, nq, Uin, FS_, pos_, thNew, props_, wflag_, grid_, faceW_, ppList_, counter_ + 1,
                            counter_, counter_ + 2, stream_);
    });
    th_ ^= 1;
    timed("flux_rk", [&] {
      return launch_flux_fused(dim, nq, Uin, faceWQ_, FS_, pos_, SB_, sb_, grid_, a, stream_);
    });
    nRhsEval++;
    return;
  }
  if (useSparseWeno) {
    // the learned selection does not read the KXRCF indicator; for the Euler
    // equations it runs inside the linear pass, on the brick that pass stages
    const bool lrn = learned_ && !forcePlain_ && hydro;
    double *wi = lrn ? nullptr : WI_;
    LearnArgs la;
    if (lrn) {
      la.eint = props_;
      la.theta = thPrev;
      la.nadFlag = nadFlag_;
      la.P = netP_;
      la.thr = netThr_;
      la.wflag = wflag_;
      la.list = list_;
      la.counter = counter_;
      la.pos = pos_;
      la.cap = wsCap_;
    }
    timed("linear", [&] {
      return launch_linear(dim, nq, nv_, Uin, FS_, wi, thNew, props_, grid_, faceW_, faceWT_, ppList_, counter_ + 1,
                           stream_, lrn ? &la : nullptr);
    });
    if (hydro) {
      if (!lrn) {
        const double wThresh = cfg_.wenoThresh * std::pow(cfg_.dmin(), 1.5);
        timed("select", [&] {
          return launch_select(dim, nq, Uin, WI_, thPrev, wflag_, list_, counter_, wThresh, grid_, pos_, wsCap_,
                               stream_);
        });
      }
      timed("weno", [&] {
        return launch_weno(dim, nq, nv_, Uin, FS_, thNew, grid_, faceW_, derivW_, faceWC_, derivWS_, list_,
                           counter_, WS_, wsCap_, stream_);
      });
      timed("weno_post", [&] {
        return launch_weno_post(dim, nq, Uin, FS_, thNew, props_, grid_, list_, counter_, WS_, wsCap_, stream_);
      });
      timed("pp", [&] {
        return launch_pp_lanes(dim, nq, Uin, FS_, thNew, props_, wflag_, grid_, ppList_, counter_ + 1, stream_);
      });
    }
  } else {
    timed("weno", [&] {
      return launch_weno(dim, nq, nv_, Uin, FS_, thNew, grid_, faceW_, derivW_, faceWC_, derivWS_, nullptr,
                         nullptr, nullptr, 0, stream_);
    });
    if (hydro)
      timed("weno_post", [&] {
        return launch_weno_post(dim, nq, Uin, FS_, thNew, props_, grid_, nullptr, nullptr, nullptr, 0, stream_);
      });
    useSparseWeno = cfg_.wenoThresh > 0.0;
  }
  if (hydro) th_ ^= 1;
  // listed Euler cells: their WENO states stay in the compact rows of WS
  RKArgs ar = a;
  if (hydro && !lastRhsFull_) {
    ar.pos = pos_;
    ar.ws = WS_;
  }
  timed("flux_rk", [&] { return launch_flux_rk(dim, nq, nv_, Uin, FS_, grid_, ar, stream_, cfg_.fluxBrick); });
  nRhsEval++;
}

double Solver::readSpeed() {
  check(hipMemcpyAsync(&hostRed_[2], vmax_, sizeof(double), hipMemcpyDeviceToHost, stream_), "read vmax");
  check(hipStreamSynchronize(stream_), "sync");
  double v;
  std::memcpy(&v, &hostRed_[2], sizeof(double));
  return v;
}

double Solver::evalRHS(int which) {
  check(launch_zero(vmax_, size_t(slices_) * sizeof(unsigned long long), stream_), "clear vmax");
  RKArgs a{};
  a.mode = RK_PRIME;
  a.K = K_;
  a.vmaxBits = vmax_;
  const double *Uin = which == 1 ? Uprev_ : (stateInCur_ ? U_[cur_] : Uprev_);
  pipeline(Uin, a);
  return readSpeed();
}

size_t Solver::carveEval(unsigned char *base, EvalBufs &b) const {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    off = (off + 255) / 256 * 256;
    void *p = base ? base + off : nullptr;
    off += bytes;
    return p;
  };
  const size_t nc = size_t(ncell_), FQ = size_t(wts_.F) * wts_.Q, D = sizeof(double), I = sizeof(int);
  b.FS = static_cast<double *>(take(fs_elems(nc, int(FQ) * nv_) * D));
  b.WS = WS_ ? static_cast<double *>(take(size_t(wsCap_) * FQ * nv_ * D)) : nullptr;
  if (SB_) {
    Surf sb{};
    b.SB = static_cast<double *>(take(surf_elems(cfg_.dim, cfg_.nQuad, grid_, sb) * D));
  }
  b.WI = static_cast<double *>(take(fs_elems(nc, int(FQ)) * D));
  b.props = static_cast<double *>(take(nc * 2 * D));
  b.thNew = static_cast<double *>(take(nc * D));
  b.wflag = static_cast<double *>(take(nc * D));
  b.SRC = SRC_ ? static_cast<double *>(take(nc * size_t(wts_.nqCell) * nv_ * D)) : nullptr;
  b.spd = spd_ ? static_cast<double *>(take(nc * 2 * D)) : nullptr;
  b.list = static_cast<int *>(take(nc * I));
  b.counter = static_cast<int *>(take(4 * I));
  b.ppList = static_cast<int *>(take(nc * I));
  b.pos = pos_ ? static_cast<int *>(take(nc * I)) : nullptr;
  b.chBits = chBits_ ? static_cast<unsigned long long *>(take(2 * sizeof(unsigned long long))) : nullptr;
  b.vmax = static_cast<unsigned long long *>(take(size_t(slices_) * sizeof(unsigned long long)));
  return off;
}

size_t Solver::noAllocBytes() const {
  EvalBufs b;
  return carveEval(nullptr, b);
}

double Solver::evalRhsNoAlloc(const double *U, double *rhs) {
  const size_t need = noAllocBytes();
  if (need > kNoAllocScratchBytes)
    throw std::runtime_error("this grid needs " + std::to_string(need) + " bytes of scratch, more than the " +
                             std::to_string(kNoAllocScratchBytes) + " of the no-allocation entry point");
  unsigned char *base = static_cast<unsigned char *>(noalloc_scratch());
  EvalBufs b;
  carveEval(base, b);
  // the solver's own buffers and evaluation state, put back on every exit
  struct Restore {
    Solver &s;
    EvalBufs keep;
    int th;
    bool sparse, lastFull, prof;
    int64_t nRhs;
    ~Restore() {
      s.FS_ = keep.FS;
      s.WS_ = keep.WS;
      s.SB_ = keep.SB;
      s.WI_ = keep.WI;
      s.props_ = keep.props;
      s.theta_[th ^ 1] = keep.thNew;
      s.wflag_ = keep.wflag;
      s.SRC_ = keep.SRC;
      s.spd_ = keep.spd;
      s.list_ = keep.list;
      s.counter_ = keep.counter;
      s.ppList_ = keep.ppList;
      s.pos_ = keep.pos;
      s.chBits_ = keep.chBits;
      s.vmax_ = keep.vmax;
      s.th_ = th;
      s.useSparseWeno = sparse;
      s.lastRhsFull_ = lastFull;
      s.profOn_ = prof;
      s.nRhsEval = nRhs;
    }
  } restore{*this,
            {FS_, WS_, SB_, WI_, props_, theta_[th_ ^ 1], wflag_, SRC_, spd_, list_, counter_, ppList_, pos_,
             chBits_, vmax_},
            th_,
            useSparseWeno,
            lastRhsFull_,
            profOn_,
            nRhsEval};
  FS_ = b.FS;
  WS_ = b.WS;
  SB_ = b.SB;
  WI_ = b.WI;
  props_ = b.props;
  theta_[th_ ^ 1] = b.thNew;
  wflag_ = b.wflag;
  SRC_ = b.SRC;
  spd_ = b.spd;
  list_ = b.list;
  counter_ = b.counter;
  ppList_ = b.ppList;
  pos_ = b.pos;
  chBits_ = b.chBits;
  vmax_ = b.vmax;
  profOn_ = false;  // event timing creates events
  check(launch_zero(vmax_, size_t(slices_) * sizeof(unsigned long long), stream_), "clear vmax");
  RKArgs a{};
  a.mode = RK_PRIME;
  a.K = rhs;
  a.vmaxBits = vmax_;
  pipeline(U, a);
  return readSpeed();
}

int Solver::takeStep() {
  Region region("Solver::TakeStep");
  if (slices_ > 1) return takeStepSlices();
  const bool fixed = cfg_.fixedDt > 0.0;
  if (fixed) dt = cfg_.fixedDt;
  // limit dt to hit plot times or the final time (Solver.cpp:197-205)
  if (time + dt > cfg_.finalTime) {
    dt = cfg_.finalTime - time;
    lastTimeStep = true;
    writePlot_ = true;
  } else if (cfg_.plotFreq > 0.0 && time + dt > plotNum * cfg_.plotFreq) {
    dt = plotNum * cfg_.plotFreq - time;
    writePlot_ = true;
  }
  const double dmin = cfg_.dmin();
  const double nDofs = double(ncell_) * nv_;
  bool accepted = false, firstUnphys = true;
  int status = ACCEPTED;
  stateInCur_ = false;
  stagePos_ = 0;
  for (int nT = 0; nT < cfg_.rejectionLimit; nT++) {
    check(launch_zero(vmax_, sizeof(unsigned long long), stream_), "clear vmax");
    if (nT > 0) {
      RKArgs a{};
      a.mode = RK_PRIME;
      a.K = K_;
      a.vmaxBits = vmax_;
      pipeline(Uprev_, a);
    }
    timed("stage_pre", [&] {
      return launch_stage_pre(nv_, Uprev_, K_, U_[cur_], Uhat_, Utmp_, rk_.beta[0] * dt,
                              rk_.bhat[0] * dt, int(ncell_), stream_);
    });
    for (int s = 1; s < rk_.nStages; s++) {
      RKArgs a{};
      a.mode = RK_STAGE;
      a.g1 = rk_.gam1[s];
      a.g2 = rk_.gam2[s];
      a.g3 = rk_.gam3[s];
      a.delta = rk_.delta[s];
      a.betaDt = rk_.beta[s] * dt;
      a.bhatDt = rk_.bhat[s] * dt;
      a.Uprev = Uprev_;
      a.Uhat = Uhat_;
      a.Utmp = Utmp_;
      a.Unext = U_[cur_ ^ 1];
      a.vmaxBits = vmax_;
      pipeline(U_[cur_], a);
      cur_ ^= 1;
    }
    {
      RKArgs a{};
      a.mode = RK_FSAL;
      a.bhatDt = rk_.bhatfsal * dt;
      a.atol = cfg_.atol;
      a.rtol = cfg_.rtol;
      a.K = K_;
      a.Uhat = Uhat_;
      a.errPartial = errPart_;
      a.posPartial = posPart_;
      a.vmaxBits = vmax_;
      pipeline(U_[cur_], a);
    }
    check(launch_finalize(errPart_, posPart_, nPart_, redOut_, stream_), "finalize");
    check(hipMemcpyAsync(hostRed_, redOut_, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_), "read err");
    check(hipMemcpyAsync(&hostRed_[2], vmax_, sizeof(double), hipMemcpyDeviceToHost, stream_), "read vmax");
    check(hipMemcpyAsync(&hostRed_[3], counter_, sizeof(int), hipMemcpyDeviceToHost, stream_), "read count");
    check(hipStreamSynchronize(stream_), "sync");
    const double errSum = hostRed_[0], posFlag = hostRed_[1];
    double maxVel;
    std::memcpy(&maxVel, &hostRed_[2], sizeof(double));
    int cnt;
    std::memcpy(&cnt, &hostRed_[3], sizeof(int));
    lastNWeno = lastRhsFull_ ? ncell_ : cnt;

    const double errNew = 1.0 / std::sqrt(errSum / nDofs);
    double dtfac = std::pow(errNew, rk_.ep1) * std::pow(errEst, rk_.ep2);
    dtfac = std::isnan(dtfac) ? 1.2 : 1.0 + std::atan(dtfac - 1.0);
    int lStat;
    if (posFlag < 0.0) lStat = UNPHYSICAL;
    else if (dtfac >= cfg_.rejectionThresh || fixed) lStat = ACCEPTED;
    else lStat = TOLERANCE;
    status = lStat;
    if (lStat == UNPHYSICAL) {
      dt = firstUnphys ? std::fmin(0.25 * cfg_.cfl * dmin / maxVel, dt / 4.0) : dt / 2.0;
      firstUnphys = false;
      useSparseWeno = false;
      nRejectUnphys++;
      if (lastTimeStep) {
        lastTimeStep = false;
        writePlot_ = false;
      }
    } else if (lStat == ACCEPTED && learned_ && !forcePlain_ && nadOn_ && nadFails()) {
      // a-posteriori check of the learned selection: redo the step with KXRCF
      forcePlain_ = true;
      nNadFallback++;
    } else if (lStat == ACCEPTED) {
      accepted = true;
      forcePlain_ = false;
      time += dt;
      const double dterr = dt * dtfac, dtcfl = cfg_.cfl * dmin / maxVel;
      dt = fixed ? cfg_.fixedDt : std::fmin(dterr, dtcfl);
      errEst = errNew;
      std::swap(Uprev_, U_[cur_]);  // Uprev <- U
      break;
    } else {
      dt = dt * dtfac;
      nRejectThresh++;
      if (lastTimeStep) {
        lastTimeStep = false;
        writePlot_ = false;
      }
    }
  }
  if (!accepted) {
    lastTimeStep = true;
    status = REJECT_LIMIT;
    stateInCur_ = true;
  }
  if (dt < cfg_.minDt) {
    lastTimeStep = true;
    if (accepted) status = STAGNATED;
  }
  nTS++;
  return status;
}

bool Solver::nadFails() {
  check(launch_zero(nadCount_, sizeof(int), stream_), "clear NAD count");
  timed("nad", [&] {
    return launch_nad(cfg_.dim, nv_, Uprev_, U_[cur_], nadRange_, nadValue_, nadFlag_, nadCount_, grid_, stream_);
  });
  int cnt = 0;
  check(hipMemcpyAsync(&hostRed_[3], nadCount_, sizeof(int), hipMemcpyDeviceToHost, stream_), "read NAD count");
  check(hipStreamSynchronize(stream_), "sync");
  std::memcpy(&cnt, &hostRed_[3], sizeof(int));
  nNadCells += cnt;
  return double(cnt) > nadMaxFrac_ * double(ncell_);
}

void Solver::setIndicator(const float *params, int np, double threshold) {
  check(hipStreamSynchronize(stream_), "sync");
  if (slices_ > 1 && np != 0) throw std::runtime_error("the learned indicator runs on single problems");
  if (np == 0) {
    learned_ = false;
    return;
  }
  if (nv_ != 5 && nv_ != 9)
    throw std::runtime_error("the learned indicator is built for the Euler and GLM-MHD equations");
  if (np != net_params(cfg_.dim)) throw std::runtime_error("indicator parameter count does not match the dimension");
  if (!netP_) netP_ = dalloc<float>(size_t(np), "indicator parameters");
  if (!nadCount_) nadCount_ = dalloc<int>(1, "NAD count");
  if (!nadFlag_) nadFlag_ = dalloc<double>(size_t(ncell_), "NAD flags");
  check(hipMemset(nadFlag_, 0, size_t(ncell_) * sizeof(double)), "clear NAD flags");
  check(hipMemcpy(netP_, params, size_t(np) * sizeof(float), hipMemcpyHostToDevice), "copy indicator");
  netThr_ = float(threshold);
  learned_ = true;
  forcePlain_ = false;
}

void Solver::collectSamples(float *X, float *Y, double tau) {
  if (nv_ != 5 && nv_ != 9)
    throw std::runtime_error("the learned indicator is built for the Euler and GLM-MHD equations");
  const int dim = cfg_.dim, nq = cfg_.nQuad;
  const double *Uin = stateInCur_ ? U_[cur_] : Uprev_;
  double *thScratch = theta_[th_ ^ 1];
  const size_t fsBytes = fs_elems(size_t(ncell_), wts_.F * wts_.Q * nv_) * sizeof(double);
  if (!FSlin_) FSlin_ = dalloc<double>(fsBytes / sizeof(double), "linear face states");
  if (nv_ == 9) {
    // labels compare the linear face states with the limited WENO ones
    check(launch_zero(counter_, 2 * sizeof(int), stream_), "clear counters");
    check(launch_mhd_props(dim, Uin, props_, grid_, stream_), "mhd_props");
    check(launch_mhd_linear(dim, nq, Uin, FS_, SRC_, nullptr, thScratch, spd_, props_, grid_, mhdWT_, ppList_,
                            counter_ + 1, stream_),
          "mhd_linear");
    check(hipMemcpyAsync(FSlin_, FS_, fsBytes, hipMemcpyDeviceToDevice, stream_), "copy linear states");
    check(launch_mhd_weno(dim, nq, Uin, FS_, SRC_, grid_, faceWS_, derivWS_, cellWS_, nullptr, nullptr, thScratch,
                          spd_, props_, faceWC_, stream_),
          "mhd_weno");
    check(launch_labels(dim, nq, 9, Uin, FSlin_, FS_, float(tau), Y, grid_, stream_), "labels");
    check(launch_features(dim, 9, Uin, nullptr, X, grid_, stream_), "features");
    check(hipStreamSynchronize(stream_), "sync");
    return;
  }
  check(launch_zero(counter_, 2 * sizeof(int), stream_), "clear counters");
  check(launch_cell_props(dim, Uin, props_, grid_, stream_), "cell_props");
  check(launch_linear(dim, nq, nv_, Uin, FS_, WI_, thScratch, props_, grid_, faceW_, faceWT_, ppList_, counter_ + 1,
                      stream_),
        "linear");
  check(hipMemcpyAsync(FSlin_, FS_, fsBytes, hipMemcpyDeviceToDevice, stream_), "copy linear states");
  check(launch_weno(dim, nq, nv_, Uin, FS_, thScratch, grid_, faceW_, derivW_, faceWC_, derivWS_, nullptr, nullptr,
                    nullptr, 0, stream_),
        "weno");
  check(launch_weno_post(dim, nq, Uin, FS_, thScratch, props_, grid_, nullptr, nullptr, nullptr, 0, stream_),
        "weno_post");
  check(launch_labels(dim, nq, 5, Uin, FSlin_, FS_, float(tau), Y, grid_, stream_), "labels");
  check(launch_features(dim, 5, Uin, props_, X, grid_, stream_), "features");
  check(hipStreamSynchronize(stream_), "sync");
}

void Solver::advanceStages(int64_t n, double dtFixed) {
  if (slices_ > 1) throw std::runtime_error("advance_stages is built for single problems");
  stateInCur_ = false;
  for (int64_t it = 0; it < n; it++) {
    if (stagePos_ == 0) {
      check(launch_zero(vmax_, sizeof(unsigned long long), stream_), "clear vmax");
      timed("stage_pre", [&] {
        return launch_stage_pre(nv_, Uprev_, K_, U_[cur_], Uhat_, Utmp_, rk_.beta[0] * dtFixed,
                                rk_.bhat[0] * dtFixed, int(ncell_), stream_);
      });
      stagePos_ = 1;
    }
    RKArgs a{};
    a.vmaxBits = vmax_;
    if (stagePos_ < rk_.nStages) {
      const int s = stagePos_;
      a.mode = RK_STAGE;
      a.g1 = rk_.gam1[s];
      a.g2 = rk_.gam2[s];
      a.g3 = rk_.gam3[s];
      a.delta = rk_.delta[s];
      a.betaDt = rk_.beta[s] * dtFixed;
      a.bhatDt = rk_.bhat[s] * dtFixed;
      a.Uprev = Uprev_;
      a.Uhat = Uhat_;
      a.Utmp = Utmp_;
      a.Unext = U_[cur_ ^ 1];
      pipeline(U_[cur_], a);
      cur_ ^= 1;
      stagePos_++;
    } else {
      a.mode = RK_FSAL;
      a.bhatDt = rk_.bhatfsal * dtFixed;
      a.atol = cfg_.atol;
      a.rtol = cfg_.rtol;
      a.K = K_;
      a.Uhat = Uhat_;
      a.errPartial = errPart_;
      a.posPartial = posPart_;
      pipeline(U_[cur_], a);
      check(launch_finalize(errPart_, posPart_, nPart_, redOut_, stream_), "finalize");
      std::swap(Uprev_, U_[cur_]);
      time += dtFixed;
      nTS++;
      stagePos_ = 0;
    }
  }
}

int Solver::flowStats(double *out) {
  if (slices_ > 1) throw std::runtime_error("flow statistics are built for single problems");
  if (nv_ == 1) {  // the reference's LinAdv statistics add nothing (LinAdv_K.H:53-64)
    out[0] = 0.0;
    return 1;
  }
  const Config &c = cfg_;
  const bool d3 = c.dim == 3;
  // dvol / volume of Geometry.H:35-40
  const double volume = d3 ? (c.xUp - c.xLo) * (c.yUp - c.yLo) * (c.zUp - c.zLo) : (c.xUp - c.xLo) * (c.yUp - c.yLo);
  const double dvol = d3 ? grid_.dx * grid_.dy * grid_.dz : grid_.dx * grid_.dy;
  const size_t nb = (size_t(ncell_) + 255) / 256;
  if (!statsPart_) statsPart_ = dalloc<double>(3 * nb, "flow statistics partials");
  const double *U = stateInCur_ ? U_[cur_] : Uprev_;
  check(launch_flow_stats(c.dim, nv_, U, grid_, dvol / volume, statsPart_, stream_), "flow_stats");
  std::vector<double> part(3 * nb);
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(part.data(), statsPart_, part.size() * sizeof(double), hipMemcpyDeviceToHost), "get stats");
  const int n = nv_ == 9 ? 3 : 2;
  for (int st = 0; st < n; st++) {
    double sum = 0.0;
    for (size_t b = 0; b < nb; b++) sum += part[3 * b + size_t(st)];
    out[st] = sum;
  }
  return n;
}

// One row of <basename>_stats.h5 (Solver.cpp:154-186, WriterPDI.cpp:69-86).
void Solver::evalFlowStats() {
  Region region("Solver::evalFlowStats");
  double st[3];
  const int n = flowStats(st);
  if (!writer_) writer_.reset(new Writer(cfg_));
  writer_->writeStats(statsRow_, cfg_.maxSteps, time, st, n);
  statsRow_++;
}

int64_t Solver::solve() {
  // Solver::Solve (Solver.cpp:84-140): step, then the outputs that are due
  Region region("Solver::Solve");
  if (cfg_.statsFreq > 0) evalFlowStats();
  int64_t n = 0;
  for (; nTS < cfg_.maxSteps && !lastTimeStep; n++) {
    const int64_t step = nTS;  // the reference advances its counter after the outputs
    takeStep();
    if (cfg_.statsFreq > 0 && step % cfg_.statsFreq == 0) evalFlowStats();
    const bool last = lastTimeStep || step == cfg_.maxSteps - 1;
    if (cfg_.plotFreq > 0.0 && (writePlot_ || last)) {
      writePlot_ = false;
      writePlot();
      plotNum++;
    }
    if (cfg_.ckptFreq > 0 && (step % cfg_.ckptFreq == 0 || last)) {
      const int64_t keep = nTS;
      nTS = step;
      writeCkpt();
      nTS = keep;
    }
    if (lastTimeStep && cfg_.statsFreq > 0) evalFlowStats();
  }
  return n;
}

void Solver::getState(double *host) const {
  const double *src = stateInCur_ ? U_[cur_] : Uprev_;
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(host, src, size_t(ncell_) * nv_ * sizeof(double), hipMemcpyDeviceToHost), "get state");
}
void Solver::setState(const double *host) {
  const size_t bytes = size_t(ncell_) * nv_ * sizeof(double);
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(Uprev_, host, bytes, hipMemcpyHostToDevice), "set state");
  check(hipMemcpy(U_[cur_], host, bytes, hipMemcpyHostToDevice), "set state");
  stateInCur_ = false;
}
void Solver::getRHS(double *host) const {
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(host, K_, size_t(ncell_) * nv_ * sizeof(double), hipMemcpyDeviceToHost), "get rhs");
}
void Solver::getFlags(double *host) const {
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(host, wflag_, size_t(ncell_) * sizeof(double), hipMemcpyDeviceToHost), "get flags");
  check(hipMemcpy(host + ncell_, theta_[th_], size_t(ncell_) * sizeof(double), hipMemcpyDeviceToHost), "get flags");
}
void Solver::setFlags(const double *host) {
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(wflag_, host, size_t(ncell_) * sizeof(double), hipMemcpyHostToDevice), "set flags");
  check(hipMemcpy(theta_[th_], host + ncell_, size_t(ncell_) * sizeof(double), hipMemcpyHostToDevice), "set flags");
}
void Solver::setWeights(const double *face, const double *deriv) {
  std::copy(face, face + wts_.face.size(), wts_.face.begin());
  std::copy(deriv, deriv + wts_.deriv.size(), wts_.deriv.begin());
  check(hipStreamSynchronize(stream_), "sync");
  uploadWeights();
}
// Device copies of the weight tables: the reference layouts
// faceWeights[nSub][F][Q][C] and derivWeights[nSub][I][C]; the full-stencil
// face weights regrouped per pass of P face points as [F*Q/P][C][P] for the
// linear reconstruction; and stencil-major copies [nSub][C][F*Q] and
// [nSub][C][I] (and [nSub][C][NQ^D] for GLM-MHD's interior points) for the
// WENO reconstruction.  The Euler kernels walk the
// stencil in their outer loop, so one stencil cell's weights are contiguous.
void Solver::uploadWeights() {
  const int C = wts_.C, FQ = wts_.F * wts_.Q, I = wts_.I, nSub = wts_.nSub;
  const int P = linear_pass_points(wts_.dim, wts_.nq);
  std::vector<double> wt(size_t(FQ) * C), fs(wts_.face.size()), ds(wts_.deriv.size());
  for (int ps = 0; ps < FQ / P; ps++)
    for (int n = 0; n < C; n++)
      for (int q = 0; q < P; q++) wt[(size_t(ps) * C + n) * P + q] = wts_.face[size_t(ps * P + q) * C + n];
  for (int sb = 0; sb < nSub; sb++)
    for (int n = 0; n < C; n++) {
      for (int fq = 0; fq < FQ; fq++)
        fs[(size_t(sb) * C + n) * FQ + fq] = wts_.face[(size_t(sb) * FQ + fq) * C + n];
      for (int i = 0; i < I; i++) ds[(size_t(sb) * C + n) * I + i] = wts_.deriv[(size_t(sb) * I + i) * C + n];
    }
  check(hipMemcpy(faceW_, wts_.face.data(), wts_.face.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  check(hipMemcpy(derivW_, wts_.deriv.data(), wts_.deriv.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  check(hipMemcpy(faceWT_, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  check(hipMemcpy(faceWS_, fs.data(), fs.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  check(hipMemcpy(derivWS_, ds.data(), ds.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  if (faceWC_) {
    const int SUB = wts_.dim == 3 ? SInfo<3>::C + SInfo<3>::Cc + SInfo<3>::B * SInfo<3>::Cb
                                  : SInfo<2>::C + SInfo<2>::Cc + SInfo<2>::B * SInfo<2>::Cb;
    const int nCell = nv_ == 9 ? wts_.nqCell : 0;
    std::vector<double> wc(size_t(SUB) * (FQ + I + nCell));
    compact_weno_weights(wts_.dim, FQ, I, wts_.face.data(), wts_.deriv.data(), wc.data(), nCell,
                         nCell ? wts_.cell.data() : nullptr);
    check(hipMemcpy(faceWC_, wc.data(), wc.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  }
  if (faceWQ_) {  // full-stencil face weights per point: [Q][C][F]
    const int Q = wts_.Q, F = wts_.F;
    std::vector<double> wq(size_t(Q) * C * F);
    for (int q = 0; q < Q; q++)
      for (int n = 0; n < C; n++)
        for (int f = 0; f < F; f++) wq[(size_t(q) * C + n) * F + f] = wts_.face[size_t(f * Q + q) * C + n];
    check(hipMemcpy(faceWQ_, wq.data(), wq.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  }
  if (cellW_) {
    const int NI = wts_.nqCell;
    std::vector<double> cs(wts_.cell.size());
    for (int sb = 0; sb < nSub; sb++)
      for (int n = 0; n < C; n++)
        for (int p = 0; p < NI; p++) cs[(size_t(sb) * C + n) * NI + p] = wts_.cell[(size_t(sb) * NI + p) * C + n];
    check(hipMemcpy(cellW_, wts_.cell.data(), wts_.cell.size() * sizeof(double), hipMemcpyHostToDevice),
          "copy weights");
    check(hipMemcpy(cellWS_, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  }
  if (mhdWT_) {  // full-stencil face passes [FQ/P][C][P], then interior passes [NI/PI][C][PI]
    const int NI = wts_.nqCell, PI = mhd_interior_pass_points(wts_.dim, wts_.nq);
    std::vector<double> mw(size_t(FQ + NI) * C);
    std::copy(wt.begin(), wt.end(), mw.begin());
    double *wi = mw.data() + size_t(FQ) * C;
    for (int ps = 0; ps < NI / PI; ps++)
      for (int n = 0; n < C; n++)
        for (int q = 0; q < PI; q++) wi[(size_t(ps) * C + n) * PI + q] = wts_.cell[size_t(ps * PI + q) * C + n];
    check(hipMemcpy(mhdWT_, mw.data(), mw.size() * sizeof(double), hipMemcpyHostToDevice), "copy weights");
  }
}
void Solver::sync() const { check(hipStreamSynchronize(stream_), "sync"); }

}  // namespace kfvm
