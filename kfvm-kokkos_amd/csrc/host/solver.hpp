// Host driver of the stage library: owns the device state of one problem and
// runs the RK-FSAL step controller around the HIP stage kernels.
//
//   Solver::Solver   src/Solver.cpp:36-81      Solver::Solve    src/Solver.cpp:84-140
//   Solver::TakeStep src/Solver.cpp:188-343    Solver::evalRHS  src/Solver.cpp:410-530
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include <memory>

#include "../device/device_api.hpp"
#include "common.hpp"
#include "weights.hpp"
#include "writer.hpp"

namespace kfvm {

class Solver {
 public:
  explicit Solver(const Config &cfg);
  ~Solver();
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;

  // Solver::evalRHS on the current (which=0) or step-start (which=1) state;
  // stores the FSAL register and returns the max wave speed.
  double evalRHS(int which);
  // evalRHS of the device state U (layout of the solution registers) into
  // rhs, with every per-cell scratch buffer of the evaluation (face states,
  // indicator planes, lists, counters, this evaluation's theta, speeds) carved
  // from the library's file-scope scratch array instead of the solver's own:
  // no device allocation.  Reads the solver's weights and previous theta and
  // changes none of its state.  Throws if the scratch does not fit.
  double evalRhsNoAlloc(const double *U, double *rhs);
  // scratch bytes evalRhsNoAlloc needs for this solver
  size_t noAllocBytes() const;
  int takeStep();
  // 2D batches: number of slices and their times / step sizes
  int slices() const { return slices_; }
  void sliceTimes(double *t, double *dt) const;
  // time and next step size (of every problem of a batch)
  void setTime(double t, double dtNext);
  int64_t solve();
  // Advance exactly n RK stages of fixed-dt steps, queued on the device with
  // no host synchronisation (a step completes every rkStages() stages).
  void advanceStages(int64_t n, double dtFixed);

  void getState(double *host) const;
  void setState(const double *host);
  void getRHS(double *host) const;
  void getFlags(double *host) const;
  void setFlags(const double *host);
  void setWeights(const double *face, const double *deriv);
  void sync() const;

  // Plot / checkpoint files of the current state (WriterPDI::writePlot /
  // writeCkpt); return the path of the HDF5 file.
  std::string writePlot();
  std::string writeCkpt();
  // Flow statistics of the current state (Solver::evalFlowStats,
  // Solver.cpp:154-186): Euler kinetic energy and enstrophy, GLM-MHD also
  // magnetic helicity, linear advection one zero; returns their number (<= 3).
  int flowStats(double *out);
  int plotNum = 0;  // index of the next plot (Solver.H plotNum)

  const Config &config() const { return cfg_; }
  const WeightTable &weights() const { return wts_; }
  int nv() const { return nv_; }
  int64_t ncell() const { return ncell_; }
  int rkStages() const { return rk_.nStages; }

  double time = 0.0, dt = 0.0, errEst = 1.0;
  int64_t nTS = 1, nRhsEval = 0, nRejectUnphys = 0, nRejectThresh = 0, lastNWeno = 0;
  bool lastTimeStep = false, useSparseWeno = true;

  // Learned troubled-cell indicator (Euler, GLM-MHD).  params: np FP32 network
  // parameters holding bfloat16 values, np == 0 returns to the KXRCF selector.
  // While it is on, every accepted step is checked for numerical
  // admissibility: the cells that fail get WENO in every stage of the next
  // step, and a step in which more than maxFrac of the cells fail is redone
  // with the KXRCF selector.
  void setIndicator(const float *params, int np, double threshold);
  // Admissibility check of learned steps: on / off, tolerance
  // delta = max(valueFrac |bound|, rangeFrac (max - min)), redo fraction.
  void setNad(bool on, double rangeFrac, double valueFrac, double maxFrac) {
    nadOn_ = on;
    nadRange_ = rangeFrac;
    nadValue_ = valueFrac;
    nadMaxFrac_ = maxFrac;
  }
  int64_t nNadCells = 0;  // violating cells summed over checked steps
  int64_t nGraphs = 0;    // batch step attempts captured as HIP graphs
  bool learned() const { return learned_; }
  int64_t nNadFallback = 0;
  // Training samples of the current state: X[ncell][nin] features and Y[ncell]
  // labels (device pointers), labels with threshold tau.
  void collectSamples(float *X, float *Y, double tau);

  // per-kernel event timing
  void profileEnable(bool on);
  void profileReset();
  bool profileGet(const std::string &name, double &ms, int64_t &count);

 private:
  // Per-evaluation scratch buffers of one RHS evaluation (pipeline)
  struct EvalBufs {
    double *FS = nullptr, *WS = nullptr, *SB = nullptr, *WI = nullptr, *props = nullptr, *thNew = nullptr,
           *wflag = nullptr, *SRC = nullptr, *spd = nullptr;
    int *list = nullptr, *counter = nullptr, *ppList = nullptr, *pos = nullptr;
    unsigned long long *chBits = nullptr, *vmax = nullptr;
  };
  // Lay the buffers of this solver's evaluation out from base (256-byte
  // aligned; base null: sizes only) and return the bytes they span.
  size_t carveEval(unsigned char *base, EvalBufs &b) const;
  struct Prof {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double ms = 0.0;
    int64_t count = 0;
  };
  void setup();
  void release() noexcept;
  template <class F>
  void timed(const char *name, F &&launch);
  void harvest();

  void pipeline(const double *Uin, const RKArgs &a);
  void pipelineMhd(const double *Uin, const RKArgs &a);
  void uploadWeights();
  double readSpeed();
  bool nadFails();
  int takeStepSlices();

  Config cfg_;
  RKTable rk_;
  WeightTable wts_;
  Grid grid_{};
  int nv_ = 5;
  int64_t ncell_ = 0;
  hipStream_t stream_ = nullptr;

  // device buffers
  double *U_[2] = {nullptr, nullptr};  // ping-pong stage states
  int cur_ = 0;
  double *Uprev_ = nullptr, *Uhat_ = nullptr, *Utmp_ = nullptr, *K_ = nullptr;
  double *FS_ = nullptr, *WI_ = nullptr, *wflag_ = nullptr;
  double *WS_ = nullptr;  // compact WENO states of the first wsCap_ listed cells (Euler)
  int *pos_ = nullptr;    // per cell: its row in WS_ (fused path: in FS_), or -1
  int wsCap_ = 0;
  // Euler sparse mode, fused path: the brick-boundary face states of the linear
  // pass (surface planes) and the face weights per point [Q][C][F]; listed
  // cells keep their states in rows of FS_
  bool fused_ = false;
  double *SB_ = nullptr, *faceWQ_ = nullptr;
  Surf sb_{};
  double *props_ = nullptr;  // per-cell internal energy and sound speed [2][ncell]
  double *theta_[2] = {nullptr, nullptr};  // positivity theta: previous RHS / this RHS
  int th_ = 0;                             // theta_[th_] holds the latest values
  // WENO list; counter_[0] WENO cells, [1] limiter list, [2] limiter rows (fused path)
  int *list_ = nullptr, *counter_ = nullptr;
  int *ppList_ = nullptr;                      // cells for the limiter's exact sequence
  // GLM-MHD: interior point states / sources, interior weights, differentiation
  // matrix, cleaning-speed maxima
  double *SRC_ = nullptr, *cellW_ = nullptr, *cellWS_ = nullptr, *diffM_ = nullptr;
  unsigned long long *chBits_ = nullptr;
  // GLM-MHD: per-cell cleaning-speed maxima [2][ncell]; linear-pass weights (face, then interior passes)
  double *spd_ = nullptr, *mhdWT_ = nullptr;
  double *statsPart_ = nullptr;  // flow statistics: per-block partial sums [3][blocks]
  int statsRow_ = 0;             // next row of <basename>_stats.h5
  void evalFlowStats();
  double *faceW_ = nullptr, *derivW_ = nullptr, *faceWT_ = nullptr, *faceWS_ = nullptr, *derivWS_ = nullptr;
  double *faceWC_ = nullptr;  // Euler WENO: compact face + derivative weights (compact_weno_weights)
  double *errPart_ = nullptr, *posPart_ = nullptr, *redOut_ = nullptr;
  unsigned long long *vmax_ = nullptr;
  double *hostRed_ = nullptr;  // pinned: [err, pos, vmax bits, counter]
  int nPart_ = 0;
  bool lastRhsFull_ = false;  // the last RHS evaluation ran WENO on every cell
  bool stateInCur_ = false;  // last step was not accepted: the solution is in U_[cur_]
  int stagePos_ = 0;         // position inside a step for advanceStages

  // learned indicator: network parameters, logit threshold, NAD counter, and
  // the linear face states kept while labels are collected
  float *netP_ = nullptr;
  float netThr_ = 0.0f;
  bool learned_ = false, forcePlain_ = false, nadOn_ = true;
  double nadRange_ = 0.05, nadValue_ = 1e-3, nadMaxFrac_ = 0.02;
  double *nadFlag_ = nullptr;  // cells that failed the last check
  int *nadCount_ = nullptr;
  double *FSlin_ = nullptr;

  // 2D batch of slices: per-slice controller state (host) and step sizes,
  // reductions and commit mask (device, with pinned host mirrors)
  int slices_ = 1;
  std::vector<double> sT_, sDt_, sErr_;
  std::vector<char> sLast_, sDone_;
  double *sDtDev_ = nullptr, *sRed_ = nullptr, *sHost_ = nullptr;  // host: [dt S][red 2S][vmax S]
  int *sMask_ = nullptr, *sMaskHost_ = nullptr;
  // theta and WENO flags of each slice's accepted evaluation in this step
  double *thBak_ = nullptr, *wfBak_ = nullptr;
  // step attempts of a batch captured as HIP graphs, by key (buffer parities,
  // re-prime, WENO mode), with the host state each replay leaves
  struct AttemptGraph {
    hipGraphExec_t exec = nullptr;
    int cur = 0, th = 0;
    bool sparse = true, full = false;
    int64_t dRhs = 0;
  };
  std::map<int, AttemptGraph> graphs_;
  bool graphsOn_ = false;
  template <class F>
  void runAttempt(bool prime, F &&body);

  std::unique_ptr<Writer> writer_;
  bool writePlot_ = false;  // the last step ended on a plot time

  bool profOn_ = false;
  std::map<std::string, Prof> prof_;
};

}  // namespace kfvm
