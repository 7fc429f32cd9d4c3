// Plot / checkpoint files in the layout of the reference's PDI writer
// (src/io/WriterPDI.cpp:16-325 with user/pdi_*_xdmf.yml): one HDF5 file per
// output with one [nz][ny][nx] double dataset per variable, the nodal
// coordinates, time, step and gamma, plus the XDMF description next to it,
// under <datadir>/R<r>_NX<nx>_NY<ny>[_NZ<nz>]/{plot,ckpt}/<base>_<step:07d>.
// The HDF5 library is loaded at run time (libhdf5 1.10 or newer); without it
// every output call fails with a message and the solver itself is unaffected.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace kfvm {

struct OutputFields {
  // host arrays over interior cells, [var][nz][ny][nx]
  const double *U = nullptr;      // conserved variables (NV planes)
  const double *flags = nullptr;  // WENO flag and positivity theta (2 planes)
};

struct CkptData {
  std::vector<double> U;      // [NV][nz][ny][nx]
  std::vector<double> theta;  // the "fomx" dataset when present
  int step = 0;
  double time = 0.0, dt = 0.0, errEst = 1.0;
  bool haveErrEst = false;
};

class Writer {
 public:
  explicit Writer(const Config &cfg);
  // directory prefix <datadir>/R<r>_NX.._NY..[_NZ..]/
  const std::string &prefix() const { return prefix_; }
  // plot file: conserved, auxiliary (velocities, eint, pressure(s)), weno, fomx
  std::string writePlot(const OutputFields &f, int num, double time);
  // checkpoint: conserved, weno, fomx, time, step, step size (+ errEst)
  std::string writeCkpt(const OutputFields &f, int step, double time, double dt, double errEst);
  // flow statistics row `row` of <basename>_stats.h5: time, then nstats values
  void writeStats(int row, int maxRows, double time, const double *stats, int nstats);
  // read a checkpoint written by writeCkpt (or by the reference's writer)
  CkptData readCkpt(const std::string &path) const;

  // names of the conserved datasets of the equation (Hydro/MHD_GLM/LinAdv SimVar order)
  static std::vector<std::string> consNames(Eq eq);

 private:
  void writeFile(const std::string &path, const OutputFields &f, bool plot, int step, double time,
                 double dt, double errEst);
  void writeXdmf(const std::string &path, const std::string &h5name, double time, bool plot) const;

  Config cfg_;
  int nv_;
  int64_t nx_, ny_, nz_;
  std::string prefix_;
  std::vector<double> xc_, yc_, zc_;
};

}  // namespace kfvm
