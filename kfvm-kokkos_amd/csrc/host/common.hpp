// Shared host-side definitions of the KFVM-WENO stage library: run-time
// configuration, enumerations and the fixed numerical tables (Gauss-Legendre
// rules, low-storage RK-FSAL coefficients).
//
// Reference correspondence:
//   enums            src/Types.H:43-52, src/physics/EquationTypes.H, RSTypes.H, RKTypes.H
//   QuadRuleLUT      src/numeric/Numeric.cpp:9-45 (note the {0,-c,+c} order of NQ=3)
//   RKCoeff          src/numeric/Numeric.H:117-282
//   ProblemSetup     src/ProblemSetup.H:26-82 (defaults), ParameterStructs.H:36-38
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace kfvm {

enum class Eq : int { Hydro = 0, LinAdv = 1, MhdGlm = 2 };
enum class Riemann : int { LLF = 0, HLL = 1, HLLC = 2, Exact = 3, KEPES = 4 };
enum class RK : int { ThreeStarP = 0, FourStarP = 1, SSP4_3_2 = 2, SSP10_4_3 = 3 };
enum class BC : int { Periodic = 0, Outflow = 1, Reflecting = 2, Noslip = 3 };
enum Face : int { West = 0, East, South, North, Bottom, Top };
enum StepStatus : int {
  ACCEPTED = 0,
  TOLERANCE = 1,
  UNPHYSICAL = 2,
  REJECT_LIMIT = 3,
  STAGNATED = 4
};

inline int num_vars(Eq e) { return e == Eq::Hydro ? 5 : (e == Eq::LinAdv ? 1 : 9); }

// Gauss-Legendre abscissae and weights on (-1/2, 1/2), in the reference's
// point order (centre first for odd rules).
struct QuadRule {
  int n = 0;
  double ab[5] = {0, 0, 0, 0, 0};
  double wt[5] = {0, 0, 0, 0, 0};
};
QuadRule quad_rule(int nq);

// Differentiation matrix of one cell in cell units (DiffMatLUT,
// Numeric.cpp:47-99): row q holds the derivative weights at interior node
// ab[q] of the Lagrange interpolant through the nodes
// (-1/2, ab[0], ..., ab[nq-1], +1/2).  Returned as [nq][nq + 2].
std::vector<double> diff_matrix(int nq);

// Low-storage RK-FSAL coefficient table (Numeric.H:117-282).
struct RKTable {
  int nStages = 0;
  std::vector<double> gam1, gam2, gam3, delta, beta, bhat, c;
  double bhatfsal = 0.0, ep1 = 0.0, ep2 = 0.0;
  bool needsTmp() const;  // true when any gam2/delta entry is non-zero
};
RKTable rk_table(RK rk);

// Everything the init file and the problem registry decide.  Reals that the
// reference parses with std::stof are stored after a round trip through float.
struct Config {
  // mesh
  int dim = 3;
  int64_t nX = 1, nY = 1, nZ = 1;
  double xLo = 0.0, xUp = 1.0, yLo = 0.0, yUp = 1.0, zLo = 0.0, zUp = 1.0;
  // stencil
  int rad = 2;
  int nQuad = 2;
  double lfac = 20.0;
  // physics / scheme selection (compile-time in the reference's Definitions.H)
  Eq eq = Eq::Hydro;
  Riemann rs = Riemann::HLLC;
  RK rk = RK::SSP4_3_2;
  std::array<BC, 6> bc = {BC::Outflow, BC::Outflow, BC::Outflow,
                          BC::Outflow, BC::Outflow, BC::Outflow};
  std::string problem = "riemann3d_planar";
  bool problemSet = false;  // the init file named its problem (problem:name)
  // time
  double finalTime = 1.0, minDt = 1.e-10, cfl = -1.0;  // cfl<0: dimension default
  double atol = 1.e-4, rtol = 1.e-4, rejectionThresh = 0.8;
  int rejectionLimit = 4, maxSteps = 1000;
  double fixedDt = -1.0;  // > 0: every step uses this dt and is accepted unless unphysical
  // eos
  double gamma = 1.4, wenoThresh = 10.0;
  bool haveAdvVel = false;
  double advVel[3] = {0.0, 0.0, 0.0};
  // user parameters of the problem registry
  double user[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t seed = 0;
  // user:<name> <value> lines of a reference problem's UserParameters.H,
  // resolved against the registry once the problem is known
  std::vector<std::pair<std::string, std::string>> userNamed;
  // 2D batch: `slices` independent problems of this configuration advanced
  // together (problem:slices), each with its own seed (seed + s unless given)
  int slices = 1;
  std::vector<uint64_t> sliceSeeds;
  // io
  std::string baseName = "default", dataDir = "data";
  double plotFreq = -1.0;
  int ckptFreq = -1, statsFreq = -1, verbose = 0;
  // tune:fluxbrick: the flux kernel's brick shape (2D: 0 = 16x16, 1 = 32x8, 2 = 64x4; 3D: 0)
  int fluxBrick = 0;
  std::string restartFile;

  double dx() const { return (xUp - xLo) / double(nX); }
  double dy() const { return (yUp - yLo) / double(nY); }
  double dz() const { return dim == 3 ? (zUp - zLo) / double(nZ) : 1.0; }
  double dmin() const;
  double cflValue() const { return cfl > 0.0 ? cfl : (dim == 3 ? 0.5 : 0.75); }
  int64_t nCells() const { return nX * nY * (dim == 3 ? nZ : 1); }
};

// Parse "category:key value" lines (ProblemSetup.H:142-367).  Returns false on
// an unreadable file; unknown keys are reported in *warnings and skipped.
bool parse_init_file(const std::string &path, Config &cfg, std::string *warnings);
bool parse_init_text(const std::string &text, Config &cfg, std::string *warnings);

// stof semantics: parse as float, widen to double.
double stof_d(const std::string &s);

}  // namespace kfvm
