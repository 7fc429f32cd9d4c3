// Numerical tables and the init-file parser.
//   quad_rule      follows src/numeric/Numeric.cpp:9-45
//   rk_table       follows src/numeric/Numeric.H:117-282
//   parse_init_*   follows src/ProblemSetup.H:142-367 and src/ParameterStructs.cpp:15-38
#include "common.hpp"
#include "problems.hpp"

#include <cmath>
#include <fstream>
#include <sstream>

namespace kfvm {

QuadRule quad_rule(int nq) {
  QuadRule q;
  q.n = nq;
  switch (nq) {
  case 1:
    q.ab[0] = 0.0;
    q.wt[0] = 1.0;
    break;
  case 2:
    q.ab[0] = -0.577350269189625764509148 / 2.0;
    q.ab[1] = 0.577350269189625764509148 / 2.0;
    q.wt[0] = 1.0 / 2.0;
    q.wt[1] = 1.0 / 2.0;
    break;
  case 3:
    // centre first, then -c, +c (Numeric.cpp:21-22)
    q.ab[0] = 0.0;
    q.ab[1] = -0.7745966692414833770358530 / 2.0;
    q.ab[2] = 0.7745966692414833770358530 / 2.0;
    q.wt[0] = 0.8888888888888888888888888 / 2.0;
    q.wt[1] = 0.5555555555555555555555555 / 2.0;
    q.wt[2] = 0.5555555555555555555555555 / 2.0;
    break;
  default:
    q.n = 0;
  }
  return q;
}

std::vector<double> diff_matrix(int nq) {
  const QuadRule qr = quad_rule(nq);
  // nodes in extended precision; the quadrature nodes are the rounded doubles
  std::vector<long double> X;
  X.push_back(-0.5L);
  for (int q = 0; q < nq; q++) X.push_back((long double)qr.ab[q]);
  X.push_back(0.5L);
  const int n = nq + 2;
  std::vector<double> D(size_t(nq) * n, 0.0);
  for (int q = 0; q < nq; q++) {
    const int k = q + 1;
    for (int m = 0; m < n; m++) {
      long double v;
      if (m == k) {
        v = 0.0L;
        for (int j = 0; j < n; j++)
          if (j != k) v += 1.0L / (X[k] - X[j]);
      } else {
        long double num = 1.0L, den = 1.0L;
        for (int j = 0; j < n; j++) {
          if (j != m && j != k) num *= X[k] - X[j];
          if (j != m) den *= X[m] - X[j];
        }
        v = num / den;
      }
      D[size_t(q) * n + m] = double(v);
    }
  }
  return D;
}

bool RKTable::needsTmp() const {
  for (int s = 0; s < nStages; s++)
    if (gam2[s] != 0.0 || delta[s] != 0.0) return true;
  return false;
}

RKTable rk_table(RK rk) {
  RKTable t;
  switch (rk) {
  case RK::SSP4_3_2:
    t.nStages = 4;
    t.gam1 = {0.0, 1.0, 1.0 / 3.0, 1.0};
    t.gam2 = {0.0, 0.0, 0.0, 0.0};
    t.gam3 = {0.0, 0.0, 2.0 / 3.0, 0.0};
    t.delta = {0.0, 0.0, 0.0, 0.0};
    t.beta = {1.0 / 2.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 2.0};
    t.bhat = {1.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0};
    t.bhatfsal = 0.0;
    t.c = {0.0, 1.0 / 2.0, 1.0, 1.0 / 2.0};
    t.ep1 = 0.8 / 3.0;
    t.ep2 = -0.31 / 3.0;
    break;
  case RK::SSP10_4_3:
    t.nStages = 10;
    t.gam1 = {0.0, 1.0, 1.0, 1.0, 2.0 / 5.0, 1.0, 1.0, 1.0, 1.0, 3.0 / 5.0};
    t.gam2 = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
    t.gam3 = {0.0, 0.0, 0.0, 0.0, 3.0 / 5.0, 0.0, 0.0, 0.0, 0.0, -3.0 / 2.0};
    t.delta = {0.0, 0.0, 0.0, 0.0, 0.0, 9.0 / 10.0, 0.0, 0.0, 0.0, 0.0};
    t.beta = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 15.0,
              1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 10.0};
    t.bhat = {0.0, 2.0 / 9.0, 0.0, 0.0, 5.0 / 18.0, 1.0 / 3.0, 0.0, 0.0, 0.0, 1.0 / 6.0};
    t.bhatfsal = 0.0;
    t.c = {0.0,       1.0 / 6.0, 1.0 / 3.0, 1.0 / 2.0, 2.0 / 3.0,
           1.0 / 3.0, 1.0 / 2.0, 2.0 / 3.0, 5.0 / 6.0, 1.0};
    t.ep1 = 0.8 / 3.0;
    t.ep2 = -0.31 / 3.0;
    break;
  case RK::ThreeStarP:
    t.nStages = 5;
    t.gam1 = {0.0, 2.587771979725733308135192812685323706e-01,
              -1.324380360140723382965420909764953437e-01,
              5.056033948190826045833606441415585735e-02,
              5.670532000739313812633197158607642990e-01};
    t.gam2 = {0.0, 5.528354909301389892439698870483746541e-01,
              6.731871608203061824849561782794643600e-01,
              2.803103963297672407841316576323901761e-01,
              5.521525447020610386070346724931300367e-01};
    t.gam3 = {0.0, 0.0, 0.0, 2.752563273304676380891217287572780582e-01,
              -8.950526174674033822276061734289327568e-01};
    t.delta = {1.0, 3.407655879334525365094815965895763636e-01,
               3.414382655003386206551709871126405331e-01,
               7.229275366787987419692007421895451953e-01, 0.0};
    t.beta = {2.300298624518076223899418286314123354e-01,
              3.021434166948288809034402119555380003e-01,
              8.025606185416310937583009085873554681e-01,
              4.362158943603440930655148245148766471e-01,
              1.129272530455059129782111662594436580e-01};
    t.bhat = {9.484166705035703392326247283838082847e-02,
              1.726371339430353766966762629176676070e-01,
              3.998243189084371024483169698618455770e-01,
              1.718016807580178450618829007973835152e-01,
              5.881914422155740300718268359027168467e-02};
    t.bhatfsal = 1.020760551185952388626787099944507877e-01;
    t.c = {0.0, 2.300298624518076223899418286314123354e-01,
           4.050046072094990912268498160116125481e-01,
           8.947822893693433545220710894560512805e-01,
           7.235136928826589010272834603680114769e-01};
    t.ep1 = 0.7 / 3.0;
    t.ep2 = -0.23 / 3.0;
    break;
  case RK::FourStarP:
    t.nStages = 9;
    t.gam1 = {0.0,
              -4.655641447335068552684422206224169103e+00,
              -7.720265099645871829248487209517314217e-01,
              -4.024436690519806086742256154738379161e+00,
              -2.129676284018530966221583708648634733e-02,
              -2.435022509790109546199372365866450709e+00,
              1.985627297131987000579523283542615256e-02,
              -2.810791146791038566946663374735713961e-01,
              1.689434168754859644351230590422137972e-01};
    t.gam2 = {0.0,
              2.499262792574495009336242992898153462e+00,
              5.866820377718875577451517985847920081e-01,
              1.205146086523094569925592464380295241e+00,
              3.474793722186732780030762737753849272e-01,
              1.321346060965113109321230804210670518e+00,
              3.119636464694193615946633676950358444e-01,
              4.351419539684379261368971206040518552e-01,
              2.359698130028753572503744518147537768e-01};
    t.gam3 = {0.0,
              0.0,
              0.0,
              7.621006678721315291614677352949377871e-01,
              -1.981182504339400567765766904309673119e-01,
              -6.228959218699007450469629366684127462e-01,
              -3.752248380775956442989480369774937099e-01,
              -3.355438309135169811915662336248989661e-01,
              -4.560955005031121479972862973705108039e-02};
    t.delta = {1.0,
               1.262923876648114432874834923838556100e+00,
               7.574967189685911558308119415539596711e-01,
               5.163589453140728104667573195005629833e-01,
               -2.746327421802609557034437892013640319e-02,
               -4.382673178127944142238606608356542890e-01,
               1.273587294602656522645691372699677063e+00,
               -6.294740283927400326554066998751383342e-01,
               0.0};
    t.beta = {2.836343005184365275160654678626695428e-01,
              9.736500104654741223716056170419660217e-01,
              3.382359225242515288768487569778320563e-01,
              -3.584943611106183357043212309791897386e-01,
              -4.113944068471528211627210454497620358e-03,
              1.427968894048586363415504654313371031e+00,
              1.808470948394314017665968411915568633e-02,
              1.605770645946802213926893453819236685e-01,
              2.952227015964591648775833803635147962e-01};
    t.bhat = {2.483675912451591196775756814283216443e-02,
              1.866327774562103796990092260942180726e-01,
              5.671080795936984495604436622517631183e-02,
              -3.447695439149287702616943808570747099e-03,
              3.602245056516636472203469198006404016e-03,
              4.545570622145088936800484247980581766e-01,
              -2.434665289427612407531544765622888855e-04,
              6.642755361103549971517945063138312147e-02,
              1.613697079523505006226025497715177578e-01};
    t.bhatfsal = 4.955424859358438183052504342394102722e-02;
    t.c = {0.0,
           2.836343005184365275160654678626695428e-01,
           5.484076570002894365286665352032296535e-01,
           3.687228761669438493478872632332010073e-01,
           -6.806126440140844191258463830024463902e-01,
           3.518526124230705801739919476290327750e-01,
           1.665941994879593315477304663913129942e+00,
           9.715279295934715835299192116436237065e-01,
           9.051569840159589594903399929316959062e-01};
    t.ep1 = 0.38 / 4.0;
    t.ep2 = -0.18 / 4.0;
    break;
  }
  return t;
}

double Config::dmin() const {
  double d = std::fmin(dx(), dy());
  if (dim == 3) d = std::fmin(d, dz());
  return d;
}

double stof_d(const std::string &s) { return static_cast<double>(std::stof(s)); }

namespace {

std::vector<std::string> split_ws(const std::string &s) {
  std::istringstream ls(s);
  std::vector<std::string> out;
  std::string tok;
  while (ls >> tok) out.push_back(tok);
  return out;
}

bool parse_bc(const std::string &s, BC &b) {
  if (s == "periodic") b = BC::Periodic;
  else if (s == "outflow") b = BC::Outflow;
  else if (s == "reflecting") b = BC::Reflecting;
  else if (s == "noslip") b = BC::Noslip;
  else return false;
  return true;
}

// Returns true when the line was understood.
bool apply_line(const std::string &category, const std::string &setting, Config &c) {
  std::vector<std::string> tok = split_ws(setting);
  if (tok.empty()) return false;
  const std::string &key = tok[0];
  auto val = [&](size_t i) -> const std::string & {
    static const std::string empty;
    return i < tok.size() ? tok[i] : empty;
  };
  try {
    if (category == "mesh") {
      if (key == "resolution") {
        c.nX = std::stoi(val(1));
        c.nY = std::stoi(val(2));
        if (tok.size() > 3) c.nZ = std::stoi(val(3));
      } else if (key == "numblocks") {
        // one block on one GPU; the layout is accepted and ignored
        for (size_t i = 1; i < tok.size(); i++)
          if (std::stoi(tok[i]) != 1) return false;
      } else if (key == "extent") {
        c.xLo = stof_d(val(1));
        c.xUp = stof_d(val(2));
        c.yLo = stof_d(val(3));
        c.yUp = stof_d(val(4));
        if (tok.size() > 6) {
          c.zLo = stof_d(val(5));
          c.zUp = stof_d(val(6));
        }
      } else return false;
    } else if (category == "stencil") {
      if (key == "lfac") c.lfac = stof_d(val(1));
      else return false;
    } else if (category == "time") {
      if (key == "mindt") c.minDt = stof_d(val(1));
      else if (key == "final") c.finalTime = stof_d(val(1));
      else if (key == "cfl") c.cfl = stof_d(val(1));
      else if (key == "atol") c.atol = stof_d(val(1));
      else if (key == "rtol") c.rtol = stof_d(val(1));
      else if (key == "rejectionthresh") c.rejectionThresh = stof_d(val(1));
      else if (key == "rejectionlimit") c.rejectionLimit = int(std::stod(val(1)));
      else if (key == "maxsteps") c.maxSteps = int(std::stod(val(1)));
      else if (key == "fixeddt") c.fixedDt = std::stod(val(1));
      else return false;
    } else if (category == "eos") {
      if (key == "gamma") c.gamma = stof_d(val(1));
      else if (key == "wenothresh") c.wenoThresh = stof_d(val(1));
      else if (key == "advvel") {
        c.haveAdvVel = true;
        for (int d = 0; d < 3; d++) c.advVel[d] = tok.size() > size_t(d + 1) ? std::stod(tok[d + 1]) : 0.0;
      } else if (key == "reynolds" || key == "prandtl" || key == "forceedot" ||
                 key == "forcestr" || key == "forcetoff" || key == "forceratio") {
        (void)stof_d(val(1));  // accepted, unused on this path
      } else return false;
    } else if (category == "user") {
      // user:pN value  (N = 0..7) feeds the problem registry
      if (key.size() == 2 && key[0] == 'p' && key[1] >= '0' && key[1] <= '7')
        c.user[key[1] - '0'] = std::stod(val(1));
      else if (key == "seed") c.seed = std::stoull(val(1));
      else if (tok.size() >= 2) c.userNamed.emplace_back(key, val(1));
      else return false;
    } else if (category == "io") {
      if (key == "plotfreq") c.plotFreq = stof_d(val(1));
      else if (key == "ckptfreq") c.ckptFreq = std::stoi(val(1));
      else if (key == "statsfreq") c.statsFreq = std::stoi(val(1));
      else if (key == "basename") c.baseName = val(1);
      else if (key == "datadir") c.dataDir = val(1);
      else if (key == "restartfile") c.restartFile = val(1);
      else if (key == "verbose") c.verbose = std::stoi(val(1));
      else if (key == "pdiconf") {
        // PDI yaml path: accepted, the HDF5 layout is built in
      } else return false;
    } else if (category == "tune") {
      // launch shapes chosen by the tile autotuner (not in the reference)
      if (key == "fluxbrick") c.fluxBrick = std::stoi(val(1));
      else return false;
    } else if (category == "problem") {
      // Replaces the reference's compile-time Definitions.H / ProblemSetup.cpp
      if (key == "name") {
        c.problem = val(1);
        c.problemSet = !c.problem.empty();
      }
      else if (key == "dim") c.dim = std::stoi(val(1));
      else if (key == "nquad") c.nQuad = std::stoi(val(1));
      else if (key == "radius") c.rad = std::stoi(val(1));
      else if (key == "slices") c.slices = std::stoi(val(1));
      else if (key == "equation") {
        if (val(1) == "hydro") c.eq = Eq::Hydro;
        else if (val(1) == "linadv") c.eq = Eq::LinAdv;
        else if (val(1) == "mhd_glm") c.eq = Eq::MhdGlm;
        else return false;
      } else if (key == "riemann") {
        if (val(1) == "hllc") c.rs = Riemann::HLLC;
        else if (val(1) == "hll") c.rs = Riemann::HLL;
        else if (val(1) == "llf") c.rs = Riemann::LLF;
        else if (val(1) == "exact") c.rs = Riemann::Exact;
        else if (val(1) == "kepes") c.rs = Riemann::KEPES;
        else return false;
      } else if (key == "rk") {
        if (val(1) == "ssp4_3_2") c.rk = RK::SSP4_3_2;
        else if (val(1) == "ssp10_4_3") c.rk = RK::SSP10_4_3;
        else if (val(1) == "threestarp") c.rk = RK::ThreeStarP;
        else if (val(1) == "fourstarp") c.rk = RK::FourStarP;
        else return false;
      } else if (key == "bc") {
        if (tok.size() == 2) {
          BC b;
          if (!parse_bc(val(1), b)) return false;
          c.bc.fill(b);
        } else if (tok.size() == 7) {
          for (int f = 0; f < 6; f++)
            if (!parse_bc(tok[f + 1], c.bc[f])) return false;
        } else return false;
      } else return false;
    } else {
      return false;
    }
  } catch (...) {
    return false;
  }
  return true;
}

}  // namespace

bool parse_init_text(const std::string &text, Config &cfg, std::string *warnings) {
  // First pass: the problem's name, whose compiled-in scheme choices (the
  // reference's Definitions.H) become the defaults the explicit keys override.
  {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ls(line);
      std::string category, setting;
      std::getline(ls, category, ':');
      std::getline(ls, setting, ':');
      const std::vector<std::string> tok = split_ws(setting);
      if (category == "problem" && tok.size() >= 2 && tok[0] == "name") cfg.problem = tok[1];
    }
    apply_problem_preset(cfg);
  }
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string category, setting;
    std::getline(ls, category, ':');
    std::getline(ls, setting, ':');
    if (!apply_line(category, setting, cfg)) {
      if (warnings) *warnings += "[Warning] Ignoring line from input file: " + line + "\n";
    }
  }
  resolve_user_keys(cfg, warnings);
  if (cfg.dim == 2) cfg.nZ = 1;
  return true;
}

bool parse_init_file(const std::string &path, Config &cfg, std::string *warnings) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_init_text(ss.str(), cfg, warnings);
}

}  // namespace kfvm
