// Compiled-in problem registry: point initial conditions and their Gauss-rule
// cell averages.  Replaces the reference's per-problem InitialCondition.H
// (textually included at src/numeric/Numeric_K.H:25-27) since device callbacks
// cannot cross a C ABI; the averaging follows IntegrateIC_K
// (src/numeric/Numeric_K.H:31-107).  The IC is computed on the host with the
// C library's math so that it does not depend on device math routines.
#include "problems.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <vector>

namespace kfvm {
namespace {

uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// counter-based uniform [0,1)
double uniform01(uint64_t seed, uint64_t ctr) {
  const uint64_t z = mix64(seed ^ mix64(ctr + 0x632BE59BD9B4E019ULL));
  return double(z >> 11) * (1.0 / 9007199254740992.0);
}

void prim2cons(double gamma, double rho, double u, double v, double w, double p, double *s) {
  const double gamm = gamma - 1.0;
  s[0] = rho;
  s[1] = rho * u;
  s[2] = rho * v;
  s[3] = rho * w;
  s[4] = p / gamm + rho * (u * u + v * v + w * w) / 2.0;
}

using PointIC = void (*)(const Config &, uint64_t cell, double x, double y, double z, double *s);

void ic_linadv_sine(const Config &c, uint64_t, double x, double y, double z, double *s) {
  s[0] = std::sin(2.0 * M_PI * (c.user[0] * x + c.user[1] * y + c.user[2] * z));
}
void ic_riemann_planar(const Config &c, uint64_t, double x, double y, double z, double *s) {
  const double r = c.dim == 3 ? x + y + z : x + y;
  if (r < c.user[0]) prim2cons(c.gamma, 1.0, 0.0, 0.0, 0.0, 1.0, s);
  else prim2cons(c.gamma, 0.125, 0.0, 0.0, 0.0, 0.1, s);
}
void ic_sod_x(const Config &c, uint64_t, double x, double, double, double *s) {
  if (x < c.user[0]) prim2cons(c.gamma, 1.0, 0.0, 0.0, 0.0, 1.0, s);
  else prim2cons(c.gamma, 0.125, 0.0, 0.0, 0.0, 0.1, s);
}
void ic_shear2d(const Config &c, uint64_t cell, double, double y, double, double *s) {
  const bool band = std::fabs(y - 0.5) < 0.25;
  const double v = 1.0e-2 * (2.0 * uniform01(c.seed, cell) - 1.0);
  prim2cons(c.gamma, band ? 2.0 : 1.0, band ? 0.5 : -0.5, v, 0.0, 2.5, s);
}
void ic_quadrant2d(const Config &c, uint64_t, double x, double y, double, double *s) {
  const int q = (x < 0.5 ? 0 : 1) + (y < 0.5 ? 0 : 2);
  const uint64_t b = 4 * uint64_t(q);
  const double rho = 0.1 + 1.9 * uniform01(c.seed, b + 0);
  const double p = 0.1 + 1.9 * uniform01(c.seed, b + 1);
  const double u = -1.0 + 2.0 * uniform01(c.seed, b + 2);
  const double v = -1.0 + 2.0 * uniform01(c.seed, b + 3);
  prim2cons(c.gamma, rho, u, v, 0.0, p, s);
}
void ic_uniform(const Config &c, uint64_t, double, double, double, double *s) {
  prim2cons(c.gamma, c.user[0], c.user[1], c.user[2], c.user[3], c.user[4], s);
}
void ic_smooth_wave(const Config &c, uint64_t, double x, double y, double z, double *s) {
  const double ph = 2.0 * M_PI * (x + (c.dim >= 2 ? y : 0.0) + (c.dim == 3 ? z : 0.0));
  prim2cons(c.gamma, 1.0 + 0.2 * std::sin(ph), c.user[0], c.user[1], c.user[2],
            1.0 + 0.1 * std::cos(ph), s);
}

// GLM-MHD states (rho, mx, my, mz, E, Bx, By, Bz, psi)
void mhd_state(double gamma, double rho, double u, double v, double w, double p, double bx,
               double by, double bz, double *s) {
  s[0] = rho;
  s[1] = rho * u;
  s[2] = rho * v;
  s[3] = rho * w;
  s[4] = p / (gamma - 1.0) + rho * (u * u + v * v + w * w) / 2.0 + (bx * bx + by * by + bz * bz) / 2.0;
  s[5] = bx;
  s[6] = by;
  s[7] = bz;
  s[8] = 0.0;
}
// 3D vortex: rho = gamma^2, p = gamma, v = (-sin 2pi y, sin 2pi x, 0.1 sin 2pi z),
// B = (-sin 2pi y, sin 4pi x, 0) / sqrt(4 pi)
void ic_mhd_vortex3d(const Config &c, uint64_t, double x, double y, double z, double *s) {
  const double g = c.gamma, b0 = 1.0 / std::sqrt(4.0 * M_PI);
  mhd_state(g, g * g, -std::sin(2.0 * M_PI * y), std::sin(2.0 * M_PI * x), 0.1 * std::sin(2.0 * M_PI * z), g,
            -std::sin(2.0 * M_PI * y) * b0, std::sin(4.0 * M_PI * x) * b0, 0.0, s);
}
// Orszag-Tang vortex (exec/MHD/OrszagTang/InitialCondition.H)
void ic_orszag_tang(const Config &c, uint64_t, double x, double y, double z, double *s) {
  const double g = c.gamma, eps = 0.2;
  const double u = -(1.0 + eps * std::sin(2.0 * M_PI * z)) * std::sin(2.0 * M_PI * y);
  const double v = (1.0 + eps * std::sin(2.0 * M_PI * z)) * std::sin(2.0 * M_PI * x);
  const double w = eps * std::sin(2.0 * M_PI * z);
  mhd_state(g, g * g, u, v, w, g, -std::sin(2.0 * M_PI * y), std::sin(4.0 * M_PI * x), 0.0, s);
}
// Circularly polarised Alfven wave along a tilted direction
// (exec/MHD/CircAlfven/InitialCondition.H), periodic on [0,3] x [0,1.5]^2
void ic_circ_alfven(const Config &c, uint64_t, double x, double y, double z, double *s) {
  const double sina = 2.0 / 3.0, sinb = 2.0 / std::sqrt(5.0);
  const double cosa = std::sqrt(5.0) / 3.0, cosb = 1.0 / std::sqrt(5.0);
  const double x1 = x * cosa * cosb + y * cosa * sinb + z * sina;
  const double u1 = 0.0, u2 = 0.1 * std::sin(2.0 * M_PI * x1), u3 = 0.1 * std::cos(2.0 * M_PI * x1);
  const double b1 = 1.0, b2 = 0.1 * std::sin(2.0 * M_PI * x1), b3 = 0.1 * std::cos(2.0 * M_PI * x1);
  const double u = u1 * cosa * cosb - u2 * sinb - u3 * sina * cosb;
  const double v = u1 * cosa * sinb + u2 * cosb - u3 * sina * sinb;
  const double w = u1 * sina + u3 * cosa;
  const double bx = b1 * cosa * cosb - b2 * sinb - b3 * sina * cosb;
  const double by = b1 * cosa * sinb + b2 * cosb - b3 * sina * sinb;
  const double bz = b1 * sina + b3 * cosa;
  mhd_state(c.gamma, 1.0, u, v, w, 0.1, bx, by, bz, s);
}
// uniform state: user p0..p7 = rho, u, v, w, p, Bx, By, Bz
void ic_mhd_uniform(const Config &c, uint64_t, double, double, double, double *s) {
  mhd_state(c.gamma, c.user[0], c.user[1], c.user[2], c.user[3], c.user[4], c.user[5], c.user[6],
            c.user[7], s);
}
// blast: pressure 10 inside radius p0 of the box centre, 0.1 outside;
// rho = 1, B = (p1, p2, 0)
void ic_mhd_blast(const Config &c, uint64_t, double x, double y, double z, double *s) {
  const double xc = 0.5 * (c.xLo + c.xUp), yc = 0.5 * (c.yLo + c.yUp);
  const double zc = c.dim == 3 ? 0.5 * (c.zLo + c.zUp) : z;
  const double r2 = (x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc);
  mhd_state(c.gamma, 1.0, 0.0, 0.0, 0.0, r2 < c.user[0] * c.user[0] ? 10.0 : 0.1, c.user[1], c.user[2],
            0.0, s);
}

// Octant Riemann problem of exec/Hydro/Riem_3d_1c/InitialCondition.H: eight
// constant states split at x = y = z = 0 (the state table is the problem's
// data).
void ic_riem_3d_1c(const Config &c, uint64_t, double x, double y, double z, double *s) {
  // {rho, u, v, w, p} per octant, index (x >= 0) * 4 + (y >= 0) * 2 + (z >= 0)
  static const double tab[8][5] = {
      {1.016216216, -0.401442839, -0.401442839, -0.401442839, 1.4},  // bottom south west
      {0.8, 0.0, 0.0, -0.727606875, 1.0},                             // top south west
      {0.8, 0.0, -0.727606875, 0.0, 1.0},                             // bottom north west
      {1.0, 0.727606875, 0.0, 0.0, 1.0},                              // top north west
      {0.8, -0.727606875, 0.0, 0.0, 1.0},                             // bottom south east
      {1.0, 0.0, 0.727606875, 0.0, 1.0},                              // top south east
      {1.0, 0.0, 0.0, 0.727606875, 1.0},                              // bottom north east
      {0.53125, 0.0, 0.0, 0.0, 0.4}};                                 // top north east
  const double *t = tab[(x < 0.0 ? 0 : 4) + (y < 0.0 ? 0 : 2) + (z < 0.0 ? 0 : 1)];
  prim2cons(c.gamma, t[0], t[1], t[2], t[3], t[4], s);
}
// Two-dimensional Riemann problem of exec/Hydro/Riem_2d_3c/InitialCondition.H
// (quadrants split at x = y = 0.8), in its own conserved-variable form.
void ic_riem_2d_3c(const Config &c, uint64_t, double x, double y, double, double *s) {
  const double gamm = c.gamma - 1.0, a = 4.0 / std::sqrt(11.0);
  s[3] = 0.0;
  if (y > 0.8 && x > 0.8) {
    s[0] = 1.5;
    s[1] = s[2] = 0.0;
    s[4] = 1.5 / gamm;
  } else if (y > 0.8) {
    s[0] = 33.0 / 62.0;
    s[1] = s[0] * a;
    s[2] = 0.0;
    s[4] = 0.3 / gamm + s[0] * (8.0 / 11.0);
  } else if (x > 0.8) {
    s[0] = 33.0 / 62.0;
    s[1] = 0.0;
    s[2] = s[0] * a;
    s[4] = 0.3 / gamm + s[0] * (8.0 / 11.0);
  } else {
    s[0] = 77.0 / 558.0;
    s[1] = s[0] * a;
    s[2] = s[0] * a;
    s[4] = (9.0 / 310.0) / gamm + s[0] * (16.0 / 11.0);
  }
}
// Linear-advection dispersion test of exec/LinAdv/DispRel/InitialCondition.H:
// sin(2 pi freq (x cos ang + y sin ang)), ang = p0, freq = p1.
void ic_disprel(const Config &c, uint64_t, double x, double y, double, double *s) {
  const double xpar = x * std::cos(c.user[0]) + y * std::sin(c.user[0]);
  s[0] = std::sin(2.0 * M_PI * c.user[1] * xpar);
}

// Scheme choices a reference problem compiles in (its Definitions.H and the
// bcType array of its ProblemSetup.cpp); explicit problem: keys override them.
struct Preset {
  int dim, nq;
  Eq eq;
  Riemann rs;
  RK rk;
  BC bc[6];
};
// Named user parameter of a reference problem (its UserParameters.H): the
// value goes to user slot `slot`, parsed with std::stof and multiplied by pi
// when `timesPi` (DispRel's ang).
struct UserKey {
  const char *name;
  int slot;
  bool timesPi;
  double dflt;
};
constexpr BC O = BC::Outflow, P = BC::Periodic;
const Preset kRiem3d1c{3, 2, Eq::Hydro, Riemann::HLLC, RK::SSP4_3_2, {O, O, O, O, O, O}};  // Riem_3d_1c
const Preset kRiem2d3c{2, 2, Eq::Hydro, Riemann::HLLC, RK::SSP4_3_2, {O, O, O, O, O, O}};  // Riem_2d_3c
const Preset kDispRel{3, 2, Eq::LinAdv, Riemann::Exact, RK::SSP4_3_2, {P, P, P, P, P, P}};  // DispRel
const UserKey kDispRelKeys[] = {{"ang", 0, true, 0.0}, {"freq", 1, false, 1.0}, {nullptr, 0, false, 0.0}};

struct Entry {
  const char *name;
  Eq eq;
  PointIC fn;
  const Preset *preset = nullptr;
  const UserKey *keys = nullptr;
};
const Entry kRegistry[] = {
    {"linadv_sine", Eq::LinAdv, ic_linadv_sine},
    {"riemann3d_planar", Eq::Hydro, ic_riemann_planar},
    {"sod_x", Eq::Hydro, ic_sod_x},
    {"shear2d", Eq::Hydro, ic_shear2d},
    {"quadrant2d", Eq::Hydro, ic_quadrant2d},
    {"uniform", Eq::Hydro, ic_uniform},
    {"smooth_wave", Eq::Hydro, ic_smooth_wave},
    {"mhd_vortex3d", Eq::MhdGlm, ic_mhd_vortex3d},
    {"orszag_tang", Eq::MhdGlm, ic_orszag_tang},
    {"circ_alfven", Eq::MhdGlm, ic_circ_alfven},
    {"mhd_uniform", Eq::MhdGlm, ic_mhd_uniform},
    {"mhd_blast", Eq::MhdGlm, ic_mhd_blast},
    {"riem_3d_1c", Eq::Hydro, ic_riem_3d_1c, &kRiem3d1c},
    {"riem_2d_3c", Eq::Hydro, ic_riem_2d_3c, &kRiem2d3c},
    {"disprel", Eq::LinAdv, ic_disprel, &kDispRel, kDispRelKeys},
};

const Entry *find_entry(const std::string &name) {
  for (const Entry &e : kRegistry)
    if (name == e.name) return &e;
  return nullptr;
}

}  // namespace

bool problem_known(const std::string &name, Eq eq) {
  const Entry *e = find_entry(name);
  return e && e->eq == eq;
}

std::string problem_names() {
  std::string s;
  for (const Entry &e : kRegistry) s += (s.empty() ? "" : ", ") + std::string(e.name);
  return s;
}

void apply_problem_preset(Config &c) {
  const Entry *e = find_entry(c.problem);
  if (!e) return;
  c.eq = e->eq;
  if (e->preset) {
    c.dim = e->preset->dim;
    c.nQuad = e->preset->nq;
    c.rs = e->preset->rs;
    c.rk = e->preset->rk;
    for (int f = 0; f < 6; f++) c.bc[f] = e->preset->bc[f];
  }
  if (e->keys)
    for (const UserKey *k = e->keys; k->name; k++) c.user[k->slot] = k->dflt;
}

void resolve_user_keys(Config &c, std::string *warnings) {
  const Entry *e = find_entry(c.problem);
  for (const auto &kv : c.userNamed) {
    bool done = false;
    if (e && e->keys)
      for (const UserKey *k = e->keys; k->name && !done; k++)
        if (kv.first == k->name) {
          try {
            c.user[k->slot] = stof_d(kv.second) * (k->timesPi ? M_PI : 1.0);
            done = true;
          } catch (...) {
          }
        }
    if (!done && warnings)
      *warnings += "[Warning] Ignoring line from input file: user:" + kv.first + " " + kv.second + "\n";
  }
}

void fill_initial_condition(const Config &c, std::vector<double> &U) {
  const Entry *entry = find_entry(c.problem);
  if (!entry) throw std::runtime_error("unknown problem: " + c.problem);
  if (entry->eq != c.eq) throw std::runtime_error("problem " + c.problem + " does not match problem:equation");
  const int NV = num_vars(c.eq);
  const int64_t nx = c.nX, ny = c.nY, nz = c.dim == 3 ? c.nZ : 1, ncell = nx * ny * nz;
  U.assign(size_t(ncell) * NV, 0.0);
  const QuadRule qr = quad_rule(c.nQuad);
  const double dx = c.dx(), dy = c.dy(), dz = c.dz();
  const int nQZ = c.dim == 3 ? qr.n : 1;
  auto work = [&](int64_t k0, int64_t k1) {
    double s[9], acc[9];
    for (int64_t k = k0; k < k1; k++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
          const double xc = dx * (double(i) + 0.5) + c.xLo;
          const double yc = dy * (double(j) + 0.5) + c.yLo;
          const double zc = c.dim == 3 ? dz * (double(k) + 0.5) + c.zLo : 0.0;
          const uint64_t cell = (uint64_t(k) * ny + j) * nx + i;
          for (int v = 0; v < NV; v++) acc[v] = 0.0;
          for (int a = 0; a < qr.n; a++)
            for (int b = 0; b < qr.n; b++)
              for (int e = 0; e < nQZ; e++) {
                const double px = qr.ab[a] * dx + xc, py = qr.ab[b] * dy + yc;
                const double pz = c.dim == 3 ? qr.ab[e] * dz + zc : c.zLo;
                for (int v = 0; v < 9; v++) s[v] = 0.0;
                entry->fn(c, cell, px, py, pz, s);
                const double w = c.dim == 3 ? qr.wt[a] * qr.wt[b] * qr.wt[e] : qr.wt[a] * qr.wt[b];
                for (int v = 0; v < NV; v++) acc[v] += w * s[v];
              }
          for (int v = 0; v < NV; v++) U[size_t(v) * ncell + cell] = acc[v];
        }
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nz >= int64_t(nt) && ncell > 100000) {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
      const int64_t k0 = nz * t / nt, k1 = nz * (t + 1) / nt;
      th.emplace_back(work, k0, k1);
    }
    for (auto &t : th) t.join();
  } else {
    work(0, nz);
  }
}

}  // namespace kfvm
