/* Point initial conditions of the oracle's problem registry.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  These restate, for the CPU
 * checker, the initial conditions the benchmark configurations define; the
 * product library carries its own registry.
 */
#ifndef KFVM_ORACLE_PROBLEMS_H
#define KFVM_ORACLE_PROBLEMS_H

#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
/* Uniform [0,1) from a (seed, counter) pair. */
inline double u01(uint64_t seed, uint64_t ctr) {
  const uint64_t z = splitmix64(seed ^ splitmix64(ctr + 0x632BE59BD9B4E019ULL));
  return double(z >> 11) * (1.0 / 9007199254740992.0);
}

inline void hydro_prim2cons(double gamma, double rho, double u, double v, double w, double p,
                            double *s) {
  const double gamm = gamma - 1.0;
  s[0] = rho;
  s[1] = rho * u;
  s[2] = rho * v;
  s[3] = rho * w;
  s[4] = p / gamm + rho * (u * u + v * v + w * w) / 2.0;
}

/* Returns false for an unknown problem.  cell = global interior cell index
 * (used by problems with per-cell random data). */
inline bool point_ic(const char *name, int dim, double gamma, const double *user, uint64_t seed,
                     uint64_t cell, double x, double y, double z, double *s) {
  (void)dim;
  if (!std::strcmp(name, "linadv_sine")) {
    s[0] = std::sin(2.0 * M_PI * (user[0] * x + user[1] * y + user[2] * z));
    return true;
  }
  if (!std::strcmp(name, "riemann3d_planar")) {
    const double r = dim == 3 ? x + y + z : x + y;
    if (r < user[0]) hydro_prim2cons(gamma, 1.0, 0.0, 0.0, 0.0, 1.0, s);
    else hydro_prim2cons(gamma, 0.125, 0.0, 0.0, 0.0, 0.1, s);
    return true;
  }
  if (!std::strcmp(name, "sod_x")) {
    if (x < user[0]) hydro_prim2cons(gamma, 1.0, 0.0, 0.0, 0.0, 1.0, s);
    else hydro_prim2cons(gamma, 0.125, 0.0, 0.0, 0.0, 0.1, s);
    return true;
  }
  if (!std::strcmp(name, "shear2d")) {
    const bool band = std::fabs(y - 0.5) < 0.25;
    const double v = 1.0e-2 * (2.0 * u01(seed, cell) - 1.0);
    hydro_prim2cons(gamma, band ? 2.0 : 1.0, band ? 0.5 : -0.5, v, 0.0, 2.5, s);
    return true;
  }
  if (!std::strcmp(name, "quadrant2d")) {
    const int q = (x < 0.5 ? 0 : 1) + (y < 0.5 ? 0 : 2);
    const uint64_t b = 4 * uint64_t(q);
    const double rho = 0.1 + 1.9 * u01(seed, b + 0);
    const double p = 0.1 + 1.9 * u01(seed, b + 1);
    const double u = -1.0 + 2.0 * u01(seed, b + 2);
    const double v = -1.0 + 2.0 * u01(seed, b + 3);
    hydro_prim2cons(gamma, rho, u, v, 0.0, p, s);
    return true;
  }
  if (!std::strcmp(name, "uniform")) {
    hydro_prim2cons(gamma, user[0], user[1], user[2], user[3], user[4], s);
    return true;
  }
  if (!std::strcmp(name, "smooth_wave")) {
    /* smooth density/pressure wave advected by a uniform velocity */
    const double ph = 2.0 * M_PI * (x + (dim >= 2 ? y : 0.0) + (dim == 3 ? z : 0.0));
    hydro_prim2cons(gamma, 1.0 + 0.2 * std::sin(ph), user[0], user[1], user[2],
                    1.0 + 0.1 * std::cos(ph), s);
    return true;
  }
  return false;
}

}  // namespace orc

#endif
