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

/* GLM-MHD conserved state (rho, mx, my, mz, E, Bx, By, Bz, psi = 0). */
inline void mhd_prim2cons(double gamma, double rho, double u, double v, double w, double p,
                          double bx, double by, double bz, double *s) {
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

/* Returns false for an unknown problem.  cell = global interior cell index
 * (used by problems with per-cell random data); mid = centre of the box. */
inline bool point_ic(const char *name, int dim, double gamma, const double *user, uint64_t seed,
                     uint64_t cell, const double *mid, double x, double y, double z, double *s) {
  const double tp = 2.0 * M_PI;
  if (!std::strcmp(name, "mhd_vortex3d")) {
    /* rho = gamma^2, p = gamma, v = (-sin 2pi y, sin 2pi x, 0.1 sin 2pi z),
       B = (-sin 2pi y, sin 4pi x, 0) / sqrt(4 pi) */
    const double b0 = 1.0 / std::sqrt(4.0 * M_PI);
    mhd_prim2cons(gamma, gamma * gamma, -std::sin(tp * y), std::sin(tp * x), 0.1 * std::sin(tp * z), gamma,
                  -std::sin(tp * y) * b0, std::sin(2.0 * tp * x) * b0, 0.0, s);
    return true;
  }
  if (!std::strcmp(name, "orszag_tang")) {
    const double e = 0.2;
    mhd_prim2cons(gamma, gamma * gamma, -(1.0 + e * std::sin(tp * z)) * std::sin(tp * y),
                  (1.0 + e * std::sin(tp * z)) * std::sin(tp * x), e * std::sin(tp * z), gamma,
                  -std::sin(tp * y), std::sin(2.0 * tp * x), 0.0, s);
    return true;
  }
  if (!std::strcmp(name, "circ_alfven")) {
    /* wave along x1 = x cos(a) cos(b) + y cos(a) sin(b) + z sin(a),
       sin a = 2/3, sin b = 2/sqrt 5 */
    const double sa = 2.0 / 3.0, sb = 2.0 / std::sqrt(5.0), ca = std::sqrt(5.0) / 3.0, cb = 1.0 / std::sqrt(5.0);
    const double x1 = x * ca * cb + y * ca * sb + z * sa;
    const double par = 0.0, p2 = 0.1 * std::sin(tp * x1), p3 = 0.1 * std::cos(tp * x1);
    /* rotate (parallel, perp2, perp3) back to (x, y, z) for v and for B = (1, p2, p3) */
    const double u = par * ca * cb - p2 * sb - p3 * sa * cb;
    const double v = par * ca * sb + p2 * cb - p3 * sa * sb;
    const double w = par * sa + p3 * ca;
    const double bx = 1.0 * ca * cb - p2 * sb - p3 * sa * cb;
    const double by = 1.0 * ca * sb + p2 * cb - p3 * sa * sb;
    const double bz = 1.0 * sa + p3 * ca;
    mhd_prim2cons(gamma, 1.0, u, v, w, 0.1, bx, by, bz, s);
    return true;
  }
  if (!std::strcmp(name, "mhd_uniform")) {
    mhd_prim2cons(gamma, user[0], user[1], user[2], user[3], user[4], user[5], user[6], user[7], s);
    return true;
  }
  if (!std::strcmp(name, "mhd_blast")) {
    const double zc = dim == 3 ? mid[2] : z;
    const double r2 = (x - mid[0]) * (x - mid[0]) + (y - mid[1]) * (y - mid[1]) + (z - zc) * (z - zc);
    mhd_prim2cons(gamma, 1.0, 0.0, 0.0, 0.0, r2 < user[0] * user[0] ? 10.0 : 0.1, user[1], user[2], 0.0, s);
    return true;
  }
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
  if (!std::strcmp(name, "riem_3d_1c")) {
    /* exec/Hydro/Riem_3d_1c/InitialCondition.H: octants split at the origin */
    double rho, u = 0.0, v = 0.0, w = 0.0, p = 1.0;
    const double a = 0.727606875;
    const bool ex = x >= 0.0, ny_ = y >= 0.0, tz = z >= 0.0;
    if (!ex && !ny_ && !tz) { rho = 1.016216216; u = v = w = -0.401442839; p = 1.4; }
    else if (!ex && !ny_) { rho = 0.8; w = -a; }
    else if (!ex && !tz) { rho = 0.8; v = -a; }
    else if (!ex) { rho = 1.0; u = a; }
    else if (!ny_ && !tz) { rho = 0.8; u = -a; }
    else if (!ny_) { rho = 1.0; v = a; }
    else if (!tz) { rho = 1.0; w = a; }
    else { rho = 0.53125; p = 0.4; }
    hydro_prim2cons(gamma, rho, u, v, w, p, s);
    return true;
  }
  if (!std::strcmp(name, "riem_2d_3c")) {
    /* exec/Hydro/Riem_2d_3c/InitialCondition.H, quadrants split at 0.8 */
    const double gm = gamma - 1.0, a = 4.0 / std::sqrt(11.0);
    double r, mx = 0.0, my = 0.0, e;
    if (y > 0.8) {
      if (x > 0.8) { r = 1.5; e = 1.5 / gm; }
      else { r = 33.0 / 62.0; mx = r * a; e = 0.3 / gm + r * (8.0 / 11.0); }
    } else {
      if (x > 0.8) { r = 33.0 / 62.0; my = r * a; e = 0.3 / gm + r * (8.0 / 11.0); }
      else { r = 77.0 / 558.0; mx = r * a; my = r * a; e = (9.0 / 310.0) / gm + r * (16.0 / 11.0); }
    }
    s[0] = r; s[1] = mx; s[2] = my; s[3] = 0.0; s[4] = e;
    return true;
  }
  if (!std::strcmp(name, "disprel")) {
    /* exec/LinAdv/DispRel/InitialCondition.H with ang = user[0], freq = user[1] */
    s[0] = std::sin(2.0 * M_PI * user[1] * (x * std::cos(user[0]) + y * std::sin(user[0])));
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
