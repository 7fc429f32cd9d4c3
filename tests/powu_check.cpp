// Host build of the device powu (x^y, 0 <= x <= 1, y >= 0) against long-double
// powl: prints the worst relative error (in units of 2^-53) where x^y > 1e-3,
// the worst absolute error of 1 - x^y over the whole range, and edge values.
#include <cmath>
#include <cstdio>
#include <random>

#include "../kfvm-kokkos_amd/csrc/device/powu.hpp"

int main() {
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  double worstRel = 0.0, worstAbs = 0.0;
  for (int i = 0; i < 3000000; i++) {
    const double x = i % 3 == 0 ? U(g) : (i % 3 == 1 ? std::exp(-40.0 * U(g)) : 1.0 - 1e-6 * U(g));
    const double y = i % 2 ? 60.0 * U(g) : 3.0 * U(g);
    const long double ref = powl((long double)x, (long double)y);
    const double a = kfvm::powu(x, y);
    if (ref > 1e-3L) {
      const double rel = double(fabsl((a - ref) / ref)) / 1.1102230246251565e-16;
      worstRel = rel > worstRel ? rel : worstRel;
    }
    const double ab = std::fabs(double((1.0L - (long double)a) - (1.0L - ref)));
    worstAbs = ab > worstAbs ? ab : worstAbs;
  }
  std::printf("%.6g %.6g %.17g %.17g %.17g %.17g %.17g\n", worstRel, worstAbs, kfvm::powu(1.0, 3.5),
              kfvm::powu(0.3, 0.0), kfvm::powu(0.0, 2.0), kfvm::powu(0.25, 0.5),
              kfvm::powu(1e-300, 3.0));
  return 0;
}
