// x^y for 0 <= x <= 1 and y >= 0, the only range the HLLC shear factor
// (p_min/p_max)^M needs.  exp(y ln x) with
//   ln x : x = m 2^e, m in [1/sqrt 2, sqrt 2), ln m = 2 atanh(s) with
//          s = (m-1)/(m+1), |s| <= 0.1716, by its Taylor series to s^23
//          (truncation < 1e-18 relative);
//   exp t: Cody-Waite reduction t = n ln2 + r, |r| <= ln2/2, ln2 split in a
//          20-bit head (n*head exact) and a tail, exp r by Taylor to r^13
//          (truncation < 4e-18 relative), times 2^n.
// Error (against long-double powl, tests/test_powu.py): below 3.6e-15
// relative where x^y > 1e-3 (the rounding of t = y ln x carries over), and
// 1 - x^y, the quantity HLLC uses, within one ulp of 1 everywhere.
#pragma once

#ifdef __HIPCC__
#define KFVM_POWU_HD __host__ __device__
#else
#define KFVM_POWU_HD
#endif

#include <cmath>

namespace kfvm {

KFVM_POWU_HD inline double powu(double x, double y) {
  if (!(x > 0.0)) return x == 0.0 ? (y > 0.0 ? 0.0 : 1.0) : x;  // 0^y; NaN passes through
  constexpr double LN2_HEAD = 0.6931467056274414;      // ln 2 with 32 trailing zero bits
  constexpr double LN2_TAIL = 4.7493250390316726e-07;  // ln 2 - LN2_HEAD
  constexpr double INV_LN2 = 1.4426950408889634;
  int e;
  double m = std::frexp(x, &e);
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double p = 1.0 / 23.0;
  p = std::fma(p, s2, 1.0 / 21.0);
  p = std::fma(p, s2, 1.0 / 19.0);
  p = std::fma(p, s2, 1.0 / 17.0);
  p = std::fma(p, s2, 1.0 / 15.0);
  p = std::fma(p, s2, 1.0 / 13.0);
  p = std::fma(p, s2, 1.0 / 11.0);
  p = std::fma(p, s2, 1.0 / 9.0);
  p = std::fma(p, s2, 1.0 / 7.0);
  p = std::fma(p, s2, 1.0 / 5.0);
  p = std::fma(p, s2, 1.0 / 3.0);
  const double lnm = std::fma(2.0 * s * s2, p, 2.0 * s);
  const double de = double(e);
  const double t = y * (de * LN2_HEAD + (de * LN2_TAIL + lnm));
  if (t < -746.0) return 0.0;
  const double n = std::rint(t * INV_LN2);
  const double r = (t - n * LN2_HEAD) - n * LN2_TAIL;
  double q = 1.6059043836821613e-10;  // 1/13!
  q = std::fma(q, r, 2.08767569878681e-09);
  q = std::fma(q, r, 2.505210838544172e-08);
  q = std::fma(q, r, 2.755731922398589e-07);
  q = std::fma(q, r, 2.7557319223985893e-06);
  q = std::fma(q, r, 2.48015873015873e-05);
  q = std::fma(q, r, 0.0001984126984126984);
  q = std::fma(q, r, 0.001388888888888889);
  q = std::fma(q, r, 0.008333333333333333);
  q = std::fma(q, r, 0.041666666666666664);
  q = std::fma(q, r, 0.16666666666666666);
  q = std::fma(q, r, 0.5);
  q = std::fma(q, r, 1.0);
  q = std::fma(q, r, 1.0);
  return std::ldexp(q, int(n));
}

}  // namespace kfvm
