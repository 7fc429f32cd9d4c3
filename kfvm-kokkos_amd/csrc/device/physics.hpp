// Device physics for the stage kernels: equation of state, linearised
// primitive map, Riemann solvers.  Every expression keeps the operand order of
// the reference so that, with contraction disabled, results agree with a CPU
// evaluation of the same formulas to the last bit (library pow/exp aside; the
// HLLC shear factor uses powu, a few ulp from the library pow).  The HLLC and
// the GLM-MHD HLL / LLF solvers divide and take square roots with div_n /
// sqrt_n, which give the IEEE result for each of 2^26 random operands of both
// signs spread over 2^-40 .. 2^40 (probe/div_sqrt_check.hip).
//
//   Hydro EOS        src/physics/hydro/Hydro_EOS.H:15-93
//   VarConvert       src/physics/hydro/Hydro_K.H:25-65
//   flux / speeds    src/physics/hydro/Hydro_Riemann.H:18-91, Riemann.H:26-45
//   HLLC             src/physics/hydro/Hydro_Riemann.H:97-187
//   HLL / LLF        src/physics/Common_K.H:23-63
//   LinAdv exact     src/physics/linadv/LinAdv_Riemann.H:63-90
//   GLM-MHD          src/physics/mhd_glm/MHD_GLM_EOS.H:15-144,
//                    MHD_GLM_Riemann.H:18-231
#pragma once

#include <hip/hip_runtime.h>

#include "powu.hpp"

namespace kfvm {
namespace dev {

// Reciprocal, quotient and square root of the positive, normal operands of
// the HLLC solver, from the hardware reciprocal and reciprocal-square-root
// estimates: two Newton-Raphson steps on the reciprocal, then one residual
// correction of the quotient (Markstein, IA-64 and Elementary Functions, 2000,
// ch. 8), and for the square root one Goldschmidt step followed by two Newton
// corrections (ibid., ch. 9).  They skip the operand scaling of the general
// IEEE sequences (no subnormal or overflowing operands reach them).
__device__ inline double rcp_n(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ inline double div_n(double a, double b) {
  const double r = rcp_n(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ inline double sqrt_n(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  g = fma(fma(-g, g, x), h, g);
  g = fma(fma(-g, g, x), h, g);
  return x == 0.0 ? x : g;  // rsq(0) = inf
}
// a / b and sqrt(x), the IEEE operations or (N) the sequences above
template <bool N>
__device__ inline double dv(double a, double b) {
  return N ? div_n(a, b) : a / b;
}
template <bool N>
__device__ inline double sq(double x) {
  return N ? sqrt_n(x) : sqrt(x);
}

struct Hydro {
  static constexpr int NV = 5;
  __device__ static inline double ekin(const double *U) {
    return 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
  }
  __device__ static inline double eint(const double *U) { return U[4] - ekin(U); }
  __device__ static inline double pres(double g, const double *U) { return (g - 1.0) * eint(U); }
  __device__ static inline double sound(double g, const double *U) {
    return sqrt(g * pres(g, U) / U[0]);
  }
  __device__ static inline double weno(double g, const double *U) {
    return pres(g, U) / pow(U[0], g);
  }
  __device__ static inline double positivity(const double *U) { return fmin(U[0], eint(U)); }
  __device__ static inline void cons2prim(double g, const double *U, double *V) {
    V[0] = U[0];
    V[1] = U[1] / U[0];
    V[2] = U[2] / U[0];
    V[3] = U[3] / U[0];
    V[4] = pres(g, U);
  }
  template <int DIR>
  __device__ static inline void flux(const double *U, const double *V, double *F) {
    constexpr int v0 = 1 + DIR, v1 = 1 + (DIR + 1) % 3, v2 = 1 + (DIR + 2) % 3;
    F[0] = U[v0];
    F[v0] = U[v0] * V[v0] + V[4];
    F[v1] = U[v0] * V[v1];
    F[v2] = U[v0] * V[v2];
    F[4] = V[v0] * (U[4] + V[4]);
  }
  __device__ static inline double roe(double rL, double rR, double qL, double qR) {
    return (qL * sqrt(rL) + qR * sqrt(rR)) / (sqrt(rL) + sqrt(rR));
  }
  // Batten wave-speed estimates
  template <int DIR>
  __device__ static inline double speeds(double g, const double *uL, const double *uR,
                                         const double *vL, const double *vR, double &sL,
                                         double &sR) {
    constexpr int v0 = 1 + DIR, v1 = 1 + (DIR + 1) % 3, v2 = 1 + (DIR + 2) % 3;
    const double rhoL = uL[0], rhoR = uR[0];
    const double hL = (uL[4] + vL[4]) / rhoL, hR = (uR[4] + vR[4]) / rhoR;
    const double aL = sound(g, uL), aR = sound(g, uR);
    const double uTil = roe(rhoL, rhoR, vL[v0], vR[v0]), vTil = roe(rhoL, rhoR, vL[v1], vR[v1]),
                 wTil = roe(rhoL, rhoR, vL[v2], vR[v2]),
                 vsqTil = uTil * uTil + vTil * vTil + wTil * wTil, hTil = roe(rhoL, rhoR, hL, hR),
                 aTil = sqrt((g - 1.0) * (hTil - 0.5 * vsqTil));
    sL = fmin(vL[v0] - aL, uTil - aTil);
    sR = fmax(vR[v0] + aR, uTil + aTil);
    return fmax(fabs(sL), fabs(sR));
  }
  template <int DIR>
  __device__ static inline double solve_hll(double g, bool llf, const double *uL,
                                            const double *uR, double *F) {
    double vL[5], vR[5], FR[5];
    cons2prim(g, uL, vL);
    cons2prim(g, uR, vR);
    double sL, sR;
    const double sMax = speeds<DIR>(g, uL, uR, vL, vR, sL, sR);
    if (llf) {
      sL = -sMax;
      sR = sMax;
    }
    const double tR = (fmin(0.0, sR) - fmin(0.0, sL)) / (sR - sL), tL = 1.0 - tR;
    const double tD = 0.5 * (sR * fabs(sL) - sL * fabs(sR)) / (sR - sL);
    flux<DIR>(uL, vL, F);
    flux<DIR>(uR, vR, FR);
#pragma unroll
    for (int v = 0; v < 5; v++) F[v] = tL * F[v] + tR * FR[v] - tD * (uR[v] - uL[v]);
    return sMax;
  }
  // HLLC with the low-Mach / shear corrections (Hydro_Riemann.H:97-187).  The
  // formulas are the reference's; divisions by a common denominator are
  // multiplications by its reciprocal (10 divisions instead of 29), and the
  // divisions and square roots are rcp_n / div_n / sqrt_n, so results differ
  // from the reference's by rounding only.
  template <int DIR>
  __device__ static inline double solve_hllc(double g, const double *uL, const double *uR,
                                             double *F) {
    constexpr int v0 = 1 + DIR, v1 = 1 + (DIR + 1) % 3, v2 = 1 + (DIR + 2) % 3;
    const double gm1 = g - 1.0;
    const double rhoL = uL[0], rhoR = uR[0];
    const double irL = rcp_n(rhoL), irR = rcp_n(rhoR);
    const double unL = uL[v0] * irL, vtL = uL[v1] * irL, wtL = uL[v2] * irL;
    const double unR = uR[v0] * irR, vtR = uR[v1] * irR, wtR = uR[v2] * irR;
    const double vsqL = unL * unL + vtL * vtL + wtL * wtL, vsqR = unR * unR + vtR * vtR + wtR * wtR;
    const double pL = gm1 * (uL[4] - 0.5 * rhoL * vsqL), pR = gm1 * (uR[4] - 0.5 * rhoR * vsqR);
    const double aL = sqrt_n(g * pL * irL), aR = sqrt_n(g * pR * irR);
    const double hL = (uL[4] + pL) * irL, hR = (uR[4] + pR) * irR;
    // Batten speeds from Roe averages
    const double srL = sqrt_n(rhoL), srR = sqrt_n(rhoR), iS = rcp_n(srL + srR);
    const double uTil = (unL * srL + unR * srR) * iS, vTil = (vtL * srL + vtR * srR) * iS;
    const double wTil = (wtL * srL + wtR * srR) * iS, hTil = (hL * srL + hR * srR) * iS;
    const double vsqTil = uTil * uTil + vTil * vTil + wTil * wTil;
    const double aTil = sqrt_n(gm1 * (hTil - 0.5 * vsqTil));
    const double sL = fmin(unL - aL, uTil - aTil), sR = fmax(unR + aR, uTil + aTil);
    const double sMax = fmax(fabs(sL), fabs(sR));
    double M = fmax(div_n(sqrt_n(vsqL), aL), div_n(sqrt_n(vsqR), aR));
    const double psL0 = rhoL * (sL - unL), psR0 = rhoR * (sR - unR);
    const double sS = div_n((pR - pL) + (psL0 * unL - psR0 * unR), psL0 - psR0);
    const double pS = 0.5 * ((pL + pR) + psL0 * (sS - unL) + psR0 * (sS - unR));
    const double iL = rcp_n(sL - sS), iR = rcp_n(sR - sS);
    const double sFacL = sS * iL, sFacR = sS * iR;
    const double uJump = unR - unL, vJump = vtR - vtL, wJump = wtR - wtL;
    const double psiFac = div_n(psR0 * psL0, psR0 - psL0);
    const double presRat = div_n(pL < pR ? pL : pR, pL < pR ? pR : pL);
    const double gFac = 1.0 - powu(presRat, M);  // presRat in (0, 1], M >= 0
    const double shFacL = psiFac * gFac * sL * iL, shFacR = psiFac * gFac * sR * iR;
    M = fmin(M, 1.0);
    const double oneMsq = 1.0 - M * M;
    const double fM = div_n(M * sqrt_n(4.0 + oneMsq * oneMsq), 1.0 + M * M);
    const double pFacL = psiFac * (fM - 1.0) * uJump + sL * pS * iL;
    const double pFacR = psiFac * (fM - 1.0) * uJump + sR * pS * iR;
    // regions of the reference: 0 <= sL: F(uL); else 0 <= sS: star L;
    // else 0 < sR: star R; else F(uR)
    const bool left = 0.0 <= sL || 0.0 <= sS;
    const double *u = left ? uL : uR;
    const double un = left ? unL : unR, vt = left ? vtL : vtR, wt = left ? wtL : wtR;
    const double p = left ? pL : pR;
    F[0] = u[v0];
    F[v0] = u[v0] * un + p;
    F[v1] = u[v0] * vt;
    F[v2] = u[v0] * wt;
    F[4] = un * (u[4] + p);
    if (left ? !(0.0 <= sL) : 0.0 < sR) {
      const double sK = left ? sL : sR, sFac = left ? sFacL : sFacR;
      const double shFac = left ? shFacL : shFacR, pFac = left ? pFacL : pFacR;
      F[0] = sFac * (sK * u[0] - F[0]);
      F[v0] = sFac * (sK * u[v0] - F[v0]) + pFac;
      F[v1] = sFac * (sK * u[v1] - F[v1]) + shFac * vJump;
      F[v2] = sFac * (sK * u[v2] - F[v2]) + shFac * wJump;
      F[4] = sFac * (sK * u[4] - F[4]) + sS * pFac;
    }
    return sMax;
  }
};

// GLM-MHD, variables (rho, mx, my, mz, E, Bx, By, Bz, psi)
// (MHD_GLM_EOS.H:15-144, MHD_GLM_Riemann.H:18-231).  DIR 0/1/2 rotates the
// normal direction as ord = {DIR, DIR+1, DIR+2} (mod 3).
struct Mhd {
  static constexpr int NV = 9;
  // N: divisions and square roots by div_n / sqrt_n (the Riemann solvers)
  template <bool N = false>
  __device__ static inline double ekin(const double *U) {
    return dv<N>(0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]), U[0]);
  }
  __device__ static inline double prsb(const double *U) {
    return 0.5 * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]);
  }
  template <bool N = false>
  __device__ static inline double eint(const double *U) {
    return (U[4] - prsb(U) - ekin<N>(U) - 0.5 * U[8] * U[8]);
  }
  template <bool N = false>
  __device__ static inline double pres(double g, const double *U) {
    return (g - 1.0) * eint<N>(U);
  }
  __device__ static inline double sound(double g, const double *U) {
    return sqrt(g * pres(g, U) / U[0]);
  }
  __device__ static inline double weno(double g, const double *U) {
    return pres(g, U) / pow(U[0], g);
  }
  template <bool N = false>
  __device__ static inline double fast(double g, const double *U, double bl) {
    const double asq = dv<N>(g * pres<N>(g, U), U[0]);
    const double bsq = dv<N>(2.0 * prsb(U), U[0]);
    const double blsq = dv<N>(bl * bl, U[0]);
    const double absq = asq + bsq;
    return sq<N>(0.5 * (absq + sq<N>(absq * absq - 4.0 * asq * blsq)));
  }
  template <bool N = false>
  __device__ static inline void cons2prim(double g, const double *U, double *V) {
    V[0] = U[0];
    V[1] = dv<N>(U[1], U[0]);
    V[2] = dv<N>(U[2], U[0]);
    V[3] = dv<N>(U[3], U[0]);
    V[4] = pres<N>(g, U);
    V[5] = U[5];
    V[6] = U[6];
    V[7] = U[7];
    V[8] = U[8];
  }
  template <int DIR>
  __device__ static inline void flux(double ch, const double *U, const double *V, double *F) {
    constexpr int v0 = 1 + DIR, v1 = 1 + (DIR + 1) % 3, v2 = 1 + (DIR + 2) % 3;
    constexpr int b0 = 5 + DIR, b1 = 5 + (DIR + 1) % 3, b2 = 5 + (DIR + 2) % 3;
    const double mx = U[v0];
    const double vx = V[v0], vy = V[v1], vz = V[v2];
    const double bx = U[b0], by = U[b1], bz = U[b2];
    const double et = U[4], pg = V[4];
    const double chpsi = ch * U[8], hpsisq = 0.5 * U[8] * U[8];
    const double uDb = vx * bx + vy * by + vz * bz, bsq = bx * bx + by * by + bz * bz;
    const double pt = pg + 0.5 * bsq;
    F[0] = mx;
    F[v0] = mx * vx + pt - bx * bx;
    F[v1] = mx * vy - bx * by;
    F[v2] = mx * vz - bx * bz;
    F[4] = vx * (et + pt - hpsisq) + bx * (chpsi - uDb);
    F[b0] = chpsi;
    F[b1] = vx * by - vy * bx;
    F[b2] = vx * bz - vz * bx;
    F[8] = ch * bx;
  }
  // Davis estimates with c_h, extremised with the Wu-Shu positivity estimates
  template <int DIR>
  __device__ static inline double speeds(double g, double ch, const double *uL, const double *uR,
                                         const double *vL, const double *vR, double &sL,
                                         double &sR) {
    constexpr int v0 = 1 + DIR, b0 = 5 + DIR, b1 = 5 + (DIR + 1) % 3, b2 = 5 + (DIR + 2) % 3;
    const double cfL = fast<true>(g, uL, uL[b0]);
    const double cfR = fast<true>(g, uR, uR[b0]);
    const double sdL = fmin(vL[v0] - fmax(ch, cfL), vR[v0] - fmax(ch, cfR));
    const double sdR = fmax(vL[v0] + fmax(ch, cfL), vR[v0] + fmax(ch, cfR));
    const double rtL = sqrt_n(uL[0]), rtR = sqrt_n(uR[0]);
    const double rtSum = rtL + rtR, vSum = div_n(rtL * vL[v0] + rtR * vR[v0], rtSum);
    const double bd0 = uL[b0] - uR[b0], bd1 = uL[b1] - uR[b1], bd2 = uL[b2] - uR[b2];
    const double bnDiff = div_n(sqrt_n(bd0 * bd0 + bd1 * bd1 + bd2 * bd2), rtSum);
    const double aL = fmin(vL[v0], vSum) - cfL - bnDiff;
    const double aR = fmin(vR[v0], vSum) + cfR + bnDiff;
    sL = fmin(sdL, aL);
    sR = fmax(sdR, aR);
    return fmax(fabs(sL), fabs(sR));
  }
  // generic HLL / LLF (Common_K.H:23-63)
  template <int DIR>
  __device__ static inline double solve_hll(double g, double ch, bool llf, const double *uL,
                                            const double *uR, double *F) {
    double vL[9], vR[9], FR[9];
    cons2prim<true>(g, uL, vL);
    cons2prim<true>(g, uR, vR);
    double sL, sR;
    const double sMax = speeds<DIR>(g, ch, uL, uR, vL, vR, sL, sR);
    if (llf) {
      sL = -sMax;
      sR = sMax;
    }
    const double tR = div_n(fmin(0.0, sR) - fmin(0.0, sL), sR - sL), tL = 1.0 - tR;
    const double tD = div_n(0.5 * (sR * fabs(sL) - sL * fabs(sR)), sR - sL);
    flux<DIR>(ch, uL, vL, F);
    flux<DIR>(ch, uR, vR, FR);
#pragma unroll
    for (int v = 0; v < 9; v++) F[v] = tL * F[v] + tR * FR[v] - tD * (uR[v] - uL[v]);
    return sMax;
  }
  __device__ static inline double logMean(double a, double b) {
    const double zeta = a / b, f = (zeta - 1.0) / (zeta + 1.0), u = f * f;
    const double F = u < 0.01 ? 1.0 + u / 3.0 + u * u / 5.0 + u * u * u / 7.0 : 0.5 * log(zeta) / f;
    return 0.5 * (a + b) / F;
  }
  __device__ static inline double am(double a, double b) { return 0.5 * (a + b); }
  // entropy-conserving flux plus LLF dissipation (MHD_GLM_Riemann.H:118-231)
  template <int DIR>
  __device__ static inline double solve_kepes(double g, double ch, const double *uL,
                                              const double *uR, double *F) {
    constexpr int v0 = 1 + DIR, v1 = 1 + (DIR + 1) % 3, v2 = 1 + (DIR + 2) % 3;
    constexpr int b0 = 5 + DIR, b1 = 5 + (DIR + 1) % 3, b2 = 5 + (DIR + 2) % 3;
    double vL[9], vR[9];
    cons2prim(g, uL, vL);
    cons2prim(g, uR, vR);
    double sL, sR;
    const double sMax = speeds<DIR>(g, ch, uL, uR, vL, vR, sL, sR);
    const double rhoLn = logMean(uL[0], uR[0]);
    const double rhoAv = am(uL[0], uR[0]);
    const double uAv = am(vL[v0], vR[v0]), vAv = am(vL[v1], vR[v1]), wAv = am(vL[v2], vR[v2]);
    const double uuAv = am(vL[v0] * vL[v0], vR[v0] * vR[v0]);
    const double vvAv = am(vL[v1] * vL[v1], vR[v1] * vR[v1]);
    const double wwAv = am(vL[v2] * vL[v2], vR[v2] * vR[v2]);
    const double bxAv = am(uL[b0], uR[b0]), byAv = am(uL[b1], uR[b1]), bzAv = am(uL[b2], uR[b2]);
    const double bxxAv = am(uL[b0] * uL[b0], uR[b0] * uR[b0]);
    const double byyAv = am(uL[b1] * uL[b1], uR[b1] * uR[b1]);
    const double bzzAv = am(uL[b2] * uL[b2], uR[b2] * uR[b2]);
    const double uBxAv = am(vL[v0] * uL[b0], vR[v0] * uR[b0]);
    const double vByAv = am(vL[v1] * uL[b1], vR[v1] * uR[b1]);
    const double wBzAv = am(vL[v2] * uL[b2], vR[v2] * uR[b2]);
    const double uBxxAv = am(vL[v0] * uL[b0] * uL[b0], vR[v0] * uR[b0] * uR[b0]);
    const double uByyAv = am(vL[v0] * uL[b1] * uL[b1], vR[v0] * uR[b1] * uR[b1]);
    const double uBzzAv = am(vL[v0] * uL[b2] * uL[b2], vR[v0] * uR[b2] * uR[b2]);
    const double psiAv = am(uL[8], uR[8]);
    const double bxPsiAv = am(uL[b0] * uL[8], uR[b0] * uR[8]);
    const double betaL = uL[0] / (2.0 * vL[4]);
    const double betaR = uR[0] / (2.0 * vR[4]);
    const double betaLn = logMean(betaL, betaR);
    const double betaAv = am(betaL, betaR);
    const double pTil = 0.5 * rhoAv / betaAv;
    const double pHatTot = pTil + 0.5 * (bxxAv + byyAv + bzzAv);
    F[0] = rhoLn * uAv;
    F[v0] = rhoLn * uAv * uAv + pHatTot - bxAv * bxAv;
    F[v1] = rhoLn * uAv * vAv - bxAv * byAv;
    F[v2] = rhoLn * uAv * wAv - bxAv * bzAv;
    F[b0] = ch * psiAv;
    F[b1] = uAv * byAv - vAv * bxAv;
    F[b2] = uAv * bzAv - wAv * bxAv;
    F[8] = ch * bxAv;
    F[4] = (1.0 / (2.0 * (g - 1.0) * betaLn) - 0.5 * (uuAv + vvAv + wwAv)) * F[0] + uAv * F[v0] +
           vAv * F[v1] + wAv * F[v2] + bxAv * F[b0] + byAv * F[b1] + bzAv * F[b2] + psiAv * F[8] -
           0.5 * (uBxxAv + uByyAv + uBzzAv) + bxAv * (uBxAv + vByAv + wBzAv) - ch * bxPsiAv;
#pragma unroll
    for (int v = 0; v < 9; v++) F[v] -= sMax * (uR[v] - uL[v]) / 2.0;
    return sMax;
  }
};

// Linearised primitive variables about a reference (centre-cell) state.
struct LinPrim {
  double gamm, rho, u, v, w, vsq;
  __device__ inline LinPrim(const double *ref, double gamma)
      : gamm(gamma - 1.0), rho(ref[0]), u(ref[1] / ref[0]), v(ref[2] / ref[0]),
        w(ref[3] / ref[0]), vsq(0.0) {
    vsq = 0.5 * (u * u + v * v + w * w);
  }
};

}  // namespace dev
}  // namespace kfvm
