// Device physics for the stage kernels: equation of state, linearised
// primitive map, Riemann solvers.  Every expression keeps the operand order of
// the reference so that, with contraction disabled, results agree with a CPU
// evaluation of the same formulas to the last bit (library pow/exp aside).
//
//   Hydro EOS        src/physics/hydro/Hydro_EOS.H:15-93
//   VarConvert       src/physics/hydro/Hydro_K.H:25-65
//   flux / speeds    src/physics/hydro/Hydro_Riemann.H:18-91, Riemann.H:26-45
//   HLLC             src/physics/hydro/Hydro_Riemann.H:97-187
//   HLL / LLF        src/physics/Common_K.H:23-63
//   LinAdv exact     src/physics/linadv/LinAdv_Riemann.H:63-90
#pragma once

#include <hip/hip_runtime.h>

namespace kfvm {
namespace dev {

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
  template <int DIR>
  __device__ static inline double solve_hllc(double g, const double *uL, const double *uR,
                                             double *F) {
    constexpr int v0 = 1 + DIR, v1 = 1 + (DIR + 1) % 3, v2 = 1 + (DIR + 2) % 3;
    double vL[5], vR[5];
    cons2prim(g, uL, vL);
    cons2prim(g, uR, vR);
    double sL, sR;
    const double sMax = speeds<DIR>(g, uL, uR, vL, vR, sL, sR);
    const double vsqL = vL[1] * vL[1] + vL[2] * vL[2] + vL[3] * vL[3];
    const double vsqR = vR[1] * vR[1] + vR[2] * vR[2] + vR[3] * vR[3];
    const double aL = sound(g, uL), aR = sound(g, uR);
    double M = fmax(sqrt(vsqL) / aL, sqrt(vsqR) / aR);
    const double rhoL = vL[0], rhoR = vR[0], pL = vL[4], pR = vR[4], unL = vL[v0], unR = vR[v0];
    const double sS = ((pR - pL) + (rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR))) /
                      (rhoL * (sL - unL) - rhoR * (sR - unR));
    const double psL0 = rhoL * (sL - unL), psR0 = rhoR * (sR - unR);
    const double pS = 0.5 * ((pL + pR) + psL0 * (sS - unL) + psR0 * (sS - unR));
    const double sFacL = sS / (sL - sS), sFacR = sS / (sR - sS);
    const double uJump = vR[v0] - vL[v0], vJump = vR[v1] - vL[v1], wJump = vR[v2] - vL[v2];
    const double psiL = uL[0] * (sL - vL[v0]), psiR = uR[0] * (sR - vR[v0]),
                 psiFac = psiR * psiL / (psiR - psiL);
    const double presRat = fmin(vL[4] / vR[4], vR[4] / vL[4]);
    const double gFac = 1.0 - pow(presRat, M);
    const double shFacL = psiFac * gFac * sL / (sL - sS), shFacR = psiFac * gFac * sR / (sR - sS);
    M = fmin(M, 1.0);
    const double fM = M * sqrt(4.0 + pow(1.0 - M * M, 2.0)) / (1.0 + M * M);
    const double pFacL = psiFac * (fM - 1.0) * uJump + sL * pS / (sL - sS);
    const double pFacR = psiFac * (fM - 1.0) * uJump + sR * pS / (sR - sS);
    if (0 <= sL) {
      flux<DIR>(uL, vL, F);
    } else if (0 <= sS) {
      flux<DIR>(uL, vL, F);
      F[0] = sFacL * (sL * uL[0] - F[0]);
      F[v0] = sFacL * (sL * uL[v0] - F[v0]) + pFacL;
      F[v1] = sFacL * (sL * uL[v1] - F[v1]) + shFacL * vJump;
      F[v2] = sFacL * (sL * uL[v2] - F[v2]) + shFacL * wJump;
      F[4] = sFacL * (sL * uL[4] - F[4]) + sS * pFacL;
    } else if (0 < sR) {
      flux<DIR>(uR, vR, F);
      F[0] = sFacR * (sR * uR[0] - F[0]);
      F[v0] = sFacR * (sR * uR[v0] - F[v0]) + pFacR;
      F[v1] = sFacR * (sR * uR[v1] - F[v1]) + shFacR * vJump;
      F[v2] = sFacR * (sR * uR[v2] - F[v2]) + shFacR * wJump;
      F[4] = sFacR * (sR * uR[4] - F[4]) + sS * pFacR;
    } else {
      flux<DIR>(uR, vR, F);
    }
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
