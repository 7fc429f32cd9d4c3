// CPU oracle for the KFVM-WENO per-stage update -- TEST INFRASTRUCTURE ONLY.
//
// A C++17/OpenMP restatement of the reference path in the reference's own
// unfused dataflow and memory layout (LayoutRight, halo'd cell arrays, face
// arrays with two slots).  It is the checker for the HIP library and the
// timed CPU baseline; nothing in the product links it.
//
// Arithmetic conventions shared with the HIP path (so that results agree to
// rounding): dot products accumulate with fma in ascending index order
// (the unblocked KokkosBatched gemv/gemm loop order, an assumption: that
// library is not in the reference snapshot), "y += alpha*t" is an fma, and
// everything else is evaluated exactly as written with contraction disabled
// (-ffp-contract=off).
//
// Functions and the reference code they restate:
//   cell_bcs          Solver.cpp:744-783, 866-877, 1085-1095, 1293-1303, 1462-1698;
//                     BoundaryConditions_K.H:53-527
//   face_bcs          Solver.cpp:785-824, 964-976, 1179-1192, 1381-1394, 1700-1946;
//                     BoundaryConditions_K.H:529-1043
//   recon_linear      Stencil_K.H:32-90 (KernelLinearRecon_K)
//   recon_weno_cell   Stencil_K.H:164-283, 416-454, 506-589 (KernelWenoRecon_K)
//   weno_select       Numeric_K.H:292-398, Solver.cpp:687-742
//   pos_pres          Common_K.H:299-643
//   riemann_*         Common_K.H:23-63, 180-297; hydro/Hydro_Riemann.H:18-187;
//                     Riemann.H:26-45; linadv/LinAdv_Riemann.H:18-90
//   integrate_rhs     Numeric_K.H:109-196
//   rk stages / err   Numeric_K.H:198-290, Solver.cpp:188-343
//   eval_rhs          Solver.cpp:410-530
//   set_ic            Numeric_K.H:31-107
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "problems.h"

namespace orc {

// ------------------------------------------------------------------ stencil
// Own enumeration of the radius-2 stencils (Stencil.cpp:47-229).
struct Stencil {
  int dim = 3, C = 0, Cc = 0, Cb = 0, B = 0, nSub = 0, I = 0, F = 0;
  std::vector<int> ox, oy, oz;       // offsets [C]
  std::vector<std::vector<int>> sub; // [B][Cb]

  int find(int i, int j, int k) const {
    for (int n = 0; n < C; n++)
      if (ox[n] == i && oy[n] == j && oz[n] == k) return n;
    return -1;
  }
  void add(int i, int j, int k) {
    ox.push_back(i);
    oy.push_back(j);
    oz.push_back(k);
  }
  explicit Stencil(int d) : dim(d) {
    const int rad = 2;
    if (d == 2) {
      add(0, 0, 0);
      for (int nL = 1; nL <= rad; nL++) {
        add(-nL, 0, 0); add(nL, 0, 0); add(0, -nL, 0); add(0, nL, 0);
        for (int nR = 1; nR <= nL; nR++) {
          double rsq = 2.0 * nR * nR;
          int lyr = int(std::ceil(std::sqrt(rsq)));
          if (rsq <= rad * rad && lyr == nL) {
            add(-nR, -nR, 0); add(-nR, nR, 0); add(nR, -nR, 0); add(nR, nR, 0);
          }
        }
        for (int i = 2; i <= rad; i++)
          for (int j = 1; j < i; j++) {
            double rsq = double(i * i + j * j);
            int lyr = int(std::ceil(std::sqrt(rsq)));
            if (rsq <= rad * rad && lyr == nL) {
              add(-i, j, 0); add(-i, -j, 0); add(i, j, 0); add(i, -j, 0);
              add(-j, i, 0); add(-j, -i, 0); add(j, i, 0); add(j, -i, 0);
            }
          }
      }
      C = int(ox.size());
      Cc = 5; Cb = 5; B = 4; nSub = 6; I = 5; F = 4;
      sub.assign(B, {});
      for (int h = 0; h <= rad; h++) {
        sub[0].push_back(find(-h, 0, 0)); sub[1].push_back(find(h, 0, 0));
        sub[2].push_back(find(0, -h, 0)); sub[3].push_back(find(0, h, 0));
      }
      for (int h = 0; h <= rad; h++)
        for (int l = 1; l <= h; l++)
          if (find(-h, -l, 0) >= 0) {
            sub[0].push_back(find(-h, -l, 0)); sub[0].push_back(find(-h, l, 0));
            sub[1].push_back(find(h, -l, 0)); sub[1].push_back(find(h, l, 0));
            sub[2].push_back(find(-l, -h, 0)); sub[2].push_back(find(l, -h, 0));
            sub[3].push_back(find(-l, h, 0)); sub[3].push_back(find(l, h, 0));
          }
    } else {
      const int rm = rad - 1;
      for (int i = -rm; i <= rm; i++)
        for (int j = -rm; j <= rm; j++)
          for (int k = -rm; k <= rm; k++)
            if (i * i + j * j + k * k <= rm * rm) add(i, j, k);
      for (int i = -rad; i <= rad; i++)
        for (int j = -rad; j <= rad; j++)
          for (int k = -rad; k <= rad; k++) {
            int rsq = i * i + j * j + k * k;
            if (rsq > rm * rm && rsq <= rad * rad) add(i, j, k);
          }
      C = int(ox.size());
      Cc = 7; Cb = 11; B = 6; nSub = 8; I = 19; F = 6;
      sub.assign(B, {});
      for (int n = 0; n < C; n++) {
        int l = ox[n], m = oy[n], q = oz[n];
        if (-l >= std::abs(m) && -l >= std::abs(q)) sub[0].push_back(n);
        if (l >= std::abs(m) && l >= std::abs(q)) sub[1].push_back(n);
        if (-m >= std::abs(l) && -m >= std::abs(q)) sub[2].push_back(n);
        if (m >= std::abs(l) && m >= std::abs(q)) sub[3].push_back(n);
        if (-q >= std::abs(l) && -q >= std::abs(m)) sub[4].push_back(n);
        if (q >= std::abs(l) && q >= std::abs(m)) sub[5].push_back(n);
      }
    }
  }
};

struct RKT {
  int nStages;
  std::vector<double> gam1, gam2, gam3, delta, beta, bhat, c;
  double bhatfsal, ep1, ep2;
};

RKT rk_coeffs(int rk) {
  RKT t;
  if (rk == 2) {  // SSP4_3_2
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
  } else if (rk == 3) {  // SSP10_4_3
    t.nStages = 10;
    t.gam1 = {0.0, 1.0, 1.0, 1.0, 2.0 / 5.0, 1.0, 1.0, 1.0, 1.0, 3.0 / 5.0};
    t.gam2 = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
    t.gam3 = {0.0, 0.0, 0.0, 0.0, 3.0 / 5.0, 0.0, 0.0, 0.0, 0.0, -3.0 / 2.0};
    t.delta = {0.0, 0.0, 0.0, 0.0, 0.0, 9.0 / 10.0, 0.0, 0.0, 0.0, 0.0};
    t.beta = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 15.0,
              1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 10.0};
    t.bhat = {0.0, 2.0 / 9.0, 0.0, 0.0, 5.0 / 18.0, 1.0 / 3.0, 0.0, 0.0, 0.0, 1.0 / 6.0};
    t.bhatfsal = 0.0;
    t.c = {0.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 2.0, 2.0 / 3.0, 1.0 / 3.0, 1.0 / 2.0, 2.0 / 3.0, 5.0 / 6.0, 1.0};
    t.ep1 = 0.8 / 3.0;
    t.ep2 = -0.31 / 3.0;
  } else if (rk == 0) {  // ThreeStarP
    t.nStages = 5;
    t.gam1 = {0.0, 2.587771979725733308135192812685323706e-01, -1.324380360140723382965420909764953437e-01,
              5.056033948190826045833606441415585735e-02, 5.670532000739313812633197158607642990e-01};
    t.gam2 = {0.0, 5.528354909301389892439698870483746541e-01, 6.731871608203061824849561782794643600e-01,
              2.803103963297672407841316576323901761e-01, 5.521525447020610386070346724931300367e-01};
    t.gam3 = {0.0, 0.0, 0.0, 2.752563273304676380891217287572780582e-01, -8.950526174674033822276061734289327568e-01};
    t.delta = {1.0, 3.407655879334525365094815965895763636e-01, 3.414382655003386206551709871126405331e-01,
               7.229275366787987419692007421895451953e-01, 0.0};
    t.beta = {2.300298624518076223899418286314123354e-01, 3.021434166948288809034402119555380003e-01,
              8.025606185416310937583009085873554681e-01, 4.362158943603440930655148245148766471e-01,
              1.129272530455059129782111662594436580e-01};
    t.bhat = {9.484166705035703392326247283838082847e-02, 1.726371339430353766966762629176676070e-01,
              3.998243189084371024483169698618455770e-01, 1.718016807580178450618829007973835152e-01,
              5.881914422155740300718268359027168467e-02};
    t.bhatfsal = 1.020760551185952388626787099944507877e-01;
    t.c = {0.0, 2.300298624518076223899418286314123354e-01, 4.050046072094990912268498160116125481e-01,
           8.947822893693433545220710894560512805e-01, 7.235136928826589010272834603680114769e-01};
    t.ep1 = 0.7 / 3.0;
    t.ep2 = -0.23 / 3.0;
  } else {  // FourStarP
    t.nStages = 9;
    t.gam1 = {0.0, -4.655641447335068552684422206224169103e+00, -7.720265099645871829248487209517314217e-01,
              -4.024436690519806086742256154738379161e+00, -2.129676284018530966221583708648634733e-02,
              -2.435022509790109546199372365866450709e+00, 1.985627297131987000579523283542615256e-02,
              -2.810791146791038566946663374735713961e-01, 1.689434168754859644351230590422137972e-01};
    t.gam2 = {0.0, 2.499262792574495009336242992898153462e+00, 5.866820377718875577451517985847920081e-01,
              1.205146086523094569925592464380295241e+00, 3.474793722186732780030762737753849272e-01,
              1.321346060965113109321230804210670518e+00, 3.119636464694193615946633676950358444e-01,
              4.351419539684379261368971206040518552e-01, 2.359698130028753572503744518147537768e-01};
    t.gam3 = {0.0, 0.0, 0.0, 7.621006678721315291614677352949377871e-01, -1.981182504339400567765766904309673119e-01,
              -6.228959218699007450469629366684127462e-01, -3.752248380775956442989480369774937099e-01,
              -3.355438309135169811915662336248989661e-01, -4.560955005031121479972862973705108039e-02};
    t.delta = {1.0, 1.262923876648114432874834923838556100e+00, 7.574967189685911558308119415539596711e-01,
               5.163589453140728104667573195005629833e-01, -2.746327421802609557034437892013640319e-02,
               -4.382673178127944142238606608356542890e-01, 1.273587294602656522645691372699677063e+00,
               -6.294740283927400326554066998751383342e-01, 0.0};
    t.beta = {2.836343005184365275160654678626695428e-01, 9.736500104654741223716056170419660217e-01,
              3.382359225242515288768487569778320563e-01, -3.584943611106183357043212309791897386e-01,
              -4.113944068471528211627210454497620358e-03, 1.427968894048586363415504654313371031e+00,
              1.808470948394314017665968411915568633e-02, 1.605770645946802213926893453819236685e-01,
              2.952227015964591648775833803635147962e-01};
    t.bhat = {2.483675912451591196775756814283216443e-02, 1.866327774562103796990092260942180726e-01,
              5.671080795936984495604436622517631183e-02, -3.447695439149287702616943808570747099e-03,
              3.602245056516636472203469198006404016e-03, 4.545570622145088936800484247980581766e-01,
              -2.434665289427612407531544765622888855e-04, 6.642755361103549971517945063138312147e-02,
              1.613697079523505006226025497715177578e-01};
    t.bhatfsal = 4.955424859358438183052504342394102722e-02;
    t.c = {0.0, 2.836343005184365275160654678626695428e-01, 5.484076570002894365286665352032296535e-01,
           3.687228761669438493478872632332010073e-01, -6.806126440140844191258463830024463902e-01,
           3.518526124230705801739919476290327750e-01, 1.665941994879593315477304663913129942e+00,
           9.715279295934715835299192116436237065e-01, 9.051569840159589594903399929316959062e-01};
    t.ep1 = 0.38 / 4.0;
    t.ep2 = -0.18 / 4.0;
  }
  return t;
}

// ----------------------------------------------------------------- physics
struct Hydro {
  static double ekin(const double *U) {
    return 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
  }
  static double eint(const double *U) { return U[4] - ekin(U); }
  static double pres(double g, const double *U) { return (g - 1.0) * eint(U); }
  static double sound(double g, const double *U) { return std::sqrt(g * pres(g, U) / U[0]); }
  static double weno(double g, const double *U) { return pres(g, U) / std::pow(U[0], g); }
  static double positivity(const double *U) { return std::fmin(U[0], eint(U)); }
  static void cons2prim(double g, const double *U, double *V) {
    V[0] = U[0];
    V[1] = U[1] / U[0];
    V[2] = U[2] / U[0];
    V[3] = U[3] / U[0];
    V[4] = pres(g, U);
  }
  static void flux(const int *ord, const double *U, const double *V, double *F) {
    const int v0 = 1 + ord[0], v1 = 1 + ord[1], v2 = 1 + ord[2];
    F[0] = U[v0];
    F[v0] = U[v0] * V[v0] + V[4];
    F[v1] = U[v0] * V[v1];
    F[v2] = U[v0] * V[v2];
    F[4] = V[v0] * (U[4] + V[4]);
  }
  static double roe(double rL, double rR, double qL, double qR) {
    return (qL * std::sqrt(rL) + qR * std::sqrt(rR)) / (std::sqrt(rL) + std::sqrt(rR));
  }
  static double speeds(double g, const int *ord, const double *uL, const double *uR,
                       const double *vL, const double *vR, double &sL, double &sR) {
    const int v0 = 1 + ord[0], v1 = 1 + ord[1], v2 = 1 + ord[2];
    const double rhoL = uL[0], rhoR = uR[0];
    const double hL = (uL[4] + vL[4]) / rhoL, hR = (uR[4] + vR[4]) / rhoR;
    const double aL = sound(g, uL), aR = sound(g, uR);
    const double uTil = roe(rhoL, rhoR, vL[v0], vR[v0]), vTil = roe(rhoL, rhoR, vL[v1], vR[v1]),
                 wTil = roe(rhoL, rhoR, vL[v2], vR[v2]),
                 vsqTil = uTil * uTil + vTil * vTil + wTil * wTil, hTil = roe(rhoL, rhoR, hL, hR),
                 aTil = std::sqrt((g - 1.0) * (hTil - 0.5 * vsqTil));
    sL = std::fmin(vL[v0] - aL, uTil - aTil);
    sR = std::fmax(vR[v0] + aR, uTil + aTil);
    return std::fmax(std::fabs(sL), std::fabs(sR));
  }
  // Generic HLL / LLF (Common_K.H:23-63)
  static double solve_hll(double g, bool llf, int dir, const double *uL, const double *uR,
                          const double *vL, const double *vR, double *F) {
    const int ord[3] = {dir, (dir + 1) % 3, (dir + 2) % 3};
    double sL, sR;
    const double sMax = speeds(g, ord, uL, uR, vL, vR, sL, sR);
    if (llf) {
      sL = -sMax;
      sR = sMax;
    }
    const double tR = (std::fmin(0.0, sR) - std::fmin(0.0, sL)) / (sR - sL), tL = 1.0 - tR;
    const double tD = 0.5 * (sR * std::fabs(sL) - sL * std::fabs(sR)) / (sR - sL);
    double FR[5];
    flux(ord, uL, vL, F);
    flux(ord, uR, vR, FR);
    for (int v = 0; v < 5; v++) F[v] = tL * F[v] + tR * FR[v] - tD * (uR[v] - uL[v]);
    return sMax;
  }
  // HLLC with the low-Mach and shear fixes (Hydro_Riemann.H:97-187)
  static double solve_hllc(double g, int dir, const double *uL, const double *uR,
                           const double *vL, const double *vR, double *F) {
    const int ord[3] = {dir, (dir + 1) % 3, (dir + 2) % 3};
    const int v0 = 1 + ord[0], v1 = 1 + ord[1], v2 = 1 + ord[2];
    double sL, sR;
    const double sMax = speeds(g, ord, uL, uR, vL, vR, sL, sR);
    const double vsqL = vL[1] * vL[1] + vL[2] * vL[2] + vL[3] * vL[3];
    const double vsqR = vR[1] * vR[1] + vR[2] * vR[2] + vR[3] * vR[3];
    const double aL = sound(g, uL), aR = sound(g, uR);
    double M = std::fmax(std::sqrt(vsqL) / aL, std::sqrt(vsqR) / aR);
    const double rhoL = vL[0], rhoR = vR[0], pL = vL[4], pR = vR[4], unL = vL[v0], unR = vR[v0];
    const double sS = ((pR - pL) + (rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR))) /
                      (rhoL * (sL - unL) - rhoR * (sR - unR));
    const double psL0 = rhoL * (sL - unL), psR0 = rhoR * (sR - unR);
    const double pS = 0.5 * ((pL + pR) + psL0 * (sS - unL) + psR0 * (sS - unR));
    const double sFacL = sS / (sL - sS), sFacR = sS / (sR - sS);
    const double uJump = vR[v0] - vL[v0], vJump = vR[v1] - vL[v1], wJump = vR[v2] - vL[v2];
    const double psiL = uL[0] * (sL - vL[v0]), psiR = uR[0] * (sR - vR[v0]),
                 psiFac = psiR * psiL / (psiR - psiL);
    const double presRat = std::fmin(vL[4] / vR[4], vR[4] / vL[4]);
    const double gFac = 1.0 - std::pow(presRat, M);
    const double shFacL = psiFac * gFac * sL / (sL - sS), shFacR = psiFac * gFac * sR / (sR - sS);
    M = std::fmin(M, 1.0);
    const double fM = M * std::sqrt(4.0 + std::pow(1.0 - M * M, 2)) / (1.0 + M * M);
    const double pFacL = psiFac * (fM - 1.0) * uJump + sL * pS / (sL - sS);
    const double pFacR = psiFac * (fM - 1.0) * uJump + sR * pS / (sR - sS);
    if (0 <= sL) {
      flux(ord, uL, vL, F);
    } else if (0 <= sS) {
      flux(ord, uL, vL, F);
      F[0] = sFacL * (sL * uL[0] - F[0]);
      F[v0] = sFacL * (sL * uL[v0] - F[v0]) + pFacL;
      F[v1] = sFacL * (sL * uL[v1] - F[v1]) + shFacL * vJump;
      F[v2] = sFacL * (sL * uL[v2] - F[v2]) + shFacL * wJump;
      F[4] = sFacL * (sL * uL[4] - F[4]) + sS * pFacL;
    } else if (0 < sR) {
      flux(ord, uR, vR, F);
      F[0] = sFacR * (sR * uR[0] - F[0]);
      F[v0] = sFacR * (sR * uR[v0] - F[v0]) + pFacR;
      F[v1] = sFacR * (sR * uR[v1] - F[v1]) + shFacR * vJump;
      F[v2] = sFacR * (sR * uR[v2] - F[v2]) + shFacR * wJump;
      F[4] = sFacR * (sR * uR[4] - F[4]) + sS * pFacR;
    } else {
      flux(ord, uR, vR, F);
    }
    return sMax;
  }
};

// ------------------------------------------------------------------ solver
// ---- GLM-MHD physics (MHD_GLM_EOS.H:15-144, MHD_GLM_Riemann.H:18-231)
struct Mhd {
  static double ekin(const double *U) { return 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]; }
  static double prsb(const double *U) { return 0.5 * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]); }
  static double eint(const double *U) { return (U[4] - prsb(U) - ekin(U) - 0.5 * U[8] * U[8]); }
  static double pres(double g, const double *U) { return (g - 1.0) * eint(U); }
  static double sound(double g, const double *U) { return std::sqrt(g * pres(g, U) / U[0]); }
  static double weno(double g, const double *U) { return pres(g, U) / std::pow(U[0], g); }
  static double positivity(const double *U) { return std::fmin(U[0], eint(U)); }
  static double fast(double g, const double *U, double bl) {
    const double asq = g * pres(g, U) / U[0];
    const double bsq = 2.0 * prsb(U) / U[0];
    const double blsq = bl * bl / U[0];
    const double absq = asq + bsq;
    return std::sqrt(0.5 * (absq + std::sqrt(absq * absq - 4.0 * asq * blsq)));
  }
  static void cons2prim(double g, const double *U, double *V) {
    V[0] = U[0];
    V[1] = U[1] / U[0];
    V[2] = U[2] / U[0];
    V[3] = U[3] / U[0];
    V[4] = pres(g, U);
    for (int v = 5; v < 9; v++) V[v] = U[v];
  }
  static void flux(double ch, const int *ord, const double *U, const double *V, double *F) {
    const int v0 = 1 + ord[0], v1 = 1 + ord[1], v2 = 1 + ord[2];
    const int b0 = 5 + ord[0], b1 = 5 + ord[1], b2 = 5 + ord[2];
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
  static double speeds(double g, double ch, const int *ord, const double *uL, const double *uR,
                       const double *vL, const double *vR, double &sL, double &sR) {
    const int v0 = 1 + ord[0], b0 = 5 + ord[0], b1 = 5 + ord[1], b2 = 5 + ord[2];
    const double cfL = fast(g, uL, uL[b0]);
    const double cfR = fast(g, uR, uR[b0]);
    const double sdL = std::fmin(vL[v0] - std::fmax(ch, cfL), vR[v0] - std::fmax(ch, cfR));
    const double sdR = std::fmax(vL[v0] + std::fmax(ch, cfL), vR[v0] + std::fmax(ch, cfR));
    const double rtL = std::sqrt(uL[0]), rtR = std::sqrt(uR[0]);
    const double rtSum = rtL + rtR, vSum = (rtL * vL[v0] + rtR * vR[v0]) / rtSum;
    const double bd[] = {uL[b0] - uR[b0], uL[b1] - uR[b1], uL[b2] - uR[b2]};
    const double bnDiff = std::sqrt(bd[0] * bd[0] + bd[1] * bd[1] + bd[2] * bd[2]) / rtSum;
    const double aL = std::fmin(vL[v0], vSum) - cfL - bnDiff;
    const double aR = std::fmin(vR[v0], vSum) + cfR + bnDiff;  // fmin as in the reference
    sL = std::fmin(sdL, aL);
    sR = std::fmax(sdR, aR);
    return std::fmax(std::fabs(sL), std::fabs(sR));
  }
  static double solve_hll(double g, double ch, bool llf, int dir, const double *uL, const double *uR,
                          const double *vL, const double *vR, double *F) {
    const int ord[3] = {dir, (dir + 1) % 3, (dir + 2) % 3};
    double sL, sR;
    const double sMax = speeds(g, ch, ord, uL, uR, vL, vR, sL, sR);
    if (llf) {
      sL = -sMax;
      sR = sMax;
    }
    const double tR = (std::fmin(0.0, sR) - std::fmin(0.0, sL)) / (sR - sL), tL = 1.0 - tR;
    const double tD = 0.5 * (sR * std::fabs(sL) - sL * std::fabs(sR)) / (sR - sL);
    double FR[9];
    flux(ch, ord, uL, vL, F);
    flux(ch, ord, uR, vR, FR);
    for (int v = 0; v < 9; v++) F[v] = tL * F[v] + tR * FR[v] - tD * (uR[v] - uL[v]);
    return sMax;
  }
  static double logMean(double a, double b) {
    const double zeta = a / b, f = (zeta - 1.0) / (zeta + 1.0), u = f * f;
    const double F = u < 0.01 ? 1.0 + u / 3.0 + u * u / 5.0 + u * u * u / 7.0 : 0.5 * std::log(zeta) / f;
    return 0.5 * (a + b) / F;
  }
  static double am(double a, double b) { return 0.5 * (a + b); }
  static double solve_kepes(double g, double ch, int dir, const double *uL, const double *uR,
                            const double *vL, const double *vR, double *F) {
    const int ord[3] = {dir, (dir + 1) % 3, (dir + 2) % 3};
    const int v0 = 1 + ord[0], v1 = 1 + ord[1], v2 = 1 + ord[2];
    const int b0 = 5 + ord[0], b1 = 5 + ord[1], b2 = 5 + ord[2];
    double sL, sR;
    const double sMax = speeds(g, ch, ord, uL, uR, vL, vR, sL, sR);
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
    for (int v = 0; v < 9; v++) F[v] -= sMax * (uR[v] - uL[v]) / 2.0;
    return sMax;
  }
};

// RKFSAL_ErrEst_K (Numeric_K.H:262-290) for one cell: adds the squared scaled
// differences (u - uhat) / (atol + rtol max(|u|, |uhat|)) to err and lowers
// posFlag to the cell's positivity (-1 when u holds a NaN).  eq 0 hydro,
// 1 linear advection, 2 GLM-MHD.
inline void err_est_cell(int nv, int eq, const double *u, const double *uh, double atol, double rtol,
                         double &err, double &posFlag) {
  bool hasNan = false;
  for (int v = 0; v < nv; v++) {
    hasNan = hasNan || std::isnan(u[v]);
    err += std::pow((u[v] - uh[v]) / (atol + rtol * std::fmax(std::fabs(u[v]), std::fabs(uh[v]))), 2);
  }
  const double pos = hasNan ? -1.0 : (eq == 0 ? Hydro::positivity(u) : (eq == 2 ? Mhd::positivity(u) : 1.0));
  posFlag = posFlag > pos ? pos : posFlag;
}

struct Base {
  virtual ~Base() {}
  virtual void set_state(const double *) = 0;
  virtual void get_state(double *) const = 0;
  virtual int set_ic(const char *, const double *, uint64_t) = 0;
  virtual double eval_rhs(int which) = 0;
  virtual void get_rhs(double *) const = 0;
  virtual void get_flags(double *) const = 0;
  virtual void set_flags(const double *) = 0;
  virtual void prime() = 0;
  virtual int take_step() = 0;
  virtual void set_mhd_tables(const double *, const double *) {}
  virtual void face_states(int, double *) {}
  virtual void weno_omegas(int64_t, int64_t, int64_t, double *) {}
  orc_config cfg{};
  double time = 0.0, dt = 0.0, errEst = 1.0;
  int64_t nTS = 1, nRhsEval = 0, nWeno = 0;
  bool lastTimeStep = false, useSparseWeno = true;
};

template <int D, int NV>
struct Solver : Base {
  static constexpr int r = 2;
  Stencil st;
  int NQ, Q, F;
  int64_t nx, ny, nz, hx, hy, hz;
  double dx, dy, dz, dmin, ab[5], wt[5];
  double vel[3];
  std::vector<double> faceW, derivW;
  std::vector<double> U, Uprev, Uhat, Utmp, RHS, flag;
  std::vector<double> fx, fy, fz;
  std::vector<char> wflag;  // cells needing WENO in sparse mode
  RKT rk;
  // GLM-MHD: interior quadrature points (states, then sources) [cell][NI][NV],
  // interior weights [nSub][NI][C], differentiation matrix [NQ][NQ+2], c_h
  int NI = 0, NZ = 1;
  std::vector<double> src, cellW, Dm;
  double ch = 0.0;

  explicit Solver(const orc_config &c, const double *fW, const double *dW) : st(D) {
    cfg = c;
    NQ = c.nq;
    Q = D == 3 ? NQ * NQ : NQ;
    F = 2 * D;
    nx = c.n[0];
    ny = c.n[1];
    nz = D == 3 ? c.n[2] : 1;
    hx = nx + 2 * r;
    hy = ny + 2 * r;
    hz = D == 3 ? nz + 2 * r : 1;
    dx = (c.hi[0] - c.lo[0]) / double(nx);
    dy = (c.hi[1] - c.lo[1]) / double(ny);
    dz = D == 3 ? (c.hi[2] - c.lo[2]) / double(nz) : 1.0;
    dmin = D == 3 ? std::fmin(dx, std::fmin(dy, dz)) : std::fmin(dx, dy);
    if (NQ == 2) {
      ab[0] = -0.577350269189625764509148 / 2.0;
      ab[1] = 0.577350269189625764509148 / 2.0;
      wt[0] = 1.0 / 2.0;
      wt[1] = 1.0 / 2.0;
    } else {
      ab[0] = 0.0;
      ab[1] = -0.7745966692414833770358530 / 2.0;
      ab[2] = 0.7745966692414833770358530 / 2.0;
      wt[0] = 0.8888888888888888888888888 / 2.0;
      wt[1] = 0.5555555555555555555555555 / 2.0;
      wt[2] = 0.5555555555555555555555555 / 2.0;
    }
    if (c.haveAdvVel) {
      vel[0] = c.advVel[0]; vel[1] = c.advVel[1]; vel[2] = c.advVel[2];
    } else {
      vel[0] = std::cos(c.gamma); vel[1] = std::sin(c.gamma); vel[2] = 0.0;
    }
    faceW.assign(fW, fW + size_t(st.nSub) * F * Q * st.C);
    derivW.assign(dW, dW + size_t(st.nSub) * st.I * st.C);
    U.assign(size_t(hx) * hy * hz * NV, 0.0);
    Uprev = U;
    const size_t nc = size_t(nx) * ny * nz;
    Uhat.assign(nc * NV, 0.0);
    Utmp = Uhat;
    RHS = Uhat;
    flag.assign(nc * 2, 0.0);
    for (size_t i = 0; i < nc; i++) flag[2 * i + 1] = 1.0;
    fx.assign(size_t(nx + 1) * ny * nz * 2 * Q * NV, 0.0);
    fy.assign(size_t(nx) * (ny + 1) * nz * 2 * Q * NV, 0.0);
    if (D == 3) fz.assign(size_t(nx) * ny * (nz + 1) * 2 * Q * NV, 0.0);
    wflag.assign(nc, 0);
    NZ = D == 3 ? NQ : 1;
    NI = NQ * NQ * NZ;
    if (NV == 9) src.assign(nc * NI * NV, 0.0);
    rk = rk_coeffs(c.rk);
    useSparseWeno = c.wenoThresh > 0.0;
  }

  // ---- indexing
  size_t ui(int64_t i, int64_t j, int64_t k) const { return ((size_t(i) * hy + j) * hz + k) * NV; }
  size_t uc(int64_t i, int64_t j, int64_t k) const {  // interior cell -> halo'd index
    return D == 3 ? ui(i + r, j + r, k + r) : ui(i + r, j + r, 0);
  }
  size_t ci(int64_t i, int64_t j, int64_t k) const { return (size_t(i) * ny + j) * nz + k; }
  double *FX(int64_t i, int64_t j, int64_t k, int s, int q) {
    return &fx[(((((size_t(i) * ny + j) * nz + k) * 2 + s) * Q + q) * NV)];
  }
  double *FY(int64_t i, int64_t j, int64_t k, int s, int q) {
    return &fy[(((((size_t(i) * (ny + 1) + j) * nz + k) * 2 + s) * Q + q) * NV)];
  }
  double *FZ(int64_t i, int64_t j, int64_t k, int s, int q) {
    return &fz[(((((size_t(i) * ny + j) * (nz + 1) + k) * 2 + s) * Q + q) * NV)];
  }
  // own face slots of interior cell (i,j,k), face order W,E,S,N,B,T
  double *own(int f, int64_t i, int64_t j, int64_t k, int q) {
    switch (f) {
    case 0: return FX(i, j, k, 1, q);
    case 1: return FX(i + 1, j, k, 0, q);
    case 2: return FY(i, j, k, 1, q);
    case 3: return FY(i, j + 1, k, 0, q);
    case 4: return FZ(i, j, k, 1, q);
    default: return FZ(i, j, k + 1, 0, q);
    }
  }
  const double *fw(int s, int f, int q) const { return &faceW[((size_t(s) * F + f) * Q + q) * st.C]; }
  const double *dw(int s, int i) const { return &derivW[(size_t(s) * st.I + i) * st.C]; }

  // ---- physics dispatch
  static constexpr bool hydro = NV == 5;
  static constexpr bool mhd = NV == 9;
  double gam() const { return cfg.gamma; }
  static double eintP(const double *s) { return mhd ? Mhd::eint(s) : Hydro::eint(s); }
  double soundP(const double *s) const { return mhd ? Mhd::sound(gam(), s) : Hydro::sound(gam(), s); }
  double wenoP(const double *s) const { return mhd ? Mhd::weno(gam(), s) : Hydro::weno(gam(), s); }
  double *SP(int64_t i, int64_t j, int64_t k, int p) { return &src[(ci(i, j, k) * NI + p) * NV]; }
  void set_mhd_tables(const double *cw, const double *dm) override {
    cellW.assign(cw, cw + size_t(st.nSub) * NI * st.C);
    Dm.assign(dm, dm + size_t(NQ) * (NQ + 2));
  }

  // reflecting / noslip sign of variable v for axis a
  double bc_sign(int b, int a, int v) const {
    if (!hydro && !mhd) return 1.0;
    if (mhd && v == 5 + a) return -1.0;                   // normal field (MHD_GLM_BCCoeffs.H)
    if (b == 3) return (v >= 1 && v <= 3) ? -1.0 : 1.0;  // noslip
    return (v == 1 + a) ? -1.0 : 1.0;                     // reflecting
  }

  // ---- cell BCs (EW, then NS, then TB)
  void cell_bcs(std::vector<double> &V) {
    const int64_t kLo = D == 3 ? r : 0, kHi = D == 3 ? nz + r : 1;
    // EW
    if (cfg.bc[0] == 0) {
#pragma omp parallel for collapse(2)
      for (int64_t j = r; j < ny + r; j++)
        for (int64_t k = kLo; k < kHi; k++)
          for (int64_t i = 0; i < r; i++)
            for (int v = 0; v < NV; v++) {
              V[ui(nx + r + i, j, k) + v] = V[ui(r + i, j, k) + v];
              V[ui(i, j, k) + v] = V[ui(nx + i, j, k) + v];
            }
    } else {
#pragma omp parallel for collapse(2)
      for (int64_t j = r; j < ny + r; j++)
        for (int64_t k = kLo; k < kHi; k++)
          for (int64_t h = 0; h < r; h++)
            for (int v = 0; v < NV; v++) {
              const double sw = cfg.bc[0] == 1 ? 1.0 : bc_sign(cfg.bc[0], 0, v);
              const double se = cfg.bc[1] == 1 ? 1.0 : bc_sign(cfg.bc[1], 0, v);
              V[ui(r - h - 1, j, k) + v] = cfg.bc[0] == 1 ? V[ui(r + h, j, k) + v] : sw * V[ui(r + h, j, k) + v];
              V[ui(nx + r + h, j, k) + v] = cfg.bc[1] == 1 ? V[ui(nx + r - h - 1, j, k) + v]
                                                           : se * V[ui(nx + r - h - 1, j, k) + v];
            }
    }
    // NS (all x including halo)
    if (cfg.bc[2] == 0) {
#pragma omp parallel for collapse(2)
      for (int64_t i = 0; i < hx; i++)
        for (int64_t k = kLo; k < kHi; k++)
          for (int64_t j = 0; j < r; j++)
            for (int v = 0; v < NV; v++) {
              V[ui(i, ny + r + j, k) + v] = V[ui(i, r + j, k) + v];
              V[ui(i, j, k) + v] = V[ui(i, ny + j, k) + v];
            }
    } else {
#pragma omp parallel for collapse(2)
      for (int64_t i = 0; i < hx; i++)
        for (int64_t k = kLo; k < kHi; k++)
          for (int64_t h = 0; h < r; h++)
            for (int v = 0; v < NV; v++) {
              const double ss = bc_sign(cfg.bc[2], 1, v), sn = bc_sign(cfg.bc[3], 1, v);
              V[ui(i, r - h - 1, k) + v] = cfg.bc[2] == 1 ? V[ui(i, r + h, k) + v] : ss * V[ui(i, r + h, k) + v];
              V[ui(i, ny + r + h, k) + v] = cfg.bc[3] == 1 ? V[ui(i, ny + r - h - 1, k) + v]
                                                           : sn * V[ui(i, ny + r - h - 1, k) + v];
            }
    }
    if (D == 3) {
      if (cfg.bc[4] == 0) {
#pragma omp parallel for collapse(2)
        for (int64_t i = 0; i < hx; i++)
          for (int64_t j = 0; j < hy; j++)
            for (int64_t k = 0; k < r; k++)
              for (int v = 0; v < NV; v++) {
                V[ui(i, j, nz + r + k) + v] = V[ui(i, j, r + k) + v];
                V[ui(i, j, k) + v] = V[ui(i, j, nz + k) + v];
              }
      } else {
#pragma omp parallel for collapse(2)
        for (int64_t i = 0; i < hx; i++)
          for (int64_t j = 0; j < hy; j++)
            for (int64_t h = 0; h < r; h++)
              for (int v = 0; v < NV; v++) {
                const double sb = bc_sign(cfg.bc[4], 2, v), stp = bc_sign(cfg.bc[5], 2, v);
                V[ui(i, j, r - h - 1) + v] = cfg.bc[4] == 1 ? V[ui(i, j, r + h) + v] : sb * V[ui(i, j, r + h) + v];
                V[ui(i, j, nz + r + h) + v] = cfg.bc[5] == 1 ? V[ui(i, j, nz + r - h - 1) + v]
                                                             : stp * V[ui(i, j, nz + r - h - 1) + v];
              }
      }
    }
  }

  // ---- face BCs
  void face_bcs() {
    // x faces
#pragma omp parallel for collapse(2)
    for (int64_t j = 0; j < ny; j++)
      for (int64_t k = 0; k < nz; k++)
        for (int q = 0; q < Q; q++)
          for (int v = 0; v < NV; v++) {
            if (cfg.bc[0] == 0) {
              FX(nx, j, k, 1, q)[v] = FX(0, j, k, 1, q)[v];
              FX(0, j, k, 0, q)[v] = FX(nx, j, k, 0, q)[v];
            } else {
              const double sw = cfg.bc[0] == 1 ? 1.0 : bc_sign(cfg.bc[0], 0, v);
              const double se = cfg.bc[1] == 1 ? 1.0 : bc_sign(cfg.bc[1], 0, v);
              FX(0, j, k, 0, q)[v] = cfg.bc[0] == 1 ? FX(0, j, k, 1, q)[v] : sw * FX(0, j, k, 1, q)[v];
              FX(nx, j, k, 1, q)[v] = cfg.bc[1] == 1 ? FX(nx, j, k, 0, q)[v] : se * FX(nx, j, k, 0, q)[v];
            }
          }
#pragma omp parallel for collapse(2)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t k = 0; k < nz; k++)
        for (int q = 0; q < Q; q++)
          for (int v = 0; v < NV; v++) {
            if (cfg.bc[2] == 0) {
              FY(i, ny, k, 1, q)[v] = FY(i, 0, k, 1, q)[v];
              FY(i, 0, k, 0, q)[v] = FY(i, ny, k, 0, q)[v];
            } else {
              const double ss = cfg.bc[2] == 1 ? 1.0 : bc_sign(cfg.bc[2], 1, v);
              const double sn = cfg.bc[3] == 1 ? 1.0 : bc_sign(cfg.bc[3], 1, v);
              FY(i, 0, k, 0, q)[v] = cfg.bc[2] == 1 ? FY(i, 0, k, 1, q)[v] : ss * FY(i, 0, k, 1, q)[v];
              FY(i, ny, k, 1, q)[v] = cfg.bc[3] == 1 ? FY(i, ny, k, 0, q)[v] : sn * FY(i, ny, k, 0, q)[v];
            }
          }
    if (D == 3) {
#pragma omp parallel for collapse(2)
      for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
          for (int q = 0; q < Q; q++)
            for (int v = 0; v < NV; v++) {
              if (cfg.bc[4] == 0) {
                FZ(i, j, nz, 1, q)[v] = FZ(i, j, 0, 1, q)[v];
                FZ(i, j, 0, 0, q)[v] = FZ(i, j, nz, 0, q)[v];
              } else {
                const double sb = cfg.bc[4] == 1 ? 1.0 : bc_sign(cfg.bc[4], 2, v);
                const double stp = cfg.bc[5] == 1 ? 1.0 : bc_sign(cfg.bc[5], 2, v);
                FZ(i, j, 0, 0, q)[v] = cfg.bc[4] == 1 ? FZ(i, j, 0, 1, q)[v] : sb * FZ(i, j, 0, 1, q)[v];
                FZ(i, j, nz, 1, q)[v] = cfg.bc[5] == 1 ? FZ(i, j, nz, 0, q)[v] : stp * FZ(i, j, nz, 0, q)[v];
              }
            }
    }
  }

  const double *cellp(const std::vector<double> &V, int64_t i, int64_t j, int64_t k, int c) const {
    return &V[D == 3 ? ui(i + r + st.ox[c], j + r + st.oy[c], k + r + st.oz[c])
                     : ui(i + r + st.ox[c], j + r + st.oy[c], 0)];
  }

  // ---- linear reconstruction in conserved variables (KernelLinearRecon_K)
  void recon_linear(const std::vector<double> &V) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int f = 0; f < F; f++)
            for (int q = 0; q < Q; q++) {
              const double *w = fw(0, f, q);
              double *o = own(f, i, j, k, q);
              for (int v = 0; v < NV; v++) {
                double acc = 0.0;
                for (int c = 0; c < st.C; c++) acc = std::fma(w[c], cellp(V, i, j, k, c)[v], acc);
                o[v] = acc;
              }
            }
    if (mhd) {  // interior points of the source terms (Stencil_K.H:92-118)
#pragma omp parallel for collapse(2) schedule(static)
      for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
          for (int64_t k = 0; k < nz; k++)
            for (int p = 0; p < NI; p++) {
              const double *w = &cellW[size_t(p) * st.C];
              double *o = SP(i, j, k, p);
              for (int v = 0; v < NV; v++) {
                double acc = 0.0;
                for (int c = 0; c < st.C; c++) acc = std::fma(w[c], cellp(V, i, j, k, c)[v], acc);
                o[v] = acc;
              }
            }
    }
  }

  // ---- KFVM-WENO for one cell (KernelWenoRecon_K)
  void recon_weno_cell(const std::vector<double> &V, int64_t i, int64_t j, int64_t k,
                       double *omegasOut = nullptr) {
    const int C = st.C, Cc = st.Cc, Cb = st.Cb, nSub = st.nSub, I = st.I;
    double Vf[33][NV];
    const double *ref = &V[uc(i, j, k)];
    // linearised-primitive context about the centre cell (Hydro_K.H:25-65)
    double gamm = 0, rho = 1, u = 0, vv = 0, w = 0, vsq = 0, bx = 0, by = 0, bz = 0, rPsi = 0;
    if (hydro) {
      gamm = gam() - 1.0;
      rho = ref[0];
      u = ref[1] / rho;
      vv = ref[2] / rho;
      w = ref[3] / rho;
      vsq = 0.5 * (u * u + vv * vv + w * w);
    }
    if (mhd) {  // VarConvert<MHD_GLM> (MHD_GLM_K.H:24-69)
      gamm = gam() - 1.0;
      rho = ref[0];
      u = ref[1] / rho;
      vv = ref[2] / rho;
      w = ref[3] / rho;
      bx = ref[5];
      by = ref[6];
      bz = ref[7];
      rPsi = ref[8];
      vsq = (u * u + vv * vv + w * w) / 2.0;
    }
    for (int c = 0; c < C; c++) {
      const double *s = cellp(V, i, j, k, c);
      for (int v = 0; v < NV; v++) Vf[c][v] = s[v];
      if (hydro) {
        Vf[c][4] = gamm * (vsq * Vf[c][0] - u * Vf[c][1] - vv * Vf[c][2] - w * Vf[c][3] + Vf[c][4]);
        Vf[c][1] = (Vf[c][1] - u * Vf[c][0]) / rho;
        Vf[c][2] = (Vf[c][2] - vv * Vf[c][0]) / rho;
        Vf[c][3] = (Vf[c][3] - w * Vf[c][0]) / rho;
      }
      if (mhd) {
        Vf[c][4] = gamm * (vsq * Vf[c][0] + Vf[c][4] - u * Vf[c][1] - vv * Vf[c][2] - w * Vf[c][3] -
                           bx * Vf[c][5] - by * Vf[c][6] - bz * Vf[c][7] - rPsi * Vf[c][8]);
        Vf[c][1] = (Vf[c][1] - u * Vf[c][0]) / rho;
        Vf[c][2] = (Vf[c][2] - vv * Vf[c][0]) / rho;
        Vf[c][3] = (Vf[c][3] - w * Vf[c][0]) / rho;
      }
    }
    // smoothness indicators (findSmoothness)
    double om[8][NV] = {};
    for (int s = 0; s < nSub; s++) {
      const int n = s == 0 ? C : (s == 1 ? Cc : Cb);
      for (int v = 0; v < NV; v++) om[s][v] = 0.0;
      for (int ii = 0; ii < I; ii++) {
        const double *d = dw(s, ii);
        for (int v = 0; v < NV; v++) {
          double t = 0.0;
          for (int c = 0; c < n; c++) {
            const int cc = s < 2 ? c : st.sub[s - 2][c];
            t = std::fma(d[c], Vf[cc][v], t);
          }
          om[s][v] = std::fma(t, t, om[s][v]);
        }
      }
    }
    // nonlinear weights (createOmegas)
    const double gHi = 0.7, gLo = 0.7;
    double gam_[8];
    gam_[0] = gHi;
    gam_[1] = (1.0 - gHi) * gLo;
    for (int s = 2; s < nSub; s++) gam_[s] = (1.0 - gHi) * (1.0 - gLo) / double(2 * D);
    for (int v = 0; v < NV; v++) {
      double tot = 0.0;
      for (int s = 0; s < nSub; s++) {
        om[s][v] = gam_[s] / (om[s][v] * om[s][v] + 1.e-40);
        tot += om[s][v];
      }
      for (int s = 0; s < nSub; s++) om[s][v] /= tot;
      om[0][v] /= gam_[0];
      for (int s = 1; s < nSub; s++) om[s][v] -= om[0][v] * gam_[s];
    }
    if (omegasOut)
      for (int s = 0; s < nSub; s++)
        for (int v = 0; v < NV; v++) omegasOut[s * NV + v] = om[s][v];
    // face states (fillRiemannStates) then back to conserved variables
    for (int f = 0; f < F; f++)
      for (int q = 0; q < Q; q++) {
        double o[NV];
        for (int v = 0; v < NV; v++) {
          double acc = 0.0;
          for (int s = 0; s < nSub; s++) {
            const int n = s == 0 ? C : (s == 1 ? Cc : Cb);
            const double *wq = fw(s, f, q);
            double t = 0.0;
            for (int c = 0; c < n; c++) {
              const int cc = s < 2 ? c : st.sub[s - 2][c];
              t = std::fma(wq[c], Vf[cc][v], t);
            }
            acc = std::fma(om[s][v], t, acc);
          }
          o[v] = acc;
        }
        if (hydro) {
          o[4] = (vsq * o[0] + rho * u * o[1] + rho * vv * o[2] + rho * w * o[3] + o[4] / gamm);
          o[1] = (rho * o[1] + u * o[0]);
          o[2] = (rho * o[2] + vv * o[0]);
          o[3] = (rho * o[3] + w * o[0]);
        }
        if (mhd) lin2cons(o, gamm, rho, u, vv, w, bx, by, bz, rPsi, vsq);
        double *dst = own(f, i, j, k, q);
        for (int v = 0; v < NV; v++) dst[v] = o[v];
      }
    if (mhd) {  // fillInterior (Stencil_K.H:456-504)
      for (int p = 0; p < NI; p++) {
        double o[NV];
        for (int v = 0; v < NV; v++) {
          double acc = 0.0;
          for (int s = 0; s < nSub; s++) {
            const int n = s == 0 ? C : (s == 1 ? Cc : Cb);
            const double *wq = &cellW[(size_t(s) * NI + p) * C];
            double t = 0.0;
            for (int c = 0; c < n; c++) {
              const int cc = s < 2 ? c : st.sub[s - 2][c];
              t = std::fma(wq[c], Vf[cc][v], t);
            }
            acc = std::fma(om[s][v], t, acc);
          }
          o[v] = acc;
        }
        lin2cons(o, gamm, rho, u, vv, w, bx, by, bz, rPsi, vsq);
        double *dst = SP(i, j, k, p);
        for (int v = 0; v < NV; v++) dst[v] = o[v];
      }
    }
  }

  // linPrim2Cons of VarConvert<MHD_GLM>
  static void lin2cons(double *o, double gamm, double rho, double u, double vv, double w, double bx,
                       double by, double bz, double rPsi, double vsq) {
    o[4] = (vsq * o[0] + rho * u * o[1] + rho * vv * o[2] + rho * w * o[3] + bx * o[5] + by * o[6] +
            bz * o[7] + o[4] / gamm + rPsi * o[8]);
    o[1] = (rho * o[1] + u * o[0]);
    o[2] = (rho * o[2] + vv * o[0]);
    o[3] = (rho * o[3] + w * o[0]);
  }

  // ---- KXRCF selector (WenoSelect_K); returns number flagged
  int64_t weno_select(const std::vector<double> &V, double wThresh) {
    int64_t cnt = 0;
#pragma omp parallel for collapse(2) reduction(+ : cnt)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++) {
          double maxJump = 0.0;
          const int nQZ = D == 2 ? 1 : NQ;
          for (int nR = 0; nR < nQZ; nR++)
            for (int nQ = 0; nQ < NQ; nQ++) {
              const int q = nR * NQ + nQ;
              auto jump = [&](const double *in, const double *out) {
                if (!hydro && !mhd) return 0.0;
                return std::fabs(wenoP(out) - wenoP(in));
              };
              maxJump = std::fmax(maxJump, jump(FX(i, j, k, 1, q), FX(i, j, k, 0, q)));
              maxJump = std::fmax(maxJump, jump(FX(i + 1, j, k, 0, q), FX(i + 1, j, k, 1, q)));
              maxJump = std::fmax(maxJump, jump(FY(i, j, k, 1, q), FY(i, j, k, 0, q)));
              maxJump = std::fmax(maxJump, jump(FY(i, j + 1, k, 0, q), FY(i, j + 1, k, 1, q)));
              if (D == 3) {
                maxJump = std::fmax(maxJump, jump(FZ(i, j, k, 1, q), FZ(i, j, k, 0, q)));
                maxJump = std::fmax(maxJump, jump(FZ(i, j, k + 1, 0, q), FZ(i, j, k + 1, 1, q)));
              }
            }
          // weno() of LinAdv is 0: the indicator is 0/0 and never fires
          const double wc = (hydro || mhd) ? wenoP(&V[uc(i, j, k)]) : 0.0;
          const double ind = maxJump / wc;
          const size_t c = ci(i, j, k);
          if (ind >= wThresh || flag[2 * c + 1] < 1.0) {
            flag[2 * c] = 1.0;
            wflag[c] = 1;
            cnt++;
          } else {
            flag[2 * c] = 0.0;
            wflag[c] = 0;
          }
        }
    return cnt;
  }

  // ---- positivity-preserving limiter (PositivityPreserve_K)
  void pos_pres(const std::vector<double> &V) {
    if (!hydro && !mhd) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++) {
          const double *ref = &V[uc(i, j, k)];
          const double rDens = ref[0];
          double maxDens = -std::numeric_limits<double>::max();
          double minDens = std::numeric_limits<double>::max();
          double minEint = std::numeric_limits<double>::max();
          double minSdSp = std::numeric_limits<double>::max();
          const int zl = D == 3 ? -1 : 0, zh = D == 3 ? 1 : 0;
          for (int xo = -1; xo <= 1; xo++)
            for (int yo = -1; yo <= 1; yo++)
              for (int zo = zl; zo <= zh; zo++) {
                const double *s = D == 3 ? &V[ui(i + r + xo, j + r + yo, k + r + zo)]
                                         : &V[ui(i + r + xo, j + r + yo, 0)];
                maxDens = std::fmax(maxDens, s[0]);
                minDens = std::fmin(minDens, s[0]);
                const double ei = eintP(s);
                const double sd = soundP(s);
                minEint = std::fmin(minEint, ei);
                minSdSp = std::fmin(minSdSp, sd);
              }
          auto at = [&](int xo, int yo, int zo) -> const double * {
            return D == 3 ? &V[ui(i + r + xo, j + r + yo, k + r + zo)] : &V[ui(i + r + xo, j + r + yo, 0)];
          };
          const double dxU = at(1, 0, 0)[1] / at(1, 0, 0)[0] - at(-1, 0, 0)[1] / at(-1, 0, 0)[0];
          const double dyV = at(0, 1, 0)[2] / at(0, 1, 0)[0] - at(0, -1, 0)[2] / at(0, -1, 0)[0];
          const double dzW = D == 3 ? at(0, 0, 1)[3] / at(0, 0, 1)[0] - at(0, 0, -1)[3] / at(0, 0, -1)[0]
                                    : at(0, 0, 0)[3] / at(0, 0, 0)[0] - at(0, 0, 0)[3] / at(0, 0, 0)[0];
          const double divV = (dxU + dyV + dzW) / 2.0;
          double kappa = 0.2;
          const double eta =
              std::fmin(1.0, std::fmax(0.0, -(kappa * minSdSp + divV) / (kappa * minSdSp)));
          kappa *= (1.0 - eta);
          maxDens *= (1.0 + kappa);
          minDens *= (1.0 - kappa);
          minEint *= (1.0 - kappa);

          // density theta
          const double dP = maxDens - rDens, dM = rDens - minDens;
          double thD = 1.0;
          for (int q = 0; q < Q; q++)
            for (int f = 0; f < F; f++) {
              const double rho = own(f, i, j, k, q)[0];
              thD = rho > minDens ? thD : std::fmin(thD, dM / (rDens - rho));
              thD = rho < maxDens ? thD : std::fmin(thD, dP / (rho - rDens));
            }
          for (int p = 0; p < (mhd ? NI : 0); p++) {
            const double rho = SP(i, j, k, p)[0];
            thD = rho > minDens ? thD : std::fmin(thD, dM / (rDens - rho));
            thD = rho < maxDens ? thD : std::fmin(thD, dP / (rho - rDens));
          }
          if (thD > 0.0 && thD < 1.0) mix(i, j, k, ref, thD);
          else if (thD < 1.0) replace(i, j, k, ref);
          // internal energy theta
          double thP = 1.0;
          for (int q = 0; q < Q; q++)
            for (int f = 0; f < F; f++) {
              const double *s = own(f, i, j, k, q);
              const double ei = eintP(s);
              thP = ei > minEint ? thP : std::fmin(thP, eint_pos(minEint, ref, s));
            }
          for (int p = 0; p < (mhd ? NI : 0); p++) {
            const double *s = SP(i, j, k, p);
            const double ei = eintP(s);
            thP = ei > minEint ? thP : std::fmin(thP, eint_pos(minEint, ref, s));
          }
          if (thP > 0.0 && thP < 1.0) mix(i, j, k, ref, thP);
          else if (thP < 1.0) replace(i, j, k, ref);
          flag[2 * ci(i, j, k) + 1] = std::fmin(thD, thP);
        }
  }
  void mix(int64_t i, int64_t j, int64_t k, const double *ref, double th) {
    for (int q = 0; q < Q; q++)
      for (int f = 0; f < F; f++) {
        double *s = own(f, i, j, k, q);
        for (int v = 0; v < NV; v++) s[v] = ref[v] + th * (s[v] - ref[v]);
      }
    for (int p = 0; p < (mhd ? NI : 0); p++) {
      double *s = SP(i, j, k, p);
      for (int v = 0; v < NV; v++) s[v] = ref[v] + th * (s[v] - ref[v]);
    }
  }
  void replace(int64_t i, int64_t j, int64_t k, const double *ref) {
    for (int q = 0; q < Q; q++)
      for (int f = 0; f < F; f++) {
        double *s = own(f, i, j, k, q);
        for (int v = 0; v < NV; v++) s[v] = ref[v];
      }
    for (int p = 0; p < (mhd ? NI : 0); p++) {
      double *s = SP(i, j, k, p);
      for (int v = 0; v < NV; v++) s[v] = ref[v];
    }
  }
  static double eint_pos(double emin, const double *ref, const double *s) {
    double a = 0.0, b = 1.0, c;
    double st_[NV];
    for (int it = 0; it < 10; it++) {
      c = 0.5 * (a + b);
      for (int v = 0; v < NV; v++) st_[v] = ref[v] + c * (s[v] - ref[v]);
      const double dec = eintP(st_) - emin;
      if (dec > 0.0) a = c;
      else b = c;
    }
    return a;
  }

  // ---- Riemann solve at one face point; F written over both slots
  double riemann(int dir, double *sL_, double *sR_) {
    double uL[NV], uR[NV], Fl[NV];
    for (int v = 0; v < NV; v++) {
      uL[v] = sL_[v];
      uR[v] = sR_[v];
    }
    double s = 0.0;
    if (mhd) {
      double vL[9], vR[9];
      Mhd::cons2prim(gam(), uL, vL);
      Mhd::cons2prim(gam(), uR, vR);
      if (cfg.rs == 4) s = Mhd::solve_kepes(gam(), ch, dir, uL, uR, vL, vR, Fl);
      else s = Mhd::solve_hll(gam(), ch, cfg.rs == 0, dir, uL, uR, vL, vR, Fl);
    } else if (hydro) {
      double vL[5], vR[5];
      Hydro::cons2prim(gam(), uL, vL);
      Hydro::cons2prim(gam(), uR, vR);
      if (cfg.rs == 2) s = Hydro::solve_hllc(gam(), dir, uL, uR, vL, vR, Fl);
      else s = Hydro::solve_hll(gam(), cfg.rs == 0, dir, uL, uR, vL, vR, Fl);
    } else {
      if (!cfg.haveAdvVel && dir == 2) {
        Fl[0] = 0.0;
        s = 0.0;
      } else {
        const double a = vel[dir];
        Fl[0] = a > 0.0 ? a * uL[0] : a * uR[0];
        s = std::fabs(a);
      }
    }
    for (int v = 0; v < NV; v++) {
      sL_[v] = Fl[v];
      sR_[v] = Fl[v];
    }
    return s;
  }

  double riemann_all() {
    double vmax = 0.0;
#pragma omp parallel for collapse(2) reduction(max : vmax)
    for (int64_t i = 0; i <= nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int q = 0; q < Q; q++) {
            const double s = riemann(0, FX(i, j, k, 0, q), FX(i, j, k, 1, q));
            vmax = s > vmax ? s : vmax;
          }
#pragma omp parallel for collapse(2) reduction(max : vmax)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j <= ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int q = 0; q < Q; q++) {
            const double s = riemann(1, FY(i, j, k, 0, q), FY(i, j, k, 1, q));
            vmax = s > vmax ? s : vmax;
          }
    if (D == 3) {
#pragma omp parallel for collapse(2) reduction(max : vmax)
      for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
          for (int64_t k = 0; k <= nz; k++)
            for (int q = 0; q < Q; q++) {
              const double s = riemann(2, FZ(i, j, k, 0, q), FZ(i, j, k, 1, q));
              vmax = s > vmax ? s : vmax;
            }
    }
    return vmax;
  }

  // ---- flux divergence (IntegrateRHS_K)
  void integrate_rhs() {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++) {
          double *rhs = &RHS[ci(i, j, k) * NV];
          for (int v = 0; v < NV; v++) rhs[v] = 0.0;
          if (D == 2) {
            for (int q = 0; q < NQ; q++)
              for (int v = 0; v < NV; v++) {
                rhs[v] += wt[q] * (FX(i, j, k, 1, q)[v] - FX(i + 1, j, k, 0, q)[v]) / dx;
                rhs[v] += wt[q] * (FY(i, j, k, 1, q)[v] - FY(i, j + 1, k, 0, q)[v]) / dy;
              }
          } else {
            for (int v = 0; v < NV; v++)
              for (int a = 0; a < NQ; a++)
                for (int b = 0; b < NQ; b++) {
                  const int q = a * NQ + b;
                  const double qWt = wt[a] * wt[b];
                  rhs[v] += qWt * (FX(i, j, k, 0, q)[v] - FX(i + 1, j, k, 1, q)[v]) / dx;
                  rhs[v] += qWt * (FY(i, j, k, 0, q)[v] - FY(i, j + 1, k, 1, q)[v]) / dy;
                  rhs[v] += qWt * (FZ(i, j, k, 0, q)[v] - FZ(i, j, k + 1, 1, q)[v]) / dz;
                }
          }
          if (mhd) {  // source terms at the interior points
            for (int a = 0; a < NQ; a++)
              for (int b = 0; b < NQ; b++)
                for (int c = 0; c < NZ; c++) {
                  const double *S = SP(i, j, k, (a * NQ + b) * NZ + c);
                  const double w = D == 3 ? wt[a] * wt[b] * wt[c] : wt[a] * wt[b];
                  for (int v = 0; v < NV; v++) rhs[v] += w * S[v];
                }
          }
        }
  }

  // ---- GLM cleaning speed (SpeedEstimate_K, MHD_GLM_K.H:213-310; Solver.cpp:456-465)
  double mhd_speed() {
    double maxLam = 0.0, maxVel = 0.0;
#pragma omp parallel for collapse(2) reduction(max : maxLam, maxVel)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int f = 0; f < F; f++)
            for (int q = 0; q < Q; q++) {
              const double *s = own(f, i, j, k, q);
              const double vel = std::fabs(s[1] / s[0]);  // x-components on every face
              const double lam = Mhd::fast(gam(), s, s[5]) + vel;
              maxVel = vel > maxVel ? vel : maxVel;
              maxLam = lam > maxLam ? lam : maxLam;
            }
    return maxLam - maxVel;
  }

  // ---- GLM source terms (SourceTerms_K, MHD_GLM_K.H:316-547)
  void mhd_sources(const std::vector<double> &V) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++) {
          double divB = 0.0;
          for (int nQ = 0; nQ < NQ; nQ++)
            for (int nS = 0; nS < NZ; nS++) {
              const int q = nS * NQ + nQ;
              const double twt = D == 2 ? wt[nQ] : wt[nQ] * wt[nS];
              const double bAvW = 0.5 * (FX(i, j, k, 1, q)[5] + FX(i, j, k, 0, q)[5]);
              const double bAvE = 0.5 * (FX(i + 1, j, k, 0, q)[5] + FX(i + 1, j, k, 1, q)[5]);
              divB += twt * (bAvE - bAvW) / dx;
              const double bAvS = 0.5 * (FY(i, j, k, 1, q)[6] + FY(i, j, k, 0, q)[6]);
              const double bAvN = 0.5 * (FY(i, j + 1, k, 0, q)[6] + FY(i, j + 1, k, 1, q)[6]);
              divB += twt * (bAvN - bAvS) / dy;
              if (D == 3) {
                const double bAvB = 0.5 * (FZ(i, j, k, 1, q)[7] + FZ(i, j, k, 0, q)[7]);
                const double bAvT = 0.5 * (FZ(i, j, k + 1, 0, q)[7] + FZ(i, j, k + 1, 1, q)[7]);
                divB += twt * (bAvT - bAvB) / dz;
              }
            }
          auto at = [&](int xo, int yo, int zo) -> const double * {
            return D == 3 ? &V[ui(i + r + xo, j + r + yo, k + r + zo)] : &V[ui(i + r + xo, j + r + yo, 0)];
          };
          const bool fdiff = flag[2 * ci(i, j, k) + 1] < 0.75;
          auto dmat = [&](int a, int b) { return Dm[size_t(a) * (NQ + 2) + b]; };
          double gpsiDu[27];
          for (int nQ = 0; nQ < NQ; nQ++)
            for (int nR = 0; nR < NQ; nR++)
              for (int nS = 0; nS < NZ; nS++) {
                const int p = (nQ * NQ + nR) * NZ + nS;
                const double *cv = SP(i, j, k, p);
                const double rho = cv[0];
                const double vx = cv[1] / rho, vy = cv[2] / rho, vz = cv[3] / rho;
                double dpx, dpy, dpz = 0.0;
                if (fdiff) {
                  const double tdx = 1.0 / (2.0 * dx), tdy = 1.0 / (2.0 * dy);
                  dpx = tdx * (at(1, 0, 0)[8] - at(-1, 0, 0)[8]);
                  dpy = tdy * (at(0, 1, 0)[8] - at(0, -1, 0)[8]);
                  if (D == 3) {
                    const double tdz = 1.0 / (2.0 * dz);
                    dpz = tdz * (at(0, 0, 1)[8] - at(0, 0, -1)[8]);
                  }
                } else {
                  const int qx = D == 2 ? nR : nR * NQ + nS, qy = D == 2 ? nQ : nQ * NQ + nS;
                  dpx = dmat(nQ, 0) * own(0, i, j, k, qx)[8] + dmat(nQ, NQ + 1) * own(1, i, j, k, qx)[8];
                  for (int n = 1; n <= NQ; n++) dpx += dmat(nQ, n) * SP(i, j, k, ((n - 1) * NQ + nR) * NZ + nS)[8];
                  dpx /= dx;
                  dpy = dmat(nR, 0) * own(2, i, j, k, qy)[8] + dmat(nR, NQ + 1) * own(3, i, j, k, qy)[8];
                  for (int n = 1; n <= NQ; n++) dpy += dmat(nR, n) * SP(i, j, k, (nQ * NQ + n - 1) * NZ + nS)[8];
                  dpy /= dy;
                  if (D == 3) {
                    const int qz = nQ * NQ + nR;
                    dpz = dmat(nS, 0) * own(4, i, j, k, qz)[8] + dmat(nS, NQ + 1) * own(5, i, j, k, qz)[8];
                    for (int n = 1; n <= NQ; n++) dpz += dmat(nS, n) * SP(i, j, k, (nQ * NQ + nR) * NZ + n - 1)[8];
                    dpz /= dz;
                  }
                }
                gpsiDu[p] = vx * dpx + vy * dpy + vz * dpz;
              }
          for (int p = 0; p < NI; p++) {
            double *S = SP(i, j, k, p);
            double stt[9];
            for (int v = 0; v < 9; v++) stt[v] = S[v];
            const double vx = stt[1] / stt[0], vy = stt[2] / stt[0], vz = stt[3] / stt[0];
            const double uDb = vx * stt[5] + vy * stt[6] + vz * stt[7];
            const double f0 = 0.0;
            S[0] = f0;
            S[1] = f0 - divB * stt[5];
            S[2] = f0 - divB * stt[6];
            S[3] = f0 - divB * stt[7];
            const double eForce = divB * uDb + gpsiDu[p] * stt[8];
            S[4] = f0 - eForce;
            S[5] = f0 - divB * vx;
            S[6] = f0 - divB * vy;
            S[7] = f0 - divB * vz;
            S[8] = -gpsiDu[p];
          }
        }
  }

  // ---- Solver::evalRHS
  double eval_rhs_on(std::vector<double> &V) {
    nRhsEval++;
    cell_bcs(V);
    if (useSparseWeno) {
      const double wThresh = cfg.wenoThresh * std::pow(dmin, 1.5);
      recon_linear(V);
      face_bcs();
      nWeno = weno_select(V, wThresh);
      if (nWeno > 0) {
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
        for (int64_t i = 0; i < nx; i++)
          for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
              if (wflag[ci(i, j, k)]) recon_weno_cell(V, i, j, k);
      }
    } else {
#pragma omp parallel for collapse(2) schedule(static)
      for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
          for (int64_t k = 0; k < nz; k++) recon_weno_cell(V, i, j, k);
      useSparseWeno = cfg.wenoThresh > 0.0;
    }
    pos_pres(V);
    face_bcs();
    if (mhd) {
      ch = mhd_speed();
      mhd_sources(V);
    }
    const double vmax = riemann_all();
    integrate_rhs();
    return vmax;
  }
  double eval_rhs(int which) override { return eval_rhs_on(which == 0 ? U : Uprev); }

  // Own face states of every cell from U: linear reconstruction (which = 0) or
  // full WENO followed by the positivity limiter (which = 1); layout
  // out[((f*Q + q)*NV + v)*ncell + cell].  The limiter's theta is not kept.
  void face_states(int which, double *out) override {
    cell_bcs(U);
    if (which == 0) {
      recon_linear(U);
    } else if (which == 2) {
      const std::vector<double> keep = flag;
      recon_linear(U);
      pos_pres(U);
      flag = keep;
    } else {
      const std::vector<double> keep = flag;
#pragma omp parallel for collapse(2) schedule(static)
      for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
          for (int64_t k = 0; k < nz; k++) recon_weno_cell(U, i, j, k);
      pos_pres(U);
      flag = keep;
    }
    const size_t ncl = size_t(nx) * ny * nz;
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int f = 0; f < F; f++)
            for (int q = 0; q < Q; q++)
              for (int v = 0; v < NV; v++)
                out[(size_t(f * Q + q) * NV + v) * ncl + (size_t(k) * ny + j) * nx + i] = own(f, i, j, k, q)[v];
  }

  // Nonlinear WENO-AO weights of cell (i,j,k) of U, [nSub][NV]
  void weno_omegas(int64_t i, int64_t j, int64_t k, double *out) override {
    cell_bcs(U);
    recon_weno_cell(U, i, j, k, out);
  }

  // ---- state access
  void set_state(const double *S) override {
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int v = 0; v < NV; v++) {
            const double x = S[((size_t(v) * nz + k) * ny + j) * nx + i];
            U[uc(i, j, k) + v] = x;
            Uprev[uc(i, j, k) + v] = x;
          }
  }
  void get_state(double *S) const override {
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int v = 0; v < NV; v++) S[((size_t(v) * nz + k) * ny + j) * nx + i] = U[uc(i, j, k) + v];
  }
  void get_rhs(double *S) const override {
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int v = 0; v < NV; v++)
            S[((size_t(v) * nz + k) * ny + j) * nx + i] = RHS[ci(i, j, k) * NV + v];
  }
  void get_flags(double *S) const override {
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int s = 0; s < 2; s++) S[((size_t(s) * nz + k) * ny + j) * nx + i] = flag[2 * ci(i, j, k) + s];
  }
  void set_flags(const double *S) override {
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++)
          for (int s = 0; s < 2; s++) flag[2 * ci(i, j, k) + s] = S[((size_t(s) * nz + k) * ny + j) * nx + i];
  }
  int set_ic(const char *name, const double *user, uint64_t seed) override {
    const int nQZ = D == 3 ? NQ : 1;
    const double mid[3] = {0.5 * (cfg.lo[0] + cfg.hi[0]), 0.5 * (cfg.lo[1] + cfg.hi[1]),
                           0.5 * (cfg.lo[2] + cfg.hi[2])};
    int ok = 1;
    for (int64_t i = 0; i < nx; i++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++) {
          const double xc = dx * (double(i) + 0.5) + cfg.lo[0];
          const double yc = dy * (double(j) + 0.5) + cfg.lo[1];
          const double zc = D == 3 ? dz * (double(k) + 0.5) + cfg.lo[2] : 0.0;
          const uint64_t cell = (uint64_t(k) * ny + j) * nx + i;
          double acc[NV];
          for (int v = 0; v < NV; v++) acc[v] = 0.0;
          for (int a = 0; a < NQ; a++)
            for (int b = 0; b < NQ; b++)
              for (int c = 0; c < nQZ; c++) {
                const double px = ab[a] * dx + xc, py = ab[b] * dy + yc;
                const double pz = D == 3 ? ab[c] * dz + zc : cfg.lo[2];
                double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (!point_ic(name, D, gam(), user, seed, cell, mid, px, py, pz, s)) ok = 0;
                const double w = D == 3 ? wt[a] * wt[b] * wt[c] : wt[a] * wt[b];
                for (int v = 0; v < NV; v++) acc[v] += w * s[v];
              }
          for (int v = 0; v < NV; v++) {
            U[uc(i, j, k) + v] = acc[v];
            Uprev[uc(i, j, k) + v] = acc[v];
          }
        }
    return ok ? 0 : -1;
  }

  // ---- Solver::Solver tail (FSAL prime)
  void prime() override {
    const double maxVel = eval_rhs_on(U);
    dt = 0.01 * cfg.cfl * dmin / maxVel;
  }

  // ---- Solver::TakeStep
  int take_step() override {
    const bool fixed = cfg.fixedDt > 0.0;
    if (fixed) dt = cfg.fixedDt;
    if (time + dt > cfg.finalTime) {
      dt = cfg.finalTime - time;
      lastTimeStep = true;
    }
    bool accepted = false, firstUnphys = true;
    int status = 0;
    const size_t nc = size_t(nx) * ny * nz;
    for (int nT = 0; nT < cfg.rejectionLimit; nT++) {
      double maxVel = 0.0, v;
      if (nT > 0) {
        v = eval_rhs_on(Uprev);
        maxVel = std::fmax(maxVel, v);
      }
      {
        const double betaDt = rk.beta[0] * dt, bhatDt = rk.bhat[0] * dt;
#pragma omp parallel for collapse(2)
        for (int64_t i = 0; i < nx; i++)
          for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
              for (int vv = 0; vv < NV; vv++) {
                const size_t h = uc(i, j, k) + vv, c = ci(i, j, k) * NV + vv;
                U[h] = Uprev[h] + betaDt * RHS[c];
                Uhat[c] = Uprev[h] + bhatDt * RHS[c];
                Utmp[c] = Uprev[h];
              }
      }
      for (int s = 1; s < rk.nStages; s++) {
        const double delta = rk.delta[s], g1 = rk.gam1[s], g2 = rk.gam2[s], g3 = rk.gam3[s];
        const double cDt = rk.c[s] * dt, bhatDt = rk.bhat[s] * dt, betaDt = rk.beta[s] * dt;
        (void)cDt;
        v = eval_rhs_on(U);
        maxVel = std::fmax(maxVel, v);
#pragma omp parallel for collapse(2)
        for (int64_t i = 0; i < nx; i++)
          for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
              for (int vv = 0; vv < NV; vv++) {
                const size_t h = uc(i, j, k) + vv, c = ci(i, j, k) * NV + vv;
                Utmp[c] += delta * U[h];
                U[h] = g1 * U[h] + g2 * Utmp[c] + g3 * Uprev[h] + betaDt * RHS[c];
                Uhat[c] += bhatDt * RHS[c];
              }
      }
      v = eval_rhs_on(U);
      maxVel = std::fmax(maxVel, v);
      {
        const double bhatDt = rk.bhatfsal * dt;
#pragma omp parallel for
        for (size_t c = 0; c < nc * NV; c++) Uhat[c] += bhatDt * RHS[c];
      }
      const double cflUsed = dt * maxVel / dmin;
      (void)cflUsed;
      double errSum = 0.0, posFlag = 1.0;
      for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
          for (int64_t k = 0; k < nz; k++) {
            double u[NV], uh[NV];
            for (int vv = 0; vv < NV; vv++) {
              u[vv] = U[uc(i, j, k) + vv];
              uh[vv] = Uhat[ci(i, j, k) * NV + vv];
            }
            err_est_cell(NV, hydro ? 0 : (mhd ? 2 : 1), u, uh, cfg.atol, cfg.rtol, errSum, posFlag);
          }
      const double nDofs = double(nc) * NV;
      double errNew = 1.0 / std::sqrt(errSum / nDofs);
      double dtfac = std::pow(errNew, rk.ep1) * std::pow(errEst, rk.ep2);
      dtfac = std::isnan(dtfac) ? 1.2 : 1.0 + std::atan(dtfac - 1.0);
      int lStat;
      if (posFlag < 0.0) lStat = 2;
      else if (dtfac >= cfg.rejectionThresh || fixed) lStat = 0;
      else lStat = 1;
      status = lStat;
      if (lStat == 2) {
        dt = firstUnphys ? std::fmin(0.25 * cfg.cfl * dmin / maxVel, dt / 4.0) : dt / 2.0;
        firstUnphys = false;
        useSparseWeno = false;
        if (lastTimeStep) lastTimeStep = false;
      } else if (lStat == 0) {
        accepted = true;
        time += dt;
        const double dterr = dt * dtfac, dtcfl = cfg.cfl * dmin / maxVel;
        dt = fixed ? cfg.fixedDt : std::fmin(dterr, dtcfl);
        errEst = errNew;
#pragma omp parallel for collapse(2)
        for (int64_t i = 0; i < nx; i++)
          for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
              for (int vv = 0; vv < NV; vv++) Uprev[uc(i, j, k) + vv] = U[uc(i, j, k) + vv];
        break;
      } else {
        dt = dt * dtfac;
        if (lastTimeStep) lastTimeStep = false;
      }
    }
    if (!accepted) {
      lastTimeStep = true;
      status = 3;
    }
    if (dt < cfg.minDt) {
      lastTimeStep = true;
      if (accepted) status = 4;
    }
    nTS++;
    return status;
  }
};

}  // namespace orc

struct orc_solver {
  orc::Base *impl;
};

extern "C" {

orc_solver *orc_create(const orc_config *cfg, const double *faceW, const double *derivW) {
#ifdef _OPENMP
  if (cfg->nThreads > 0) omp_set_num_threads(cfg->nThreads);
#endif
  orc::Base *b = nullptr;
  if (cfg->eq == 0) {
    if (cfg->dim == 2) b = new orc::Solver<2, 5>(*cfg, faceW, derivW);
    else b = new orc::Solver<3, 5>(*cfg, faceW, derivW);
  } else if (cfg->eq == 1) {
    if (cfg->dim == 2) b = new orc::Solver<2, 1>(*cfg, faceW, derivW);
    else b = new orc::Solver<3, 1>(*cfg, faceW, derivW);
  } else if (cfg->eq == 2) {
    if (cfg->dim == 2) b = new orc::Solver<2, 9>(*cfg, faceW, derivW);
    else b = new orc::Solver<3, 9>(*cfg, faceW, derivW);
  } else {
    return nullptr;
  }
  return new orc_solver{b};
}
void orc_destroy(orc_solver *s) {
  if (!s) return;
  delete s->impl;
  delete s;
}
void orc_set_state(orc_solver *s, const double *U) { s->impl->set_state(U); }
void orc_get_state(const orc_solver *s, double *U) { s->impl->get_state(U); }
int orc_set_ic(orc_solver *s, const char *p, const double *user, uint64_t seed) {
  return s->impl->set_ic(p, user, seed);
}
double orc_eval_rhs(orc_solver *s, int which) { return s->impl->eval_rhs(which); }
void orc_get_rhs(const orc_solver *s, double *r) { s->impl->get_rhs(r); }
void orc_get_flags(const orc_solver *s, double *f) { s->impl->get_flags(f); }
void orc_set_flags(orc_solver *s, const double *f) { s->impl->set_flags(f); }
void orc_prime(orc_solver *s) { s->impl->prime(); }
int orc_take_step(orc_solver *s) { return s->impl->take_step(); }
void orc_get_time(const orc_solver *s, double *t, double *dt, int64_t *step, double *errEst) {
  *t = s->impl->time;
  *dt = s->impl->dt;
  *step = s->impl->nTS;
  *errEst = s->impl->errEst;
}
void orc_set_time(orc_solver *s, double t, double dt) {
  s->impl->time = t;
  s->impl->dt = dt;
}
int64_t orc_rhs_evals(const orc_solver *s) { return s->impl->nRhsEval; }
void orc_force_full_weno(orc_solver *s) { s->impl->useSparseWeno = false; }
int64_t orc_num_weno(const orc_solver *s) { return s->impl->nWeno; }
int orc_last_step(const orc_solver *s) { return s->impl->lastTimeStep ? 1 : 0; }
void orc_face_states(orc_solver *s, int which, double *out) { s->impl->face_states(which, out); }
void orc_weno_omegas(orc_solver *s, int64_t i, int64_t j, int64_t k, double *out) {
  s->impl->weno_omegas(i, j, k, out);
}
void orc_error_norm(int nv, int eq, int64_t ncell, const double *U, const double *Uhat, double atol, double rtol,
                    double *errSum, double *posFlag) {
  double e = 0.0, p = 1.0;
  for (int64_t c = 0; c < ncell; c++) orc::err_est_cell(nv, eq, U + c * nv, Uhat + c * nv, atol, rtol, e, p);
  *errSum = e;
  *posFlag = p;
}
double orc_riemann_point(int eq, int rs, double gamma, double ch, int dir, const double *uL,
                         const double *uR, double *F) {
  if (eq == 2) {
    double vL[9], vR[9];
    orc::Mhd::cons2prim(gamma, uL, vL);
    orc::Mhd::cons2prim(gamma, uR, vR);
    if (rs == 4) return orc::Mhd::solve_kepes(gamma, ch, dir, uL, uR, vL, vR, F);
    return orc::Mhd::solve_hll(gamma, ch, rs == 0, dir, uL, uR, vL, vR, F);
  }
  double vL[5], vR[5];
  orc::Hydro::cons2prim(gamma, uL, vL);
  orc::Hydro::cons2prim(gamma, uR, vR);
  if (rs == 2) return orc::Hydro::solve_hllc(gamma, dir, uL, uR, vL, vR, F);
  return orc::Hydro::solve_hll(gamma, rs == 0, dir, uL, uR, vL, vR, F);
}
void orc_set_mhd_tables(orc_solver *s, const double *cellW, const double *Dm) {
  s->impl->set_mhd_tables(cellW, Dm);
}

}  // extern "C"
