// Gaussian-process (squared-exponential kernel) reconstruction weights for the
// radius-2 KFVM stencils.  Host-side set-up; the stage kernels only read the
// resulting tables.
//
//   layout   faceWeights[nSub][2D][Q][C], derivWeights[nSub][I][C],
//            cellWeights[nSub][NQ^D][C]  (src/stencil/Stencil.H:27-49)
//   method   src/stencil/Stencil.cpp:281-846 with ReconVectors.cpp:14-86,
//            SEKernel.H:14-28, Monomials.cpp:14-57, StencilSymmetry.cpp:208-283
#pragma once

#include <vector>

namespace kfvm {

struct WeightTable {
  int dim = 3, nq = 2;
  int nSub = 0, F = 0, Q = 0, C = 0, I = 0, nqCell = 0;
  std::vector<double> face;   // [nSub][F][Q][C]
  std::vector<double> deriv;  // [nSub][I][C]
  std::vector<double> cell;   // [nSub][nqCell][C], quadrature points in nested (x,y[,z]) order
  // Number of reconstruction vectors obtained by a linear solve (the rest are
  // exact permutations under stencil symmetries).
  int nSolved = 0;

  double &f(int s, int face_, int q, int c) { return face[((s * F + face_) * Q + q) * C + c]; }
  double &d(int s, int i, int c) { return deriv[(s * I + i) * C + c]; }
  double &cw(int s, int q, int c) { return cell[(s * nqCell + q) * C + c]; }
};

// lfac: GP length scale in cells (eps = 1/(lfac*sqrt(2))).  dx,dy,dz scale the
// smoothness-indicator derivative weights by dx^a dy^b dz^c.
WeightTable build_weights(int dim, int nq, double lfac, double dx, double dy, double dz);

}  // namespace kfvm
