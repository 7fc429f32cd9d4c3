// Training set and trainer of the learned troubled-cell indicator.  Samples
// (features and labels) are collected on the device from solver states; the
// network is trained there by the library's own forward / backward / Adam
// kernels (csrc/device/indicator.hpp).  No counterpart in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace kfvm {

class Solver;

class Indicator {
 public:
  explicit Indicator(int dim);
  ~Indicator();
  Indicator(const Indicator &) = delete;
  Indicator &operator=(const Indicator &) = delete;

  int dim() const { return dim_; }
  int inputs() const { return nin_; }
  int params() const { return np_; }
  int64_t samples() const { return n_; }
  int64_t positives() const { return npos_; }

  // Append the samples of the solver's current state (labels with threshold tau).
  void collect(Solver &s, double tau);
  // Copy samples [off, off+n) to the host.
  void getSamples(float *X, float *Y, int64_t off, int64_t n);
  // Deterministic He-uniform initialisation of the FP32 master parameters.
  void init(uint64_t seed);
  // iters Adam steps on batches of `batch` samples drawn by a counter hash of
  // (seed, step); class weight posWeight on positive samples.  Returns the
  // mean loss of the last batch.
  double train(int iters, int batch, double lr, double posWeight, uint64_t seed);
  void getParams(float *P) const;   // FP32 master parameters
  void setParams(const float *P);
  void getBf16(uint16_t *W) const;  // bfloat16 parameters the forward pass uses
  // Mean loss and gradient of n explicit samples at parameters P (the forward
  // pass uses P as given).
  double gradient(const float *P, const float *X, const float *Y, int64_t n, double posWeight, float *grad);
  // Logits of n feature rows with parameters P.
  void logits(const float *P, const float *X, int64_t n, float *Z);

 private:
  void reserve(int64_t n);
  void gradOn(const float *dP, const float *dX, const float *dY, const int *dIdx, int64_t n, int nb,
              uint64_t seed, double posWeight);

  int dim_, nin_, np_;
  int64_t n_ = 0, cap_ = 0, npos_ = 0;
  float *X_ = nullptr, *Y_ = nullptr;
  float *M_ = nullptr, *Pbf_ = nullptr, *grad_ = nullptr, *m1_ = nullptr, *m2_ = nullptr, *loss_ = nullptr;
  float *part_ = nullptr, *lossPart_ = nullptr;
  int partBlocks_ = 0;
  int64_t step_ = 0;
  hipStream_t stream_ = nullptr;
};

}  // namespace kfvm
