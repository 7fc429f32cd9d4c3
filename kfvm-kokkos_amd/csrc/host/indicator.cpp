// Training set and trainer of the learned troubled-cell indicator.  See
// indicator.hpp.
#include "indicator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../device/device_api.hpp"
#include "solver.hpp"

namespace kfvm {

namespace {
void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
template <class T>
T *dalloc(size_t n, const char *what) {
  void *p = nullptr;
  check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), what);
  return static_cast<T *>(p);
}
uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
float bf16_round_host(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&v, &u, 4);
  return v;
}
}  // namespace

Indicator::Indicator(int dim) : dim_(dim) {
  if (dim != 2 && dim != 3) throw std::runtime_error("indicator dimension must be 2 or 3");
  nin_ = net_inputs(dim);
  np_ = net_params(dim);
  check(hipStreamCreate(&stream_), "hipStreamCreate");
  M_ = dalloc<float>(np_, "indicator master");
  Pbf_ = dalloc<float>(np_, "indicator bf16");
  grad_ = dalloc<float>(np_, "indicator grad");
  m1_ = dalloc<float>(np_, "adam m");
  m2_ = dalloc<float>(np_, "adam v");
  loss_ = dalloc<float>(1, "loss");
  init(1);
}

Indicator::~Indicator() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  void *bufs[] = {X_, Y_, M_, Pbf_, grad_, m1_, m2_, loss_, part_, lossPart_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Indicator::reserve(int64_t n) {
  if (n <= cap_) return;
  const int64_t cap = std::max<int64_t>(n, cap_ * 2);
  float *X = dalloc<float>(size_t(cap) * nin_, "indicator features");
  float *Y = dalloc<float>(size_t(cap), "indicator labels");
  if (n_ > 0) {
    check(hipMemcpyAsync(X, X_, size_t(n_) * nin_ * sizeof(float), hipMemcpyDeviceToDevice, stream_), "grow");
    check(hipMemcpyAsync(Y, Y_, size_t(n_) * sizeof(float), hipMemcpyDeviceToDevice, stream_), "grow");
  }
  check(hipStreamSynchronize(stream_), "sync");
  if (X_) (void)hipFree(X_);
  if (Y_) (void)hipFree(Y_);
  X_ = X;
  Y_ = Y;
  cap_ = cap;
}

void Indicator::collect(Solver &s, double tau) {
  if (s.config().dim != dim_) throw std::runtime_error("solver dimension differs from the indicator's");
  const int64_t nc = s.ncell();
  reserve(n_ + nc);
  s.collectSamples(X_ + size_t(n_) * nin_, Y_ + n_, tau);
  std::vector<float> y(static_cast<size_t>(nc));
  check(hipMemcpy(y.data(), Y_ + n_, size_t(nc) * sizeof(float), hipMemcpyDeviceToHost), "read labels");
  for (float v : y) npos_ += v > 0.5f ? 1 : 0;
  n_ += nc;
}

void Indicator::getSamples(float *X, float *Y, int64_t off, int64_t n) {
  if (off < 0 || n < 0 || off + n > n_) throw std::runtime_error("sample range out of bounds");
  check(hipStreamSynchronize(stream_), "sync");
  if (X) check(hipMemcpy(X, X_ + size_t(off) * nin_, size_t(n) * nin_ * sizeof(float), hipMemcpyDeviceToHost), "X");
  if (Y) check(hipMemcpy(Y, Y_ + off, size_t(n) * sizeof(float), hipMemcpyDeviceToHost), "Y");
}

void Indicator::init(uint64_t seed) {
  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases
  const int H = 32;
  std::vector<float> P(size_t(np_), 0.0f);
  const int fan[3] = {nin_, H, H};
  const int off[3] = {0, H * nin_ + H, H * nin_ + H + H * H + H};
  const int cnt[3] = {H * nin_, H * H, H};
  uint64_t ctr = 0;
  for (int l = 0; l < 3; l++) {
    const double b = std::sqrt(6.0 / fan[l]);
    for (int i = 0; i < cnt[l]; i++) {
      const double u = double(splitmix(seed * 0x9E3779B97F4A7C15ull + ctr++) >> 11) / 9007199254740992.0;
      P[size_t(off[l] + i)] = float((2.0 * u - 1.0) * b);
    }
  }
  setParams(P.data());
}

void Indicator::setParams(const float *P) {
  std::vector<float> bf(static_cast<size_t>(np_));
  for (int i = 0; i < np_; i++) bf[size_t(i)] = bf16_round_host(P[i]);
  check(hipMemcpy(M_, P, size_t(np_) * sizeof(float), hipMemcpyHostToDevice), "params");
  check(hipMemcpy(Pbf_, bf.data(), size_t(np_) * sizeof(float), hipMemcpyHostToDevice), "params");
  check(hipMemset(m1_, 0, size_t(np_) * sizeof(float)), "adam");
  check(hipMemset(m2_, 0, size_t(np_) * sizeof(float)), "adam");
  step_ = 0;
}

void Indicator::getParams(float *P) const {
  check(hipStreamSynchronize(stream_), "sync");
  check(hipMemcpy(P, M_, size_t(np_) * sizeof(float), hipMemcpyDeviceToHost), "params");
}

void Indicator::getBf16(uint16_t *W) const {
  check(hipStreamSynchronize(stream_), "sync");
  std::vector<float> bf(static_cast<size_t>(np_));
  check(hipMemcpy(bf.data(), Pbf_, size_t(np_) * sizeof(float), hipMemcpyDeviceToHost), "params");
  for (int i = 0; i < np_; i++) {
    uint32_t u;
    std::memcpy(&u, &bf[size_t(i)], 4);
    W[i] = uint16_t(u >> 16);
  }
}

void Indicator::gradOn(const float *dP, const float *dX, const float *dY, const int *dIdx, int64_t n, int nb,
                       uint64_t seed, double posWeight) {
  const int tb = net_train_block();
  const int nblk = (nb + tb - 1) / tb;
  if (nblk > partBlocks_) {
    check(hipStreamSynchronize(stream_), "sync");
    if (part_) (void)hipFree(part_);
    if (lossPart_) (void)hipFree(lossPart_);
    part_ = dalloc<float>(size_t(nblk) * np_, "gradient partials");
    lossPart_ = dalloc<float>(size_t(nblk), "loss partials");
    partBlocks_ = nblk;
  }
  check(launch_train_grad(dim_, dX, dY, dIdx, int(n), nb, seed, dP, float(posWeight), part_, lossPart_, stream_),
        "train_grad");
  check(launch_train_reduce(part_, lossPart_, nblk, np_, nb, grad_, loss_, stream_), "train_reduce");
}

double Indicator::train(int iters, int batch, double lr, double posWeight, uint64_t seed) {
  if (n_ == 0) throw std::runtime_error("no samples collected");
  if (iters < 0 || batch <= 0) throw std::runtime_error("iters >= 0 and batch > 0 required");
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < iters; it++) {
    step_++;
    gradOn(Pbf_, X_, Y_, nullptr, n_, batch, splitmix(seed + uint64_t(step_)), posWeight);
    const float c1 = float(1.0 - std::pow(double(b1), double(step_)));
    const float c2 = float(1.0 - std::pow(double(b2), double(step_)));
    check(launch_adam(M_, Pbf_, grad_, m1_, m2_, np_, float(lr), b1, b2, eps, c1, c2, stream_), "adam");
  }
  float l = 0.0f;
  check(hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, stream_), "loss");
  check(hipStreamSynchronize(stream_), "sync");
  return double(l);
}

double Indicator::gradient(const float *P, const float *X, const float *Y, int64_t n, double posWeight,
                           float *grad) {
  float *dP = dalloc<float>(size_t(np_), "P");
  float *dX = dalloc<float>(size_t(n) * nin_, "X");
  float *dY = dalloc<float>(size_t(n), "Y");
  int *dIdx = dalloc<int>(size_t(n), "idx");
  std::vector<int> idx(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; i++) idx[size_t(i)] = int(i);
  double l = 0.0;
  try {
    check(hipMemcpy(dP, P, size_t(np_) * sizeof(float), hipMemcpyHostToDevice), "P");
    check(hipMemcpy(dX, X, size_t(n) * nin_ * sizeof(float), hipMemcpyHostToDevice), "X");
    check(hipMemcpy(dY, Y, size_t(n) * sizeof(float), hipMemcpyHostToDevice), "Y");
    check(hipMemcpy(dIdx, idx.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice), "idx");
    gradOn(dP, dX, dY, dIdx, n, int(n), 0, posWeight);
    float lf = 0.0f;
    check(hipMemcpyAsync(grad, grad_, size_t(np_) * sizeof(float), hipMemcpyDeviceToHost, stream_), "grad");
    check(hipMemcpyAsync(&lf, loss_, sizeof(float), hipMemcpyDeviceToHost, stream_), "loss");
    check(hipStreamSynchronize(stream_), "sync");
    l = lf;
  } catch (...) {
    (void)hipFree(dP), (void)hipFree(dX), (void)hipFree(dY), (void)hipFree(dIdx);
    throw;
  }
  (void)hipFree(dP), (void)hipFree(dX), (void)hipFree(dY), (void)hipFree(dIdx);
  return l;
}

void Indicator::logits(const float *P, const float *X, int64_t n, float *Z) {
  float *dP = dalloc<float>(size_t(np_), "P");
  float *dX = dalloc<float>(size_t(n) * nin_, "X");
  float *dZ = dalloc<float>(size_t(n), "Z");
  try {
    check(hipMemcpy(dP, P, size_t(np_) * sizeof(float), hipMemcpyHostToDevice), "P");
    check(hipMemcpy(dX, X, size_t(n) * nin_ * sizeof(float), hipMemcpyHostToDevice), "X");
    check(launch_logits(dim_, dX, dP, dZ, int(n), stream_), "logits");
    check(hipStreamSynchronize(stream_), "sync");
    check(hipMemcpy(Z, dZ, size_t(n) * sizeof(float), hipMemcpyDeviceToHost), "Z");
  } catch (...) {
    (void)hipFree(dP), (void)hipFree(dX), (void)hipFree(dZ);
    throw;
  }
  (void)hipFree(dP), (void)hipFree(dX), (void)hipFree(dZ);
}

}  // namespace kfvm
