// Check of the Riemann solvers' reciprocal, quotient and square root
// (rcp_n / div_n / sqrt_n in device/physics.hpp) against the IEEE operations:
// 2^26 operand pairs, log-uniform magnitudes in [2^-40, 2^40) with both
// signs (square roots: magnitudes only), plus zero and a negative square-root
// operand.  Prints the number of results that differ from the IEEE result
// and the largest difference in units in the last place.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../device/physics.hpp"

using kfvm::dev::div_n;
using kfvm::dev::rcp_n;
using kfvm::dev::sqrt_n;

__device__ inline uint64_t mix(uint64_t z) {  // splitmix64 finaliser
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline double operand(uint64_t h, bool sign) {
  const double u = double(h >> 11) * 0x1.0p-53;
  const double x = exp2(-40.0 + 80.0 * u);
  return sign && (h & 1u) ? -x : x;
}
__device__ inline unsigned long long ulps(double a, double b) {
  const long long d = __double_as_longlong(a) - __double_as_longlong(b);
  return (unsigned long long)(d < 0 ? -d : d);
}

// out: [0..2] mismatches of rcp, div, sqrt; [3..5] largest ulp differences;
// [6] sqrt_n(0) == 0; [7] sqrt_n(-1) is NaN
__global__ void check(unsigned long long *out, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t h1 = mix(2 * uint64_t(t)), h2 = mix(2 * uint64_t(t) + 1);
  const double x = operand(h1, true), y = operand(h2, true), p = operand(h1, false);
  const unsigned long long e[3] = {ulps(rcp_n(x), 1.0 / x), ulps(div_n(y, x), y / x), ulps(sqrt_n(p), sqrt(p))};
  for (int k = 0; k < 3; k++) {
    if (e[k]) atomicAdd(out + k, 1ull);
    atomicMax(out + 3 + k, e[k]);
  }
  if (t == 0) {
    out[6] = sqrt_n(0.0) == 0.0;
    const double q = sqrt_n(-1.0);
    out[7] = q != q;
  }
}

int main() {
  const int n = 1 << 26, threads = 256;
  unsigned long long *d = nullptr, h[8];
  if (hipMalloc(&d, sizeof(h)) != hipSuccess || hipMemset(d, 0, sizeof(h)) != hipSuccess) return 2;
  hipLaunchKernelGGL(check, dim3(n / threads), dim3(threads), 0, 0, d, n);
  if (hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 2;
  hipFree(d);
  std::printf("operands %d\nrcp_n mismatches %llu max_ulp %llu\ndiv_n mismatches %llu max_ulp %llu\n"
              "sqrt_n mismatches %llu max_ulp %llu\nsqrt_n(0) zero %llu\nsqrt_n(-1) nan %llu\n",
              n, h[0], h[3], h[1], h[4], h[2], h[5], h[6], h[7]);
  return 0;
}
