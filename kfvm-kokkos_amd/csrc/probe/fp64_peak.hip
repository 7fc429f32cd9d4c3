// Throughput of FP64 vector FMA (v_fma_f64) against FP64 matrix FMA
// (v_mfma_f64_16x16x4_f64) on the current GPU, all operands in registers:
// the measurement behind the choice of vector FMA for the reconstruction
// contractions.  Run it under rocprofv3 --kernel-trace --stats: the two
// kernels do known FP64 work (printed), so their trace durations give the
// rates.
//
//   valu_fma  eight independent FMA chains per lane
//   mfma_f64  four independent 16x16x4 accumulations per wave
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef double v4d __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));            \
      std::exit(1);                                                           \
    }                                                                         \
  } while (0)

__global__ __launch_bounds__(256) void valu_fma(double *out, double a, double b, int iters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double x[8];
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = t * 1e-9 + k;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = fma(x[k], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; k++) s += x[k];
  out[t] = s;
}

__global__ __launch_bounds__(256) void mfma_f64(double *out, double a0, int iters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const double a = a0 + (t & 63) * 1e-12, b = 1.0 - (t & 63) * 1e-13;
  v4d c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, c2 = c0, c3 = c0;
  for (int i = 0; i < iters; i++) {
    c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c3, 0, 0, 0);
  }
  out[t] = c0[0] + c1[1] + c2[2] + c3[3];
}

int main() {
  hipDeviceProp_t p;
  CHECK(hipGetDeviceProperties(&p, 0));
  const int nblk = p.multiProcessorCount * 8, nthr = 256, n = nblk * nthr;
  double *out = nullptr;
  CHECK(hipMalloc(&out, size_t(n) * sizeof(double)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const int iters = 20000;
  const double flVALU = 2.0 * 8 * double(iters) * n;                 // 2 flops per FMA
  const double flMFMA = 2.0 * 16 * 16 * 4 * 4 * double(iters) * (n / 64);  // per wave
  std::printf("%d CUs; valu_fma %.4g FP64 flops per launch, mfma_f64 %.4g\n", p.multiProcessorCount, flVALU,
              flMFMA);
  for (int rep = 0; rep < 3; rep++) {
    float ms = 0.f;
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(valu_fma, dim3(nblk), dim3(nthr), 0, 0, out, 0.999999, 1e-7, iters);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("valu_fma  %.3f ms  %.2f TFLOP/s\n", ms, flVALU / ms / 1e9);
  }
  for (int rep = 0; rep < 3; rep++) {
    float ms = 0.f;
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(mfma_f64, dim3(nblk), dim3(nthr), 0, 0, out, 0.5, iters);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("mfma_f64  %.3f ms  %.2f TFLOP/s\n", ms, flMFMA / ms / 1e9);
  }
  CHECK(hipFree(out));
  return 0;
}
