// Calibration for HBM traffic counters: a plain copy of a known size.
// FETCH_SIZE / WRITE_SIZE of k_kopie divided by the bytes it must move gives
// the factor that converts those counters into bytes for the other kernels.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void k_kopie(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? (size_t)atoi(argv[1]) : 512;
  const size_t n = mib * 1024 * 1024 / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess) return 1;
  if (hipMemset(a, 1, n * sizeof(float4)) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  for (int rep = 0; rep < 6; ++rep) {
    if (hipEventRecord(e0) != hipSuccess) return 1;
    hipLaunchKernelGGL(k_kopie, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    printf("copy %zu MiB: read %zu KiB + write %zu KiB per dispatch, %.3f ms, %.1f GB/s\n", mib, mib * 1024, mib * 1024, ms,
           2.0 * n * sizeof(float4) / (ms * 1e-3) / 1e9);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
