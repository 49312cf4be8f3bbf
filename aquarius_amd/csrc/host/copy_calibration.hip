// copy_calibration.hip -- a plain device copy of a known size (512 MiB read and
// written, three launches), used to calibrate HBM traffic counters (FETCH_SIZE,
// WRITE_SIZE) against the bytes a kernel is known to move.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_f4(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) b[i] = a[i];
}

int main() {
  const long long bytes = 512ll << 20, n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(copy_f4, dim3(8192), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("copy_f4: %lld bytes read and %lld written per launch, 3 launches\n", bytes, bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
