// copy_calib: a plain copy of 64 MiB (one uint4 per thread), three times.  Its
// FETCH_SIZE / WRITE_SIZE counter readings under rocprofv3 --pmc give the factors
// that turn those counters into bytes for the NTT kernels' traffic figures.
#include <hip/hip_runtime.h>
#include <stdio.h>
__global__ void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
int main() {
  const size_t bytes = 64ull << 20, n = bytes / 16;
  uint4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  (void)hipMemset(a, 1, bytes);
  for (int i = 0; i < 3; i++) hipLaunchKernelGGL(copy_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 2;
  printf("copied %zu bytes x3\n", bytes);
  return 0;
}
