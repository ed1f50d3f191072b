// copy_calib.hip -- a plain copy of 256 MiB (one uint4 per thread): the
// rocprofv3 FETCH_SIZE / WRITE_SIZE readings of this kernel give the factor
// that turns those counters into HBM bytes for the walk kernel.
#include <hip/hip_runtime.h>
#include <stdio.h>
__global__ void copyk(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
int main() {
  const size_t bytes = (size_t)256 << 20, n = bytes / sizeof(uint4);
  uint4 *a, *b;
  (void)hipMalloc(&a, bytes); (void)hipMalloc(&b, bytes);
  (void)hipMemset(a, 1, bytes);
  (void)hipDeviceSynchronize();
  hipLaunchKernelGGL(copyk, dim3((n + 255) / 256), dim3(256), 0, 0, a, b, n);
  (void)hipDeviceSynchronize();
  printf("copied %zu bytes\n", bytes);
  return 0;
}
