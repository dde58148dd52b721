// Copy kernel of known size (1 GiB read + 1 GiB written per launch) that
// calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters: the bytes per unit of
// each counter are 1 GiB divided by what the counter reads for this kernel.
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ void copy_kernel(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}
int main() {
  const size_t bytes = size_t(1) << 30, n = bytes / 16;
  double2 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  (void)hipMemset(a, 0, bytes);
  hipEvent_t e0, e1; (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  for (int it = 0; it < 5; it++) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1); (void)hipEventSynchronize(e1);
    float ms; (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy 1 GiB: %.3f ms, %.1f GB/s (read+write)\n", ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
  }
  return 0;
}
