// copy_probe.hip -- plain device copy of a known size.  Used to calibrate the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) against known
// HBM bytes: a copy of B bytes reads B and writes B.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const double4* __restrict__ src,
                                                   double4* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
  size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 256;
  size_t bytes = mib << 20;
  size_t n = bytes / sizeof(double4);
  double4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
  for (int rep = 0; rep < 5; rep++) copy_kernel<<<4096, 256>>>(a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  printf("copy_probe: %zu MiB per launch read, %zu MiB written, 5 launches\n", mib, mib);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
