// Calibration of the HBM byte counters: a plain device copy of 1 GiB (read 1 GiB, write
// 1 GiB), run three times.  rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch of k_copy,
// against the known 2^30 bytes each way, give the scale factors applied to the counters of
// the opacity kernels.  Built by fastspec_amd/build.py as fastspec_amd/hbm_copy_cal.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x, s = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) b[i] = a[i];
}

int main() {
  const size_t bytes = (size_t)1 << 30, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  for (int r = 0; r < 3; r++) k_copy<<<8192, 256>>>(a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 2;
  std::printf("copied %zu bytes x 3\n", bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
