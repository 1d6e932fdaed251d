// Calibration program for HBM byte counters: copies 1 GiB (64 Mi float4)
// three times with a plain streaming kernel.  The FETCH_SIZE and WRITE_SIZE
// that rocprofv3 reports for one dispatch, against the 1 GiB known to move
// each way, give the factor that turns those counters into bytes.
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void copy_k(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int main() {
  const size_t n = (size_t)64 << 20;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess) return 1;
  if (hipMemset(a, 0, n * sizeof(float4)) != hipSuccess) return 1;
  for (int r = 0; r < 3; ++r) copy_k<<<(unsigned)((n + 255) / 256), 256>>>(a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  printf("copied %zu bytes per dispatch\n", n * sizeof(float4));
  hipFree(a);
  hipFree(b);
  return 0;
}
