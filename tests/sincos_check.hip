// Accuracy of sincosrev (mhre_device.h), the hardware sine and cosine in
// revolutions after the reduction h - rint(h), and of sincosrev_hw (the same
// without the reduction, for |h| <= 255), against float64 on the host.
// Prints one line per argument range: "range R max_abs_err_sin E1 cos E2",
// then "range_hw R ..." for the unreduced form.
// Exit status 0 on success, 1 on a HIP error.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "mhre_device.h"   // cuda_mcmc_amd/csrc (on the include path)

namespace {

__global__ void k_sincosrev(const float* h, float* s, float* c, int n, int hw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    if (hw) mhre::sincosrev_hw(h[i], &s[i], &c[i]);
    else mhre::sincosrev(h[i], &s[i], &c[i]);
  }
}

bool ok(hipError_t e) {
  if (e != hipSuccess) fprintf(stderr, "HIP error: %s\n", hipGetErrorString(e));
  return e == hipSuccess;
}

}  // namespace

int main() {
  constexpr int n = 1 << 22;   // a multiple of the 256-thread block
  const float ranges[] = {0.5f, 4.0f, 255.0f, 4096.0f};
  const float ranges_hw[] = {0.5f, 2.0f, 64.0f, 255.0f};
  float *dh = nullptr, *ds = nullptr, *dc = nullptr;
  if (!ok(hipMalloc((void**)&dh, n * sizeof(float))) || !ok(hipMalloc((void**)&ds, n * sizeof(float))) ||
      !ok(hipMalloc((void**)&dc, n * sizeof(float))))
    return 1;
  std::vector<float> h(n), s(n), c(n);
  auto run = [&](float r, int hw) -> bool {
    for (int i = 0; i < n; ++i) h[i] = (float)(-r + 2.0 * r * i / (n - 1));
    if (!ok(hipMemcpy(dh, h.data(), n * sizeof(float), hipMemcpyHostToDevice))) return false;
    k_sincosrev<<<n / 256, 256>>>(dh, ds, dc, n, hw);
    if (!ok(hipGetLastError()) || !ok(hipMemcpy(s.data(), ds, n * sizeof(float), hipMemcpyDeviceToHost)) ||
        !ok(hipMemcpy(c.data(), dc, n * sizeof(float), hipMemcpyDeviceToHost)))
      return false;
    double es = 0.0, ec = 0.0;
    for (int i = 0; i < n; ++i) {
      const double t = 2.0 * M_PI * (double)h[i];   // h[i] is exact in float64
      es = std::fmax(es, std::fabs(s[i] - std::sin(t)));
      ec = std::fmax(ec, std::fabs(c[i] - std::cos(t)));
    }
    printf("%s %g max_abs_err_sin %.4e cos %.4e\n", hw ? "range_hw" : "range", r, es, ec);
    return true;
  };
  for (float r : ranges)
    if (!run(r, 0)) return 1;
  for (float r : ranges_hw)
    if (!run(r, 1)) return 1;
  hipFree(dh);
  hipFree(ds);
  hipFree(dc);
  return 0;
}
