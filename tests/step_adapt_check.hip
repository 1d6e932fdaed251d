// The step-size adaptation forms of mhre_device.h against the reference's
// divisions (src/gpu_mcmc.cu:493, :518-522), on the host, exhaustively where
// the domain allows:
//   step_shrink(sd) == (float)((double)sd / 1.01) for all 2^32 float bit
//     patterns (NaNs compared as NaN);
//   rate_below(sh, den, rate_threshold(T)) == ((float)sh / (float)den < T)
//     for sh in [-3, 2100], den in [0, 2048] and 14 targets T, among them
//     0.30 and 0.25 and targets whose predecessor is odd and even;
//   mod_small(cnt, n, 1.0f / n) == cnt % n for cnt in [0, 2^23) and 19 n.
// Prints one line per check: "<name> checked N mismatches M"; exit status 0
// when every count of mismatches is 0.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "mhre_device.h"   // cuda_mcmc_amd/csrc (on the include path)

using namespace mhre;

namespace {

float from_bits(uint32_t b) {
  float f;
  memcpy(&f, &b, sizeof f);
  return f;
}

uint32_t to_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, sizeof b);
  return b;
}

// Runs fn(lo, hi) over [0, n) split among the host's threads; sums the results.
template <class F>
uint64_t parallel_count(uint64_t n, F fn) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 4;
  if (nt > 32) nt = 32;
  std::vector<uint64_t> bad(nt, 0);
  std::vector<std::thread> th;
  for (unsigned i = 0; i < nt; ++i)
    th.emplace_back([&, i] { bad[i] = fn(n * i / nt, n * (i + 1) / nt); });
  for (auto& t : th) t.join();
  uint64_t s = 0;
  for (uint64_t b : bad) s += b;
  return s;
}

}  // namespace

int main() {
  int fails = 0;
  {
    const uint64_t n = 1ull << 32;
    const uint64_t bad = parallel_count(n, [](uint64_t lo, uint64_t hi) {
      uint64_t b = 0;
      for (uint64_t i = lo; i < hi; ++i) {
        const float sd = from_bits((uint32_t)i);
        const float want = (float)((double)sd / 1.01), got = step_shrink(sd);
        if (std::isnan(want) ? !std::isnan(got) : to_bits(want) != to_bits(got)) ++b;
      }
      return b;
    });
    printf("step_shrink checked %llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)bad);
    fails += bad != 0;
  }
  {
    const float targets[] = {0.30f, 0.25f, 0.5f, 1.0f / 3.0f, 0.2f, 0.44f, 0.234f, 0.7f, 0.1f, 1.0f,
                             from_bits(0x3e99999au), from_bits(0x3e99999bu), 1e-3f, 0.999f};
    uint64_t n = 0, bad = 0;
    for (float T : targets) {
      double mid;
      int tie;
      rate_threshold(T, &mid, &tie);
      for (int den = 0; den <= 2048; ++den)
        for (int sh = -3; sh <= 2100; ++sh) {
          const bool want = (float)sh / (float)den < T;
          bad += rate_below(sh, den, mid, tie) != want;
          ++n;
        }
    }
    printf("rate_below checked %llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)bad);
    fails += bad != 0;
  }
  {
    const int ns[] = {1, 2, 3, 5, 7, 10, 31, 32, 33, 64, 100, 127, 1000, 4096, 65535, 99991, 1 << 20,
                      (1 << 22) - 1, 1 << 22};
    uint64_t n = 0, bad = 0;
    for (int m : ns) {
      const float rn = 1.0f / (float)m;
      const uint64_t cnt_n = 1ull << 23;
      bad += parallel_count(cnt_n, [&](uint64_t lo, uint64_t hi) {
        uint64_t b = 0;
        for (uint64_t c = lo; c < hi; ++c) b += mod_small((int)c, m, rn) != (int)c % m;
        return b;
      });
      n += cnt_n;
    }
    printf("mod_small checked %llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)bad);
    fails += bad != 0;
  }
  return fails ? 1 : 0;
}
