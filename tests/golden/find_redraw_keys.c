/* Searches Philox4x32-10 keys whose first output word is 0xFFFFFFFF for a
 * given counter: the one value for which the engine's index pick
 * floor(n (w + 1) / 2^32) equals n and the pick must be redrawn (the
 * reference's do-while redraw loops, src/gpu_mcmc.cu:369-372 and :549-552).
 *
 *   gcc -O3 -fopenmp -o /tmp/find_redraw_keys tests/golden/find_redraw_keys.c
 *   /tmp/find_redraw_keys C0 C1 C2 C3 KEY1      -> prints KEY0 values (a few)
 *
 * The counters used by tests/test_gpu_failures.py and the keys found are in
 * tests/golden/redraw_keys.json; tests/test_oracle_philox.py re-checks them
 * with the oracle's Philox.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint32_t first_word(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return c0;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s C0 C1 C2 C3 KEY1\n", argv[0]);
    return 2;
  }
  const uint32_t c0 = (uint32_t)strtoul(argv[1], 0, 0), c1 = (uint32_t)strtoul(argv[2], 0, 0);
  const uint32_t c2 = (uint32_t)strtoul(argv[3], 0, 0), c3 = (uint32_t)strtoul(argv[4], 0, 0);
  const uint32_t k1 = (uint32_t)strtoul(argv[5], 0, 0);
  int found = 0;
#pragma omp parallel for schedule(dynamic, 1 << 20)
  for (int64_t k = 0; k <= 0xFFFFFFFFll; ++k) {
    if (found >= 2) continue;
    if (first_word(c0, c1, c2, c3, (uint32_t)k, k1) == 0xFFFFFFFFu) {
#pragma omp critical
      {
        printf("%u\n", (uint32_t)k);
        fflush(stdout);
        ++found;
      }
    }
  }
  return 0;
}
