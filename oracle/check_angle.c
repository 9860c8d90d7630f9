/* Exhaustive check, over every float x in [-1, 1], of the facts the GPU's
 * branch-free angle test (angle_over_157 in fd2_device.h) rests on:
 *   fd2_acos(x) >= 0;
 *   x <= 7.90e-4  ->  fd2_acos(x) > 1.57;
 *   x >= 8.03e-4  ->  fd2_acos(x) <= 1.57;
 *   7.90e-4 < x < 8.03e-4  ->  fd2_acos(x) == fd2_acos_small(x).
 * Test infrastructure; prints the number of violations (0 expected). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../hybridmantis_amd/csrc/fd2_math.h"

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, sizeof f);
  return f;
}

int main(void) {
  const float lo = 7.90e-4f, hi = 8.03e-4f;
  /* [-1, -0] then [+0, 1] as bit patterns */
  const int64_t n_neg = 0xBF800000ll - 0x80000000ll + 1, n_pos = 0x3F800000ll + 1;
  long long bad = 0, band = 0;
#pragma omp parallel for reduction(+ : bad, band) schedule(static)
  for (int64_t k = 0; k < n_neg + n_pos; ++k) {
    const uint32_t u = k < n_neg ? (uint32_t)(0x80000000ll + k) : (uint32_t)(k - n_neg);
    const float x = from_bits(u);
    const float a = fd2_acos(x);
    if (!(a >= 0.0f)) bad++;
    if (x <= lo && !(a > 1.57f)) bad++;
    if (x >= hi && !(a <= 1.57f)) bad++;
    if (x > lo && x < hi) {
      band++;
      if (fd2_acos_small(x) != a) bad++;
    }
  }
  printf("checked %lld floats, %lld in the band, %lld violations\n",
         (long long)(n_neg + n_pos), band, bad);
  return bad != 0;
}
