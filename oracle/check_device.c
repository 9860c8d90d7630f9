/* Exhaustive checks of the facts two shortcuts of the GPU transport rest on
 * (test infrastructure; prints the number of violations, 0 expected).
 *
 * 1. The branch-free angle test (angle_over_157 in fd2_device.h), over every
 *    float x in [-1, 1]:
 *      fd2_acos(x) >= 0;
 *      x <= 7.90e-4  ->  fd2_acos(x) > 1.57;
 *      x >= 8.03e-4  ->  fd2_acos(x) <= 1.57;
 *      7.90e-4 < x < 8.03e-4  ->  fd2_acos(x) == fd2_acos_small(x).
 * 2. The bulk-absorption bound (isotropic in fd2_device.h), over every
 *    uniform the kernel draws (u = k 2^-24, k = 1 .. 2^24) and a range of
 *    absorption coefficients: with da = (-1/bulk) fd2_log(u) and
 *    t = ((1 - u) (1/bulk)) (1 - 2^-20) in float arithmetic, t <= da, so a
 *    step ds < t is certainly shorter than the absorption length da. */
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

static long long check_absorption(void) {
  static const float bulks[] = {1e-6f, 3.3e-5f, 1e-4f, 1.234e-4f, 1e-3f, 0.0173f, 0.1f,
                                1.0f, 7.5f, 100.0f};
  long long bad = 0;
  for (unsigned b = 0; b < sizeof bulks / sizeof bulks[0]; ++b) {
    const float nib = -1.0f / bulks[b];
    const float margin = 0.99999904632568359375f; /* 1 - 2^-20 */
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t k = 1; k <= (1 << 24); ++k) {
      const float u = (float)k * (1.0f / 16777216.0f);
      const float da = nib * fd2_log(u);
      const float t = ((1.0f - u) * -nib) * margin;
      if (!(t <= da)) bad++;
    }
  }
  return bad;
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
  printf("angle test: checked %lld floats, %lld in the band, %lld violations\n",
         (long long)(n_neg + n_pos), band, bad);
  const long long bad_abs = check_absorption();
  printf("absorption bound: %lld violations\n", bad_abs);
  return bad != 0 || bad_abs != 0;
}
