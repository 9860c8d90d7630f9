/* Checks of the faithful arithmetic (hybridmantis_amd/csrc/fd2_ff.h) against
 * the host C library the reference C path calls (test infrastructure;
 * prints one line per check, "violations 0" expected).
 *
 * 1. fd2x_acosf / fd2x_asinf against the C library's acosf / asinf for every
 *    float in [-1, 1]: equal bits.  Also reports how many of those floats
 *    the C library does not round correctly (its result differs from the
 *    correctly rounded double acos/asin), which is why a correctly rounded
 *    arc function would not reproduce the reference.
 * 2. fd2x_sin / fd2x_cos (|x| <= 8) and fd2x_log (x in (0, 1]) against the
 *    double functions: relative error below 2^-45 on 2^24 samples each,
 *    including arguments next to multiples of pi/2 and next to 1.
 * 3. Each rounded expression of fd2_ff.h against the same expression
 *    evaluated in double and rounded to float, as the reference writes it,
 *    on 2^22 random arguments each: at most 16 differing roundings per
 *    expression (an exact value within ~2^-46 of a float rounding boundary;
 *    expected rate ~1e-6 or below).
 * 4. The crosstalk pieces (float)((double)q + 0.5) and (+ 1.0) equal the
 *    float sums q + 0.5f and q + 1.0f for every float q in [-2, 2]. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../hybridmantis_amd/csrc/fd2_ff.h"

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, sizeof f);
  return f;
}
static uint32_t to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, sizeof u);
  return u;
}

/* splitmix64 (Steele, Lea, Flood 2014), one stream per index */
static uint64_t mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unit(uint64_t i, int k) { return (double)(mix(i * 8u + (uint64_t)k) >> 11) * 0x1p-53; }

static double ffval(fd2_ff a) { return (double)a.hi + (double)a.lo; }

int main(void) {
  int fail = 0;
  /* 1. arc functions, every float in [-1, 1] */
  {
    const int64_t n_neg = 0xBF800000ll - 0x80000000ll + 1, n_pos = 0x3F800000ll + 1;
    long long bad_a = 0, bad_s = 0, ncr_a = 0, ncr_s = 0;
#pragma omp parallel for reduction(+ : bad_a, bad_s, ncr_a, ncr_s) schedule(static)
    for (int64_t k = 0; k < n_neg + n_pos; ++k) {
      const uint32_t u = k < n_neg ? (uint32_t)(0x80000000ll + k) : (uint32_t)(k - n_neg);
      const float x = from_bits(u);
      const float ga = acosf(x), gs = asinf(x);
      if (to_bits(fd2x_acosf(x)) != to_bits(ga)) bad_a++;
      if (to_bits(fd2x_asinf(x)) != to_bits(gs)) bad_s++;
      if (to_bits(ga) != to_bits((float)acos((double)x))) ncr_a++;
      if (to_bits(gs) != to_bits((float)asin((double)x))) ncr_s++;
    }
    printf("acosf: %lld floats, C library not correctly rounded on %lld, violations %lld\n",
           (long long)(n_neg + n_pos), ncr_a, bad_a);
    printf("asinf: %lld floats, C library not correctly rounded on %lld, violations %lld\n",
           (long long)(n_neg + n_pos), ncr_s, bad_s);
    fail |= bad_a != 0 || bad_s != 0;
  }
  /* 2. sin, cos, log relative errors */
  {
    const int64_t n = 1 << 24;
    double ws = 0.0, wc = 0.0, wl = 0.0;
#pragma omp parallel for reduction(max : ws, wc, wl) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      float x;
      if (i % 3 == 0)
        x = (float)((unit(i, 0) * 2.0 - 1.0) * 8.0);
      else if (i % 3 == 1) /* next to a multiple of pi/2 */
        x = (float)(1.5707963267948966 * (double)(int)(unit(i, 0) * 11.0 - 5.0) +
                    (unit(i, 1) - 0.5) * 1e-5);
      else
        x = (float)(unit(i, 0) * 1e-3);
      const double s = sin((double)x), c = cos((double)x);
      fd2_ff fs, fc;
      fd2x_sincos(x, &fs, &fc);
      if (s != 0.0) ws = fmax(ws, fabs(ffval(fs) - s) / fabs(s));
      if (c != 0.0) wc = fmax(wc, fabs(ffval(fc) - c) / fabs(c));
      float u = (float)unit(i, 2);
      if (i % 2) u = 1.0f - (float)(unit(i, 3) * 1e-4);
      if (u > 0.0f && u < 1.0f) {
        const double l = log((double)u);
        wl = fmax(wl, fabs(ffval(fd2x_log(u)) - l) / fabs(l));
      }
    }
    const double lim = 0x1p-45;
    printf("sin/cos/log max relative error %.3g %.3g %.3g (bound %.3g), violations %d\n", ws, wc,
           wl, lim, (ws > lim) + (wc > lim) + (wl > lim));
    fail |= ws > lim || wc > lim || wl > lim;
    if (fd2x_log(1.0f).hi != 0.0f) {
      printf("log(1) != 0, violations 1\n");
      fail = 1;
    }
  }
  /* 3. rounded expressions against their double evaluation */
  {
    const int64_t n = 1 << 22;
    long long mm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0, m5 = 0, m6 = 0, m7 = 0;
#pragma omp parallel for reduction(+ : m0, m1, m2, m3, m4, m5, m6, m7) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const float s = (float)unit(i, 0), rr = (float)unit(i, 1);
      const float th = (float)(unit(i, 2) * 6.283185307179586);
      /* theta = s * 2.0f * pi (double pi), K:3275 */
      if (fd2x_mul_dconst(s * 2.0f, FD2X_PI_D) != (float)((double)(s * 2.0f) * 3.14159265)) m0++;
      if (fd2x_mul_dconst(s, FD2X_TWOPI_D) != (float)((double)s * 6.283185308)) m0++;
      float x, y;
      fd2x_polar(rr, th, &x, &y);
      if (x != (float)((double)rr * cos((double)th)) || y != (float)((double)rr * sin((double)th)))
        m1++;
      const float p0 = (float)(unit(i, 3) * 2 - 1), p1 = (float)(unit(i, 4) * 2 - 1),
                  p2 = (float)(unit(i, 5) * 2 - 1);
      if (fd2x_norm3(p0, p1, p2) != (float)sqrt(pow(p0, 2) + pow(p1, 2) + pow(p2, 2))) m2++;
      const float r = p0;
      if (fd2x_sqrt_one_minus(r * r) != (float)sqrt(1.0 - (double)(r * r))) m3++;
      const float uu = (float)unit(i, 6), nib = -1.0f / (float)(1e-5 + unit(i, 7) * 0.1);
      if (uu > 0.0f && fd2x_scaled_log(nib, uu) != (float)((double)nib * log((double)uu))) m4++;
      /* Fresnel from the angles, K:1811-1871 */
      const float t1 = (float)(unit(i, 8) * 1.57), ratio = (float)(0.3 + unit(i, 9) * 0.7);
      const float t2 = asinf((float)((double)ratio * sin((double)t1)));
      float pr, pt;
      fd2x_fresnel(t1, t2, &pr, &pt);
      const double dm = (double)(t1 - t2), dp = (double)(t1 + t2);
      const float rpr =
          (float)(0.5f * (pow(tan(dm), 2) / pow(tan(dp), 2) + pow(sin(dm), 2) / pow(sin(dp), 2)));
      const double s12 = sin((double)(2.0f * t1)) * sin((double)(2.0f * t2));
      const float rpt =
          (float)(0.5f * (s12 / (pow(sin(dp), 2) * pow(cos(dm), 2)) + s12 / pow(sin(dp), 2)));
      if (to_bits(pr) != to_bits(rpr) || to_bits(pt) != to_bits(rpt)) m5++;
      if (fd2x_scaled_sin(ratio, t1) != (float)((double)ratio * sin((double)t1))) m6++;
      /* the ray/cylinder roots, K:992-998, and a refracted component */
      const float qa = (float)unit(i, 10), qb = (float)((unit(i, 11) * 2 - 1) * 10),
                  qc = (float)((unit(i, 12) * 2 - 1) * 6);
      const float disc = (qb * qb) - (4.0f * qa * qc);
      if (disc >= 0.0f) {
        float r1, r2;
        fd2x_roots(qb, disc, 2.0f * qa, &r1, &r2);
        const double sq = sqrt((double)disc);
        if (r1 != (float)(((double)(-qb) + sq) / (double)(2.0f * qa)) ||
            r2 != (float)(((double)(-qb) - sq) / (double)(2.0f * qa)))
          m7++;
      }
      const float t2b = t2 > 0.0f ? t2 : 0.5f;
      const double ct = cos((double)t2b), rt = sin((double)t2b) / sin((double)t1);
      const fd2_ff fct = fd2x_cos(t2b), frt = fd2x_div(fd2x_sin(t2b), fd2x_sin(t1));
      if (t1 > 0.0f &&
          fd2x_refract_component(p0, p1, fct, frt) != (float)((double)(-p0) * ct - rt * (double)p1))
        m7++;
    }
    mm[0] = m0; mm[1] = m1; mm[2] = m2; mm[3] = m3; mm[4] = m4; mm[5] = m5; mm[6] = m6; mm[7] = m7;
    const char* names[8] = {"double-constant products", "polar", "norm3", "sqrt(1-x)",
                            "scaled log", "Fresnel Pr/Pt", "Snell sine", "roots+refraction"};
    long long over = 0;
    for (int k = 0; k < 8; ++k) {
      printf("  %-24s %lld of %lld differ\n", names[k], mm[k], (long long)n);
      if (mm[k] > 16) over++;
    }
    printf("expressions: violations %lld\n", over);
    fail |= over != 0;
  }
  /* 4. crosstalk pieces */
  {
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t k = 0; k < 0x100000000ll; ++k) {
      const float q = from_bits((uint32_t)k);
      if (!(q >= -2.0f && q <= 2.0f)) continue;
      if ((float)((double)q + 0.5) != q + 0.5f) bad++;
      if ((float)((double)q + 1.0) != q + 1.0f) bad++;
    }
    printf("crosstalk sums: violations %lld\n", bad);
    fail |= bad != 0;
  }
  return fail;
}
