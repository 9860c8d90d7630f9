/*
 * fd2_math.h -- single-precision elementary functions of the transport.
 *
 * Written only with IEEE-754 single-precision +, -, *, /, sqrt and explicit
 * fmaf, so the same source gives the same bits on gfx950 (HIP, built with
 * -ffp-contract=off) and on the host (C, -ffp-contract=off -mfma).  The
 * arguments the transport passes are bounded, which keeps the reductions
 * short:
 *   fd2_sincos  |x| <= a few hundred (used on [-pi, 2*pi]); Cody-Waite
 *               reduction by pi/2 in three parts, then the Taylor-based
 *               kernels sin r = r + r^3 P(r^2), cos r = 1 - r^2/2 + r^4 Q(r^2)
 *               on |r| <= pi/4;
 *   fd2_sincos_turns  sin and cos of 2 pi u (an azimuth drawn as a uniform
 *               u): the quarter-turn count 4u - rint(4u) is exact, so the
 *               reduction is one product by pi/2, then the same kernels;
 *   fd2_rsqrt   1/sqrt(x), x > 0 normal: the bit-level estimate
 *               0x5f375a86 - (bits(x) >> 1) (C. Lomont, "Fast inverse square
 *               root", 2003) and three Newton steps y (3/2 - (x/2) y^2);
 *   fd2_log     x > 0 normal; x = 2^e m, m in [sqrt(1/2), sqrt(2)),
 *               log m = 2s + s z L(z) with s = (m-1)/(m+1), z = s^2;
 *   fd2_fresnel_r  unpolarised Fresnel reflectance from sines/cosines;
 *   fd2_sqrt_unit  sqrtf for 0 and x >= 2^-96, with fewer instructions
 *               on the GPU (hardware root plus the correctly rounded fix-up);
 *   fd2_asin / fd2_acos  |x| <= 1 (NaN outside), asin x = x + x z A(z) on
 *               |x| <= 1/2 and the half-angle identity
 *               asin x = pi/2 - 2 asin(sqrt((1-x)/2)) above.
 * P, Q, L, A are least-squares fits (Chebyshev nodes, double precision) to
 * the functions' own power series; their relative errors are below 2.3e-8,
 * so each result is within a few ulp of the exact value.
 */
#ifndef FD2_MATH_H
#define FD2_MATH_H

#ifdef __HIPCC__
#define FD2_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <stdint.h>
#include <string.h>
#define FD2_HD static inline
#endif

#ifdef __HIPCC__
FD2_HD float fd2_bits_to_float(unsigned int u) { return __uint_as_float(u); }
FD2_HD unsigned int fd2_float_to_bits(float f) { return __float_as_uint(f); }
#else
FD2_HD float fd2_bits_to_float(unsigned int u) {
  float f;
  memcpy(&f, &u, sizeof f);
  return f;
}
FD2_HD unsigned int fd2_float_to_bits(float f) {
  unsigned int u;
  memcpy(&u, &f, sizeof u);
  return u;
}
#endif

/* round to nearest integer (ties away from zero) using only float ops:
 * adding and subtracting 1.5*2^23 rounds |x| < 2^22 to nearest-even */
FD2_HD float fd2_rint(float x) {
  const float big = 12582912.0f; /* 1.5 * 2^23 */
  return (x + big) - big;
}

/* sin and cos of r, |r| <= pi/4, rotated by k quarter turns (k an exact
 * small integer) */
FD2_HD void fd2_sincos_quadrant(float r, float k, float* s, float* c) {
  const float z = r * r;
  float ps = fmaf(z, 2.724937303e-06f, -1.984008122e-04f);
  ps = fmaf(z, ps, 8.333331905e-03f);
  ps = fmaf(z, ps, -1.666666716e-01f);
  const float sr = fmaf(r * z, ps, r);
  float qc = fmaf(z, -2.730065773e-07f, 2.480059811e-05f);
  qc = fmaf(z, qc, -1.388888806e-03f);
  qc = fmaf(z, qc, 4.166666791e-02f);
  const float cr = fmaf(z * z, qc, fmaf(-0.5f, z, 1.0f));
  /* quadrant q = k mod 4 (k is an exact small integer) */
  const int q = ((int)k) & 3;
  float ss, cc;
  if (q == 0) { ss = sr; cc = cr; }
  else if (q == 1) { ss = cr; cc = -sr; }
  else if (q == 2) { ss = -sr; cc = -cr; }
  else { ss = -cr; cc = sr; }
  *s = ss;
  *c = cc;
}

FD2_HD void fd2_sincos(float x, float* s, float* c) {
  if (!(fabsf(x) <= 65536.0f)) { /* outside the reduction's range, inf, NaN */
    *s = NAN;
    *c = NAN;
    return;
  }
  const float two_over_pi = 0.63661975f;
  const float pio2_1 = 1.57079637f;      /* float(pi/2)            */
  const float pio2_2 = -4.37113883e-08f; /* pi/2 - pio2_1          */
  const float pio2_3 = -1.77635684e-15f; /* remainder              */
  const float k = fd2_rint(x * two_over_pi);
  float r = fmaf(-k, pio2_1, x);
  r = fmaf(-k, pio2_2, r);
  r = fmaf(-k, pio2_3, r);
  fd2_sincos_quadrant(r, k, s, c);
}

/* sin(2 pi u), cos(2 pi u) for |u| <= 2^20 */
FD2_HD void fd2_sincos_turns(float u, float* s, float* c) {
  const float t = 4.0f * u;
  const float k = fd2_rint(t);
  fd2_sincos_quadrant((t - k) * 1.57079637f, k, s, c);
}

FD2_HD float fd2_rsqrt(float x) {
  float y = fd2_bits_to_float(0x5f375a86u - (fd2_float_to_bits(x) >> 1));
  const float hx = 0.5f * x;
  y = y * fmaf(-hx * y, y, 1.5f);
  y = y * fmaf(-hx * y, y, 1.5f);
  y = y * fmaf(-hx * y, y, 1.5f);
  return y;
}

FD2_HD float fd2_sin(float x) {
  float s, c;
  fd2_sincos(x, &s, &c);
  return s;
}

FD2_HD float fd2_cos(float x) {
  float s, c;
  fd2_sincos(x, &s, &c);
  return c;
}

FD2_HD float fd2_log(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  unsigned int b = fd2_float_to_bits(x);
  int e = (int)(b >> 23) - 127;
  if (e == -127) { /* subnormal: scale up by 2^23 */
    b = fd2_float_to_bits(x * 8388608.0f);
    e = (int)(b >> 23) - 127 - 23;
  }
  float m = fd2_bits_to_float((b & 0x007FFFFFu) | 0x3F800000u); /* [1,2) */
  if (m > 1.41421354f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float l = fmaf(z, 2.333009392e-01f, 2.855083644e-01f);
  l = fmaf(z, l, 4.000012279e-01f);
  l = fmaf(z, l, 6.666666865e-01f);
  const float ef = (float)e;
  const float ln2_hi = 0.693147182f, ln2_lo = -1.90465421e-09f;
  /* e*ln2 + 2s + s*z*L, small parts first */
  const float lo = fmaf(s * z, l, ef * ln2_lo);
  return fmaf(ef, ln2_hi, fmaf(2.0f, s, lo));
}

/* asin on [0, 1/2] */
/* sqrtf(x) for x = +-0, x >= 2^-96, +inf or NaN (NaN for x < 0).  On the GPU:
 * the hardware square root (within 1 ulp) moved to the correctly rounded
 * result by the signs of its neighbours' residuals -- the compiler's own
 * sqrtf sequence without the scaling it adds for x < 2^-96 and without its
 * +-0/+inf pass-through, which the sequence already gets right.  On the CPU:
 * sqrtf.  Callers: 1 - s^2 and 1 - r^2 of 24-bit uniforms (0 or >= 2^-24),
 * (1 - |x|)/2 in asin/acos, squared norms of unit-length vectors. */
FD2_HD float fd2_sqrt_unit(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float y = __builtin_amdgcn_sqrtf(x);
  const float ydn = __uint_as_float(__float_as_uint(y) - 1u);
  const float yup = __uint_as_float(__float_as_uint(y) + 1u);
  const float rdn = fmaf(-ydn, y, x), rup = fmaf(-yup, y, x);
  float r = y;
  if (rdn <= 0.0f) r = ydn;
  if (rup > 0.0f) r = yup;
  return r;
#else
  return sqrtf(x);
#endif
}

/* Unpolarised Fresnel reflectance from the sines and cosines of the angles
 * of incidence (1) and refraction (2):
 *   R = (Rs + Rp)/2,  Rs = sin^2(t1-t2)/sin^2(t1+t2),  Rp = tan^2(t1-t2)/tan^2(t1+t2)
 *     = sin^2(t1-t2) (cos^2(t1-t2) + cos^2(t1+t2)) / (2 sin^2(t1+t2) cos^2(t1-t2)),
 * with the angle sums from the products (one division).  R + T = 1 exactly
 * for Snell-related angles, so R is the reflection probability. */
FD2_HD float fd2_fresnel_r(float c1, float s1, float c2, float s2) {
  const float sm = s1 * c2 - c1 * s2, sp = s1 * c2 + c1 * s2;
  const float cm = c1 * c2 + s1 * s2, cp = c1 * c2 - s1 * s2;
  const float sm2 = sm * sm, cm2 = cm * cm;
  return (sm2 * (cm2 + cp * cp)) / (2.0f * (sp * sp) * cm2);
}

/* sin and cos of 0.00042 rad, the angle the reference substitutes for an
 * incidence angle below eps (float of the exact values) */
#define FD2_SIN_00042 0x1.b866e4p-12f
#define FD2_COS_00042 0x1.fffffep-1f

FD2_HD float fd2_asin_core(float x) {
  const float z = x * x;
  float a = fmaf(z, 3.368483111e-02f, 1.715300046e-02f);
  a = fmaf(z, a, 3.109983914e-02f);
  a = fmaf(z, a, 4.459947348e-02f);
  a = fmaf(z, a, 7.500094175e-02f);
  a = fmaf(z, a, 1.666666567e-01f);
  return fmaf(x * z, a, x);
}

FD2_HD float fd2_asin(float x) {
  const float ax = fabsf(x);
  float r;
  if (ax <= 0.5f) {
    r = fd2_asin_core(ax);
  } else {
    const float s = fd2_sqrt_unit((1.0f - ax) * 0.5f); /* NaN when |x| > 1 */
    const float pio2_hi = 1.57079637f, pio2_lo = -4.37113883e-08f;
    r = pio2_hi - fmaf(2.0f, fd2_asin_core(s), -pio2_lo);
  }
  return x < 0.0f ? -r : r;
}

/* acos x for |x| <= 1/2 (fd2_acos's first branch) */
FD2_HD float fd2_acos_small(float x) {
  const float pio2_hi = 1.57079637f, pio2_lo = -4.37113883e-08f;
  return pio2_hi - (fd2_asin_core(x) - pio2_lo);
}

FD2_HD float fd2_acos(float x) {
  const float pi_hi = 3.14159274f, pi_lo = -8.74227766e-08f;
  const float ax = fabsf(x);
  if (ax <= 0.5f) return fd2_acos_small(x);
  const float s = fd2_sqrt_unit((1.0f - ax) * 0.5f); /* NaN when |x| > 1 or NaN */
  const float t = 2.0f * fd2_asin_core(s);
  if (x > 0.0f) return t;
  return pi_hi - (t - pi_lo);
}

#endif
