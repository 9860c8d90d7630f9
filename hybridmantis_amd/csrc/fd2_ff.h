/*
 * fd2_ff.h -- the reference C path's double-precision expressions, and its
 * float arc functions, evaluated with single-precision arithmetic only.
 *
 * The reference's C transport keeps its state in float but calls the double
 * <math.h> functions (sin, cos, tan, pow, log, sqrt), so each expression
 * holding one of them is evaluated in double and rounded to float on
 * assignment (main/kernel_cuda_c_ver1_0.cu, C branches; e.g. :581, :712,
 * :992-998, :1811-1871, :3271-3285).  The functions below return the float
 * such an expression yields, computed from float-float numbers (an unevaluated
 * sum hi + lo of two floats, about 48 significant bits; Dekker 1971, "A
 * floating-point technique for extending the available precision"; the
 * two-product uses the fused multiply-add).  The double evaluation and this
 * one round to the same float unless the exact value lies within about 2^-46
 * (relative) of a float rounding boundary, a chance of about 1 in 4e6 per
 * rounding.  No FP64 instruction is used, so the same source gives the same
 * bits on gfx950 (HIP, -ffp-contract=off) and on the host (C,
 * -ffp-contract=off -mfma).
 *
 * fd2x_acosf / fd2x_asinf restate the single-precision arc cosine and arc
 * sine of the host C library the reference C path calls (GNU libc 2.35:
 * the float conversions of the Sun fdlibm 1993 e_acos.c -- a rational
 * approximation with the half-angle reduction above 1/2 -- and of
 * S. L. Moshier's Cephes asinf polynomial).  Neither is correctly rounded
 * (about 5e6 of the 2^31 floats in [-1, 1] differ from the correctly rounded
 * value), so the trajectory test needs these exact sequences; every float
 * in [-1, 1] gives the C library's bits (checked exhaustively on the host).
 *
 * Argument ranges the transport needs: sin/cos/tan of floats with
 * |x| <= 8 (angles in [-pi, 2 pi]); log of floats in (0, 1].
 */
#ifndef FD2_FF_H
#define FD2_FF_H

#include "fd2_math.h"

typedef struct {
  float hi, lo;
} fd2_ff;

/* ------------------------------------------------------ error-free steps */

FD2_HD fd2_ff fd2x_pair(float hi, float lo) {
  fd2_ff r;
  r.hi = hi;
  r.lo = lo;
  return r;
}

/* a + b = s + e exactly (Knuth's two-sum, no ordering needed) */
FD2_HD fd2_ff fd2x_two_sum(float a, float b) {
  const float s = a + b;
  const float bb = s - a;
  const float e = (a - (s - bb)) + (b - bb);
  return fd2x_pair(s, e);
}

/* a + b = s + e exactly when |a| >= |b| (Dekker's fast two-sum) */
FD2_HD fd2_ff fd2x_fast_two_sum(float a, float b) {
  const float s = a + b;
  const float e = b - (s - a);
  return fd2x_pair(s, e);
}

/* a * b = p + e exactly (no underflow) */
FD2_HD fd2_ff fd2x_two_prod(float a, float b) {
  const float p = a * b;
  const float e = fmaf(a, b, -p);
  return fd2x_pair(p, e);
}

/* ------------------------------------------------ float-float arithmetic */

FD2_HD fd2_ff fd2x_add(fd2_ff a, fd2_ff b) {
  fd2_ff s = fd2x_two_sum(a.hi, b.hi);
  const fd2_ff t = fd2x_two_sum(a.lo, b.lo);
  s.lo = s.lo + t.hi;
  s = fd2x_fast_two_sum(s.hi, s.lo);
  s.lo = s.lo + t.lo;
  return fd2x_fast_two_sum(s.hi, s.lo);
}

FD2_HD fd2_ff fd2x_neg(fd2_ff a) { return fd2x_pair(-a.hi, -a.lo); }

FD2_HD fd2_ff fd2x_sub(fd2_ff a, fd2_ff b) { return fd2x_add(a, fd2x_neg(b)); }

FD2_HD fd2_ff fd2x_mul(fd2_ff a, fd2_ff b) {
  fd2_ff p = fd2x_two_prod(a.hi, b.hi);
  p.lo = fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo));
  return fd2x_fast_two_sum(p.hi, p.lo);
}

FD2_HD fd2_ff fd2x_mul_f(fd2_ff a, float b) {
  fd2_ff p = fd2x_two_prod(a.hi, b);
  p.lo = fmaf(a.lo, b, p.lo);
  return fd2x_fast_two_sum(p.hi, p.lo);
}

FD2_HD fd2_ff fd2x_sqr(fd2_ff a) { return fd2x_mul(a, a); }

/* a / b: three quotient digits.  A zero or non-finite quotient digit (b = 0,
 * a = 0, overflow, NaN) is returned as it stands, as the double division
 * would give it. */
FD2_HD fd2_ff fd2x_div(fd2_ff a, fd2_ff b) {
  const float q1 = a.hi / b.hi;
  if (!(fabsf(q1) <= 3.0e38f) || q1 == 0.0f) return fd2x_pair(q1, 0.0f);
  fd2_ff r = fd2x_sub(a, fd2x_mul_f(b, q1));
  const float q2 = r.hi / b.hi;
  r = fd2x_sub(r, fd2x_mul_f(b, q2));
  const float q3 = r.hi / b.hi;
  fd2_ff q = fd2x_fast_two_sum(q1, q2);
  return fd2x_add(q, fd2x_pair(q3, 0.0f));
}

/* sqrt(a), a >= 0 (NaN below): one Newton correction of the float root */
FD2_HD fd2_ff fd2x_sqrt(fd2_ff a) {
  if (!(a.hi > 0.0f)) return fd2x_pair(sqrtf(a.hi), 0.0f);
  const float s = sqrtf(a.hi);
  const fd2_ff r = fd2x_sub(a, fd2x_two_prod(s, s));
  return fd2x_fast_two_sum(s, r.hi / (2.0f * s));
}

/* ------------------------------------------------- elementary functions */

/* Taylor coefficients as float pairs (hi + lo = the coefficient to 2^-48):
 * sin: (-1)^k/(2k+1)!, k = 1..8; cos: (-1)^k/(2k)!, k = 1..9 */
#define FD2X_NSIN 8
#define FD2X_NCOS 9
#define FD2X_NATANH 10

/* sin(r) and cos(r), r = hi + lo with |r| <= pi/4 (1 + 2^-20): Horner in
 * r^2 over the Taylor series to r^17 / r^18 (truncation below 2^-53) */
FD2_HD void fd2x_sincos_kernel(fd2_ff r, fd2_ff* s, fd2_ff* c) {
  const float sin_hi[FD2X_NSIN] = {-0x1.555556p-3f,  0x1.111112p-7f,  -0x1.a01a02p-13f,
                                   0x1.71de3ap-19f,  -0x1.ae6456p-26f, 0x1.612462p-33f,
                                   -0x1.ae7f3ep-41f, 0x1.952c78p-49f};
  const float sin_lo[FD2X_NSIN] = {0x1.555556p-28f,  -0x1.dddddep-32f, 0x1.7f97fap-39f,
                                   0x1.55b1ccp-45f,  -0x1.fd5138p-52f, -0x1.8af25ep-58f,
                                   -0x1.ccee08p-67f, -0x1.f9ea56p-74f};
  const float cos_hi[FD2X_NCOS] = {-0x1p-1f,         0x1.555556p-5f,   -0x1.6c16c2p-10f,
                                   0x1.a01a02p-16f,  -0x1.27e4fcp-22f, 0x1.1eed8ep-29f,
                                   -0x1.93974ap-37f, 0x1.ae7f3ep-45f,  -0x1.682786p-53f};
  const float cos_lo[FD2X_NCOS] = {0.0f,             -0x1.555556p-30f, 0x1.27d27ep-35f,
                                   -0x1.7f97fap-42f, 0x1.10ec14p-47f,  0x1.ff1b12p-54f,
                                   -0x1.180f94p-62f, 0x1.ccee08p-71f,  -0x1.dcbeccp-80f};
  const fd2_ff z = fd2x_sqr(r);
  fd2_ff ps = fd2x_pair(sin_hi[FD2X_NSIN - 1], sin_lo[FD2X_NSIN - 1]);
  for (int k = FD2X_NSIN - 2; k >= 0; --k)
    ps = fd2x_add(fd2x_mul(ps, z), fd2x_pair(sin_hi[k], sin_lo[k]));
  /* sin r = r + r z P(z) */
  *s = fd2x_add(r, fd2x_mul(r, fd2x_mul(z, ps)));
  fd2_ff pc = fd2x_pair(cos_hi[FD2X_NCOS - 1], cos_lo[FD2X_NCOS - 1]);
  for (int k = FD2X_NCOS - 2; k >= 0; --k)
    pc = fd2x_add(fd2x_mul(pc, z), fd2x_pair(cos_hi[k], cos_lo[k]));
  /* cos r = 1 + z Q(z) */
  *c = fd2x_add(fd2x_pair(1.0f, 0.0f), fd2x_mul(z, pc));
}

/* sin(x) and cos(x) of a float, |x| <= 8: x - k pi/2 with pi/2 in three
 * float parts (72 bits; the products by the small integer k are exact
 * two-products), then the quadrant's kernel values */
FD2_HD void fd2x_sincos(float x, fd2_ff* s, fd2_ff* c) {
  const float p1 = 0x1.921fb6p+0f, p2 = -0x1.777a5cp-25f, p3 = -0x1.ee59dap-50f;
  const float k = fd2_rint(x * 0x1.45f306p-1f);
  fd2_ff r = fd2x_add(fd2x_pair(x, 0.0f), fd2x_two_prod(-k, p1));
  r = fd2x_add(r, fd2x_two_prod(-k, p2));
  r = fd2x_add(r, fd2x_pair(-k * p3, 0.0f));
  fd2_ff sr, cr;
  fd2x_sincos_kernel(r, &sr, &cr);
  const int q = ((int)k) & 3;
  if (q == 0) {
    *s = sr;
    *c = cr;
  } else if (q == 1) {
    *s = cr;
    *c = fd2x_neg(sr);
  } else if (q == 2) {
    *s = fd2x_neg(sr);
    *c = fd2x_neg(cr);
  } else {
    *s = fd2x_neg(cr);
    *c = sr;
  }
}

FD2_HD fd2_ff fd2x_sin(float x) {
  fd2_ff s, c;
  fd2x_sincos(x, &s, &c);
  return s;
}

FD2_HD fd2_ff fd2x_cos(float x) {
  fd2_ff s, c;
  fd2x_sincos(x, &s, &c);
  return c;
}

/* natural logarithm of a normal float x > 0: x = 2^e m with m in
 * [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) = 2 s (1 + z/3 + ... + z^10/21),
 * s = (m - 1)/(m + 1), z = s^2 <= 0.0295 (truncation below 2^-51), plus
 * e ln 2 with ln 2 in three float parts.  log(1) = 0 exactly. */
FD2_HD fd2_ff fd2x_log(float x) {
  const float at_hi[FD2X_NATANH] = {0x1.555556p-2f, 0x1.99999ap-3f, 0x1.24924ap-3f,
                                    0x1.c71c72p-4f, 0x1.745d18p-4f, 0x1.3b13b2p-4f,
                                    0x1.111112p-4f, 0x1.e1e1e2p-5f, 0x1.af286cp-5f,
                                    0x1.861862p-5f};
  const float at_lo[FD2X_NATANH] = {-0x1.555556p-27f, -0x1.99999ap-29f, -0x1.b6db6ep-28f,
                                    -0x1.c71c72p-31f, -0x1.745d18p-29f, -0x1.89d89ep-29f,
                                    -0x1.dddddep-29f, -0x1.e1e1e2p-33f, -0x1.af286cp-32f,
                                    -0x1.e79e7ap-31f};
  const float l1 = 0x1.62e430p-1f, l2 = -0x1.05c610p-29f, l3 = -0x1.950d88p-54f;
  const unsigned int b = fd2_float_to_bits(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = fd2_bits_to_float((b & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421353816986083984375f) { /* largest float below sqrt(2) */
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f; /* exact (Sterbenz) */
  const fd2_ff s = fd2x_div(fd2x_pair(f, 0.0f), fd2x_two_sum(m, 1.0f));
  const fd2_ff z = fd2x_sqr(s);
  fd2_ff pl = fd2x_pair(at_hi[FD2X_NATANH - 1], at_lo[FD2X_NATANH - 1]);
  for (int k = FD2X_NATANH - 2; k >= 0; --k)
    pl = fd2x_add(fd2x_mul(pl, z), fd2x_pair(at_hi[k], at_lo[k]));
  /* 2s + 2s z L(z) */
  const fd2_ff s2 = fd2x_pair(2.0f * s.hi, 2.0f * s.lo);
  fd2_ff lm = fd2x_add(s2, fd2x_mul(s2, fd2x_mul(z, pl)));
  if (e == 0) return lm;
  const float fe = (float)e;
  fd2_ff el = fd2x_add(fd2x_two_prod(fe, l1), fd2x_two_prod(fe, l2));
  el = fd2x_add(el, fd2x_pair(fe * l3, 0.0f));
  return fd2x_add(el, lm);
}

/* --------------------------------- the reference's expressions, as floats */

/* (float)((double)s * C) for the reference's double constants
 * pi = 3.14159265 (main/kernel_cuda_c_ver1_0.cu:57) and twopipen =
 * 6.283185308 (:56), each the exact sum of its three float parts */
FD2_HD float fd2x_mul_dconst(float s, float c0, float c1, float c2) {
  fd2_ff p = fd2x_add(fd2x_two_prod(s, c0), fd2x_two_prod(s, c1));
  p = fd2x_add(p, fd2x_two_prod(s, c2));
  return p.hi;
}
#define FD2X_PI_D 0x1.921fb6p+1f, -0x1.86e562p-24f, 0x1p-51f
#define FD2X_TWOPI_D 0x1.921fb6p+2f, -0x1.75b756p-23f, 0.0f

/* (float)((double)rr * cos((double)th)) and the sine likewise */
FD2_HD void fd2x_polar(float rr, float th, float* x, float* y) {
  fd2_ff s, c;
  fd2x_sincos(th, &s, &c);
  *x = fd2x_mul_f(c, rr).hi;
  *y = fd2x_mul_f(s, rr).hi;
}

/* (float)sqrt(pow(a,2) + pow(b,2) + pow(c,2)) */
FD2_HD float fd2x_norm3(float a, float b, float c) {
  fd2_ff t = fd2x_add(fd2x_two_prod(a, a), fd2x_two_prod(b, b));
  t = fd2x_add(t, fd2x_two_prod(c, c));
  return fd2x_sqrt(t).hi;
}

/* (float)sqrt(1.0 - (double)x) */
FD2_HD float fd2x_sqrt_one_minus(float x) {
  return fd2x_sqrt(fd2x_two_sum(1.0f, -x)).hi;
}

/* (float)((double)c * log((double)u)) */
FD2_HD float fd2x_scaled_log(float c, float u) { return fd2x_mul_f(fd2x_log(u), c).hi; }

/* The two ray/cylinder roots (float)(((double)(-b) +- sqrt((double)disc)) /
 * (double)(two_a)) */
FD2_HD void fd2x_roots(float b, float disc, float two_a, float* r1, float* r2) {
  const fd2_ff sq = fd2x_sqrt(fd2x_pair(disc, 0.0f));
  const fd2_ff den = fd2x_pair(two_a, 0.0f);
  *r1 = fd2x_div(fd2x_add(fd2x_pair(-b, 0.0f), sq), den).hi;
  *r2 = fd2x_div(fd2x_sub(fd2x_pair(-b, 0.0f), sq), den).hi;
}

/* (float)((double)ratio * sin((double)t)) */
FD2_HD float fd2x_scaled_sin(float ratio, float t) { return fd2x_mul_f(fd2x_sin(t), ratio).hi; }

/* Fresnel reflectance and transmittance from the angles t1, t2 as the
 * reference writes them (main/kernel_cuda_c_ver1_0.cu:1811-1871), with
 * dm = (float)(t1 - t2), dp = (float)(t1 + t2):
 *   Pr = 0.5 (tan^2 dm / tan^2 dp + sin^2 dm / sin^2 dp)
 *   Pt = 0.5 (s12 / (sin^2 dp cos^2 dm) + s12 / sin^2 dp),
 *   s12 = sin(2 t1) sin(2 t2) */
FD2_HD void fd2x_fresnel(float t1, float t2, float* pr, float* pt) {
  const float dm = t1 - t2, dp = t1 + t2;
  fd2_ff sdm, cdm, sdp, cdp;
  fd2x_sincos(dm, &sdm, &cdm);
  fd2x_sincos(dp, &sdp, &cdp);
  const fd2_ff tq = fd2x_div(fd2x_div(sdm, cdm), fd2x_div(sdp, cdp));
  const fd2_ff sq = fd2x_div(sdm, sdp);
  const fd2_ff r = fd2x_add(fd2x_sqr(tq), fd2x_sqr(sq));
  *pr = 0.5f * r.hi;
  const fd2_ff s12 = fd2x_mul(fd2x_sin(2.0f * t1), fd2x_sin(2.0f * t2));
  const fd2_ff sdp2 = fd2x_sqr(sdp);
  const fd2_ff a = fd2x_div(s12, fd2x_mul(sdp2, fd2x_sqr(cdm)));
  const fd2_ff b = fd2x_div(s12, sdp2);
  *pt = 0.5f * fd2x_add(a, b).hi;
}

/* one refracted direction component, trans_dir_cos
 * (main/kernel_cuda_c_ver1_0.cu:3169-3205):
 * (float)((double)(-n) * cos(t2) - (sin(t2)/sin(t1)) * (double)w), with
 * cos_t2 = cos(t2) and ratio = sin(t2)/sin(t1) as float pairs */
FD2_HD float fd2x_refract_component(float n, float w, fd2_ff cos_t2, fd2_ff ratio) {
  return fd2x_sub(fd2x_mul_f(cos_t2, -n), fd2x_mul_f(ratio, w)).hi;
}

/* ----------------------------- the C library's float arc cosine and sine */

FD2_HD float fd2x_acosf(float x) {
  const float pi = 3.1415925026e+00f, pio2_hi = 1.5707962513e+00f,
              pio2_lo = 7.5497894159e-08f;
  const float p0 = 1.6666667163e-01f, p1 = -3.2556581497e-01f, p2 = 2.0121252537e-01f,
              p3 = -4.0055535734e-02f, p4 = 7.9153501429e-04f, p5 = 3.4793309169e-05f;
  const float q1 = -2.4033949375e+00f, q2 = 2.0209457874e+00f, q3 = -6.8828397989e-01f,
              q4 = 7.7038154006e-02f;
  const unsigned int hx = fd2_float_to_bits(x), ix = hx & 0x7fffffffu;
  if (ix == 0x3f800000u) return (hx >> 31) ? pi + 2.0f * pio2_lo : 0.0f;
  if (ix > 0x3f800000u) return (x - x) / (x - x);
  if (ix < 0x3f000000u) { /* |x| < 1/2 */
    if (ix <= 0x23000000u) return pio2_hi + pio2_lo;
    const float z = x * x;
    const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
    const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
    const float r = p / q;
    return pio2_hi - (x - (pio2_lo - x * r));
  }
  if (hx >> 31) { /* x <= -1/2 */
    const float z = (1.0f + x) * 0.5f;
    const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
    const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
    const float s = sqrtf(z);
    const float r = p / q;
    const float w = r * s - pio2_lo;
    return pi - 2.0f * (s + w);
  }
  /* x >= 1/2: acos x = 2 asin(sqrt((1-x)/2)), the root split as df + c */
  const float z = (1.0f - x) * 0.5f;
  const float s = sqrtf(z);
  const float df = fd2_bits_to_float(fd2_float_to_bits(s) & 0xfffff000u);
  const float c = (z - df * df) / (s + df);
  const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
  const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
  const float r = p / q;
  const float w = r * s + c;
  return 2.0f * (df + w);
}

FD2_HD float fd2x_asinf(float x) {
  const float pio2_hi = 1.57079637050628662109375f, pio2_lo = -4.37113900018624283e-8f,
              pio4_hi = 0.785398185253143310546875f;
  const float p0 = 1.666675248e-1f, p1 = 7.495297643e-2f, p2 = 4.547037598e-2f,
              p3 = 2.417951451e-2f, p4 = 4.216630880e-2f;
  const unsigned int hx = fd2_float_to_bits(x), ix = hx & 0x7fffffffu;
  if (ix == 0x3f800000u) return x * pio2_hi + x * pio2_lo;
  if (ix > 0x3f800000u) return (x - x) / (x - x);
  if (ix < 0x3f000000u) { /* |x| < 1/2: x + x^3 P(x^2) */
    if (ix < 0x32000000u) return x;
    const float t = x * x;
    const float w = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
    return x + x * w;
  }
  /* 1/2 <= |x| < 1: pi/2 - 2 asin(sqrt((1-|x|)/2)) */
  float t = (1.0f - fabsf(x)) * 0.5f;
  float p = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
  const float s = sqrtf(t);
  if (ix >= 0x3f79999au) { /* |x| > 0.975 */
    t = pio2_hi - (2.0f * (s + s * p) - pio2_lo);
  } else {
    const float w = fd2_bits_to_float(fd2_float_to_bits(s) & 0xfffff000u);
    const float c = (t - w * w) / (s + w);
    const float r = p;
    p = 2.0f * s * r - (pio2_lo - 2.0f * c);
    const float q = pio4_hi - 2.0f * w;
    t = pio4_hi - (p - q);
  }
  return (hx >> 31) ? -t : t;
}

#endif
