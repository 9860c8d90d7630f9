/*
 * fdo.c -- CPU oracle: plain-C restatement of the reference C transport path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See fdo.h.
 *
 * Citations are main/kernel_cuda_c_ver1_0.cu ("K:") C branches and
 * main/hybridMANTIS_c_ver1_0.c ("drv:") unless stated otherwise.
 *
 * Precision rule followed everywhere (the reference file includes <math.h>
 * only): variables are float; sqrt/sin/cos/tan/pow/log/asin are the double
 * functions, so an expression containing one of them, or a double literal,
 * is evaluated in double and rounded to float when assigned; acosf/asinf
 * stay float.  The casts below make each promotion explicit.
 */
#include "fdo.h"

#include <math.h>

#include "../hybridmantis_amd/csrc/fd2_math.h"
#include "../hybridmantis_amd/csrc/fd2_ff.h"
#include <stdlib.h>
#include <string.h>

/* K:56-58 -- double literals, so comparisons against them are in double */
#define FDO_TWOPI 6.283185308
#define FDO_PI 3.14159265
#define FDO_EPS 8.1929093e-6
/* Safety cap on one photon's loop passes (in-column steps plus the
 * boundary_analysis / top-surface loops the reference leaves unbounded),
 * identical to the GPU kernel's: past it the photon is ended as lost. */
#define FDO_MAX_LOOPS (1 << 20)

/* ------------------------------------------------------------------ RNG */

/* RANECU, K:3694-3711: L'Ecuyer's combined MLCG, float result in (0,1]. */
float fdo_ranecu(int32_t s[2]) {
  int32_t q1 = s[0] / 53668;
  s[0] = 40014 * (s[0] - q1 * 53668) - q1 * 12211;
  int32_t q2 = s[1] / 52774;
  s[1] = 40692 * (s[1] - q2 * 52774) - q2 * 3791;
  if (s[0] < 0) s[0] += 2147483563;
  if (s[1] < 0) s[1] += 2147483399;
  int32_t z = s[0] - s[1];
  if (z < 1) z += 2147483562;
  const float scale = (float)(1.0 / 2147483563.0);
  return (float)z * scale;
}

/* (a*s) mod m, K:3626-3660.  The reference avoids 64-bit products with
 * L'Ecuyer-Cote approximate factoring; for 0 <= a,s < m both compute the
 * exact residue, which 64-bit arithmetic gives directly. */
int32_t fdo_abmodm(int32_t m, int32_t a, int32_t s) {
  int64_t r = ((int64_t)a * (int64_t)s) % (int64_t)m;
  if (r < 0) r += m;
  return (int32_t)r;
}

static int32_t powmod(int32_t a, uint64_t e, int32_t m) {
  uint64_t result = 1, base = (uint64_t)a % (uint64_t)m;
  while (e) {
    if (e & 1u) result = (result * base) % (uint64_t)m;
    base = (base * base) % (uint64_t)m;
    e >>= 1;
  }
  return (int32_t)result;
}

/* K:3516-3572: jump each MLCG ahead by leap = (batch+1)*(hpt*1000) steps:
 * seed_k = seed_input * a_k^leap mod m_k. */
void fdo_init_prng(int history_batch, int histories_per_thread, int seed_input,
                   int32_t seed[2]) {
  uint64_t leap = (uint64_t)(history_batch + 1) *
                  (uint64_t)(histories_per_thread * 1000);
  seed[0] = fdo_abmodm(2147483563, seed_input, powmod(40014, leap, 2147483563));
  seed[1] = fdo_abmodm(2147483399, seed_input, powmod(40692, leap, 2147483399));
}

/* Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void fdo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ----------------------------------------------------------- state */

typedef struct {
  const fdo_geom *g;
  fdo_tally *t;
  int rng_mode;
  int dev;            /* FDO_MATH_DEVICE: FP32 expressions, fd2_math.h */
  int emu;            /* FDO_MATH_EMULATED: faithful, double parts by fd2_ff.h */
  int32_t ranecu[2];
  uint32_t key[2];
  uint32_t event_hi;  /* counter mode: event number >> 32 (counter word 1) */
  uint32_t xs[4];     /* per-photon xoshiro128+ state */

  const fdo_event *ev;
  float pos[3], dcos[3], normal[3];
  float xc, yc;
  uint64_t nreb;      /* num_rebound of the current photon */
  int32_t loops;      /* loop passes of the current photon */
  int fate;
  float fx, fy;       /* detection point */
  int32_t fpix;
} ctx;

static float uni(ctx *c) {
  c->t->work[FDO_W_DRAW]++;
  if (c->rng_mode == FDO_RNG_RANECU) return fdo_ranecu(c->ranecu);
  /* xoshiro128+ (Blackman & Vigna, "Scrambled linear pseudorandom number
   * generators", 2018), 24 high bits -> (0,1], exact in float */
  uint32_t *s = c->xs;
  uint32_t out = s[0] + s[3];
  uint32_t t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = (s[3] << 11) | (s[3] >> 21);
  return (float)((out >> 8) + 1u) * (1.0f / 16777216.0f);
}

/* counter mode: photon k of event e draws from a xoshiro128+ stream whose
 * 128-bit state is Philox4x32-10(counter = (k, e >> 32, 0, 0), key = (seed,
 * e mod 2^32)). */
static void counter_photon(ctx *c, uint32_t photon) {
  uint32_t ctr[4] = {photon, c->event_hi, 0u, 0u};
  fdo_philox4x32_10(ctr, c->key, c->xs);
  if ((c->xs[0] | c->xs[1] | c->xs[2] | c->xs[3]) == 0u) c->xs[0] = 0x9E3779B9u;
}

/* ------------------------------------------------------------ tallies */

static void count(ctx *c, int which, int fate) {
  c->t->counters[which]++;
  c->fate = fate;
}

/* detection tally, K:3402-3419 (also K:2562-2577, 2735-2750, 2840-2855):
 * pixel (floor((x-lb)/pitch), floor((y-lb)/pitch)), inside iff lb<=x<=ub.
 * Per-history counts use hist-1 (the C branch's hist index is off by one). */
static void detect(ctx *c) {
  const fdo_geom *g = c->g;
  const float *p = c->pos;
  c->t->counters[FDO_C_DETECT]++;
  int i = (int)floor((double)((p[0] - g->lbx) / (float)g->pixel));
  int j = (int)floor((double)((p[1] - g->lby) / (float)g->pixel));
  c->fate = FDO_FATE_DETECT_OUT;
  c->fpix = -1;
  c->fx = p[0];
  c->fy = p[1];
  if (p[0] <= g->ubx && p[1] <= g->uby && p[0] >= g->lbx && p[1] >= g->lby) {
    if (i >= 0 && i < g->xdim && j >= 0 && j < g->ydim) {
      c->t->prf[(int64_t)i * g->ydim + j]++;
      c->fate = FDO_FATE_DETECT;
      c->fpix = i * g->ydim + j;
    }
  }
  int32_t h = c->ev->hist;
  if (h >= 1 && h <= c->t->num_hist) c->t->hist_count[h - 1]++;
}

/* ------------------------------------------------------- vector helpers */

static float dot3(const float *a, const float *b) {  /* K:3110-3117 */
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* renormalise when |v| leaves [1-eps, 1+eps]; the comparison is in double */
static void renorm(const ctx *c, float *v) {
  float nrm = (float)sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
  if ((double)nrm < 1.0 - FDO_EPS || (double)nrm > 1.0 + FDO_EPS) {
    if (c->dev) { /* kernel: one reciprocal */
      const float inv = 1.0f / nrm;
      v[0] = v[0] * inv;
      v[1] = v[1] * inv;
      v[2] = v[2] * inv;
    } else {
      v[0] = v[0] / nrm;
      v[1] = v[1] / nrm;
      v[2] = v[2] / nrm;
    }
  }
}

/* K:3169-3205 trans_dir_cos: specular reflection d' = d + 2(-d.n)n, or
 * refraction d' = -n cos(t2) - (sin t2 / sin t1)(d + (-d.n)n). */
static void redirect(ctx *c, float t1, float t2, int refract) {
  float *u = c->dcos;
  const float *n = c->normal;
  float m[3] = {-u[0], -u[1], -u[2]};
  float ca = dot3(m, n);
  if (!refract) {
    c->t->work[FDO_W_REFLECT]++;
    u[0] = 2.0f * ca * n[0] + u[0];
    u[1] = 2.0f * ca * n[1] + u[1];
    u[2] = 2.0f * ca * n[2] + u[2];
  } else {
    c->t->work[FDO_W_REFRACT]++;
    /* faithful mode only (device mode: refract_dev) */
    if (c->emu) {
      const fd2_ff ct = fd2x_cos(t2);
      const fd2_ff ratio = fd2x_div(fd2x_sin(t2), fd2x_sin(t1));
      u[0] = fd2x_refract_component(n[0], u[0] + ca * n[0], ct, ratio);
      u[1] = fd2x_refract_component(n[1], u[1] + ca * n[1], ct, ratio);
      u[2] = fd2x_refract_component(n[2], u[2] + ca * n[2], ct, ratio);
    } else {
      double ct = cos((double)t2);
      double ratio = sin((double)t2) / sin((double)t1);
      u[0] = (float)((double)(-n[0]) * ct - ratio * (double)(u[0] + ca * n[0]));
      u[1] = (float)((double)(-n[1]) * ct - ratio * (double)(u[1] + ca * n[1]));
      u[2] = (float)((double)(-n[2]) * ct - ratio * (double)(u[2] + ca * n[2]));
    }
  }
  renorm(c, u);
}

/* K:3261-3300 RoughSurface: perturbation p from two draws (z = 2s-1 but the
 * radius uses sqrt(1-s^2), a non-uniform direction), normalised; then
 * n <- normalise(beta*p + n). */
static void rough_n(ctx *c, float *n, float beta, int normalise);
static void rough(ctx *c, float *n, float beta) { rough_n(c, n, beta, 1); }

/* normalise = 0 (device mode only): n <- beta p + n without the final
 * normalisation, for callers that normalise a projection of it */
static void rough_n(ctx *c, float *n, float beta, int normalise) {
  c->t->work[FDO_W_ROUGH]++;
  float s = uni(c);
  float p2 = 2.0f * s - 1.0f;
  float rr = (float)sqrt((double)(1.0f - s * s));
  s = uni(c);
  float p0, p1, nb;
  if (c->dev) { /* kernel: azimuth in turns, reciprocal square root */
    float sn, cs;
    fd2_sincos_turns(s, &sn, &cs);
    p0 = rr * cs;
    p1 = rr * sn;
    nb = 0.0f;
  } else if (c->emu) {
    const float th = fd2x_mul_dconst(s * 2.0f, FD2X_PI_D);
    fd2x_polar(rr, th, &p0, &p1);
    nb = fd2x_norm3(p0, p1, p2);
  } else {
    float th = (float)((double)(s * 2.0f) * FDO_PI);
    p0 = (float)((double)rr * cos((double)th));
    p1 = (float)((double)rr * sin((double)th));
    nb = (float)sqrt(pow(p0, 2) + pow(p1, 2) + pow(p2, 2));
  }
  if (c->dev) {
    const float inv = fd2_rsqrt(p0 * p0 + p1 * p1 + p2 * p2);
    p0 = p0 * inv;
    p1 = p1 * inv;
    p2 = p2 * inv;
  } else {
    p0 = p0 / nb;
    p1 = p1 / nb;
    p2 = p2 / nb;
  }
  float r0 = beta * p0 + n[0];
  float r1 = beta * p1 + n[1];
  float r2 = beta * p2 + n[2];
  if (c->dev && !normalise) {
    n[0] = r0;
    n[1] = r1;
    n[2] = r2;
  } else if (c->dev) {
    const float inv = fd2_rsqrt(r0 * r0 + r1 * r1 + r2 * r2);
    n[0] = r0 * inv;
    n[1] = r1 * inv;
    n[2] = r2 * inv;
  } else {
    nb = c->emu ? fd2x_norm3(r0, r1, r2) : (float)sqrt(pow(r0, 2) + pow(r1, 2) + pow(r2, 2));
    n[0] = r0 / nb;
    n[1] = r1 / nb;
    n[2] = r2 / nb;
  }
}

/* x,y direction cosines rr*(cos th, sin th) with th = u*2pi: double in the
 * C path, FP32 (fd2_math.h) in device mode */
static void azimuth(ctx *c, float rr, float u, float *dx, float *dy) {
  if (c->dev) { /* kernel: azimuth in turns */
    float sn, cs;
    fd2_sincos_turns(u, &sn, &cs);
    *dx = rr * cs;
    *dy = rr * sn;
  } else if (c->emu) {
    fd2x_polar(rr, fd2x_mul_dconst(u, FD2X_TWOPI_D), dx, dy);
  } else {
    float th = (float)((double)u * FDO_TWOPI);
    *dx = (float)((double)rr * cos((double)th));
    *dy = (float)((double)rr * sin((double)th));
  }
}

/* isotropic emission direction used at photon termination and for the
 * trapped-photon restart, K:578-592 / K:616-630: cos(theta) = 2u-1,
 * radius sqrt(1.0 - r*r) evaluated in double. */
static void emit_dir(ctx *c) {
  c->t->work[FDO_W_INIT]++;
  float *u = c->dcos;
  float r = uni(c) * 2.0f - 1.0f;
  u[2] = r;
  float rr = c->dev   ? sqrtf(1.0f - r * r)
             : c->emu ? fd2x_sqrt_one_minus(r * r)
                      : (float)sqrt(1.0 - (double)(r * r));
  azimuth(c, rr, uni(c), &u[0], &u[1]);
  renorm(c, u);
}

/* first photon of an event, drv:257-277: |2u-1| > 0.01 by rejection and the
 * radius sqrt(1.0f - r*r) evaluated in float before the double sqrt. */
static void emit_first_dir(ctx *c) {
  c->t->work[FDO_W_INIT]++;
  float *u = c->dcos;
  float r = uni(c) * 2.0f - 1.0f;
  while (fabsf(r) <= 0.01f) r = uni(c) * 2.0f - 1.0f;
  u[2] = r;
  float rr = (float)sqrt((double)(1.0f - r * r));
  azimuth(c, rr, uni(c), &u[0], &u[1]);
  renorm(c, u);
}

/* ------------------------------------------------------ dist_to_surface */

static int out_of_slab(const fdo_geom *g, const float *p) {
  const float H = g->H;
  return (double)p[0] < FDO_EPS || p[0] > g->xdet || (double)p[1] < FDO_EPS ||
         p[1] > g->ydet || p[2] < -H / 2.0f || p[2] > H / 2.0f;
}

/* K:966-1185: move to the nearer of the column wall (ray/cylinder roots,
 * larger root) and the plane z=+-H/2; then lost check, then the
 * out-of-column nudge (step back 0.5 um along -d, halving on overshoot, at
 * most 10+10 tries). Returns the distance or -99 (lost / out of column). */
static float dist_to_surface(ctx *c) {
  const fdo_geom *g = c->g;
  float *p = c->pos;
  const float *u = c->dcos;
  const float R = g->R, H = g->H;
  c->t->work[FDO_W_DIST]++;
  float t0 = p[0], t1 = p[1], t2 = p[2];
  if (c->nreb == 0) { /* first column on the event (K:985-989) */
    c->xc = c->ev->col_dx == 0.0f ? c->ev->x : c->ev->x - c->ev->col_dx;
    c->yc = c->ev->y;
  }
  const float xc = c->xc, yc = c->yc;
  float qa = u[0] * u[0] + u[1] * u[1];
  float qb = 2.0f * (((p[0] - xc) * u[0]) + ((p[1] - yc) * u[1]));
  float qc = (p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc) - (R * R);
  float r1, r2;
  if (c->dev) {
    float sqf = sqrtf((qb * qb) - (4.0f * qa * qc));
    const float inv2a = 1.0f / (2.0f * qa);
    r1 = (-qb + sqf) * inv2a;
    r2 = (-qb - sqf) * inv2a;
  } else if (c->emu) {
    fd2x_roots(qb, (qb * qb) - (4.0f * qa * qc), 2.0f * qa, &r1, &r2);
  } else {
    double sq = sqrt((double)((qb * qb) - (4.0f * qa * qc)));
    r1 = (float)(((double)(-qb) + sq) / (double)(2.0f * qa));
    r2 = (float)(((double)(-qb) - sq) / (double)(2.0f * qa));
  }
  float dcyl = 999.0f;
  if (r1 >= r2)
    dcyl = r1;
  else if (r2 > r1)
    dcyl = r2;

  float d;
  if (u[2] > 0.0f) {
    if (fabs((double)(u[0] - 0.0f)) < FDO_EPS && fabs((double)(u[1] - 0.0f)) < FDO_EPS &&
        fabs((double)(u[2] - 1.0f)) < FDO_EPS) {
      d = (H / 2.0f - p[2]) / u[2];
      p[2] = H / 2.0f;
    } else {
      float dpl = (H / 2.0f - p[2]) / u[2];
      if (dpl >= dcyl) {
        d = dcyl;
        p[2] = t2 + d * u[2];
      } else {
        d = dpl;
        p[2] = H / 2.0f;
      }
    }
  } else if (u[2] < 0.0f) {
    if (fabs((double)(u[0] - 0.0f)) < FDO_EPS && fabs((double)(u[1] - 0.0f)) < FDO_EPS &&
        fabs((double)(u[2] - (-1.0f))) < FDO_EPS) {
      d = (-H / 2.0f - p[2]) / u[2];
      p[2] = -H / 2.0f;
    } else {
      float dpl = (-H / 2.0f - p[2]) / u[2];
      if (dpl >= dcyl) {
        d = dcyl;
        p[2] = t2 + d * u[2];
      } else {
        d = dpl;
        p[2] = -H / 2.0f;
      }
    }
  } else {
    d = dcyl;
    p[2] = t2 + d * u[2];
  }
  p[0] = t0 + d * u[0];
  p[1] = t1 + d * u[1];

  if (out_of_slab(g, p)) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return -99.0f;
  }

  float r_now = (float)sqrt((double)((p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc)));
  int moved = 0, shrunk = 0;
  float step = 0.5f;
  while ((double)r_now > (double)R - 1e-5 && moved < 10 && shrunk < 10) {
    c->t->work[FDO_W_NUDGE]++;
    float s0 = p[0], s1 = p[1];
    p[0] = p[0] + step * (-u[0]);
    p[1] = p[1] + step * (-u[1]);
    float r_new = (float)sqrt((double)((p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc)));
    if (r_new > r_now) {
      p[0] = s0;
      p[1] = s1;
      step = step / 2.0f;
      shrunk++;
    } else {
      r_now = r_new;
      moved++;
    }
  }
  if ((double)r_now > (double)R - 1e-5) {
    count(c, FDO_C_OUTOFCOL, FDO_FATE_OUTOFCOL);
    return -99.0f;
  }
  if (out_of_slab(g, p)) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return -99.0f;
  }
  return d;
}

/* K:702-732 isotropic: step to the surface, then bulk absorption with
 * distance -ln(u)/mu (drawn even when the step was lost). 1 = terminated. */
static int isotropic(ctx *c) {
  const fdo_geom *g = c->g;
  float ds = dist_to_surface(c);
  float da;
  if (g->bulk > 0.0f) {
    c->t->work[FDO_W_ABSLOG]++;
    if (c->dev)
      da = (-1.0f / g->bulk) * fd2_log(uni(c));
    else if (c->emu)
      da = fd2x_scaled_log(-1.0f / g->bulk, uni(c));
    else
      da = (float)((double)(-1.0f / g->bulk) * log((double)uni(c)));
  } else {
    da = 999999.0f;
  }
  if (fabs((double)(ds - (-99.0f))) < FDO_EPS) return 1;
  if (ds < da && ds >= 0.0f) return 0;
  if (ds >= da && ds >= 0.0f) {
    count(c, FDO_C_ABS_BULK, FDO_FATE_ABS_BULK);
    return 1;
  }
  return 0;
}

/* K:3372-3427 detection: at z=-H/2 the sensor reflects specularly with
 * probability sensorRefl (then the photon moves on in-column), else the
 * photon is detected.  1 = terminated. */
static int detection(ctx *c) {
  const fdo_geom *g = c->g;
  float *p = c->pos;
  if (fabs((double)(p[2] - (float)(-g->H / 2.0f))) < FDO_EPS) {
    c->t->work[FDO_W_SENSOR]++;
    if (uni(c) < g->sensor_refl) {
      c->normal[0] = 0.0f;
      c->normal[1] = 0.0f;
      c->normal[2] = 1.0f;
      redirect(c, 0.0f, 0.0f, 0);
      return isotropic(c) == 1 ? 1 : 0;
    }
    detect(c);
    return 1;
  }
  return 0;
}

/* new column orientation, K:2501-2538 (also 2716-2748 in boundary_analysis
 * and 3036-3070 in refl_bottom): normal = -d perturbed with beta=1, projected
 * to xy and renormalised, then perturbed with beta (always_rough: apply even
 * when beta <= 0, as refl_bottom does).  Accepts when acos(-d.n) <= 1.57;
 * `tries` is the caller's retry counter (<=100).  1 = photon killed. */
static int column_normal(ctx *c, const float *md, int *tries, int always_rough) {
  const fdo_geom *g = c->g;
  float *n = c->normal;
  for (;;) {
    c->t->work[FDO_W_NEWNORMAL]++;
    n[0] = md[0];
    n[1] = md[1];
    n[2] = md[2];
    /* kernel: the xy part of the unnormalised p + (-d) is normalised below,
     * which is the same direction as normalising p + (-d) first */
    rough_n(c, n, 1.0f, !c->dev);
    float den = (float)sqrt((double)(n[0] * n[0] + n[1] * n[1]));
    n[2] = 0.0f;
    if (c->dev) {
      const float inv = fd2_rsqrt(n[0] * n[0] + n[1] * n[1]);
      n[0] = n[0] * inv;
      n[1] = n[1] * inv;
    } else {
      n[0] = n[0] / den;
      n[1] = n[1] / den;
    }
    if (always_rough || g->beta > 0.0f) rough(c, n, g->beta);
    float ang = c->dev ? fd2_acos(dot3(md, n)) : c->emu ? fd2x_acosf(dot3(md, n)) : acosf(dot3(md, n));
    if (ang < 0.0f || ang > 1.57f) {
      if (*tries < 100) {
        (*tries)++;
        continue;
      }
      count(c, FDO_C_THETA1, FDO_FATE_THETA1);
      return 1;
    }
    return 0;
  }
}

static int at_top_up(const ctx *c) {
  return fabs((double)(c->pos[2] - (c->g->H / 2.0f))) < FDO_EPS && c->dcos[2] > 0.0f;
}

/* K:2985-3087 refl_bottom: after a sensor reflection between columns, fly
 * U(dmin,dmax); above the top -> stop on the top plane (or lost); else a new
 * column normal.  1 = terminated. */
static int refl_bottom(ctx *c) {
  const fdo_geom *g = c->g;
  float *p = c->pos;
  const float *u = c->dcos;
  const float H = g->H;
  c->t->work[FDO_W_FLIGHT]++;
  float tp[3] = {p[0], p[1], p[2]};
  float md[3] = {-u[0], -u[1], -u[2]};
  float d = uni(c) * (g->d_max - g->d_min) + g->d_min;
  float dtop = ((H / 2.0f) - tp[2]) / u[2];
  p[0] = tp[0] + u[0] * d;
  p[1] = tp[1] + u[1] * d;
  p[2] = tp[2] + u[2] * d;
  if ((double)p[0] < FDO_EPS || p[0] > g->xdet || (double)p[1] < FDO_EPS || p[1] > g->ydet) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return 1;
  }
  if (p[2] > H / 2.0f) {
    if (dtop < d && (double)dtop > FDO_EPS) {
      p[0] = tp[0] + u[0] * dtop;
      p[1] = tp[1] + u[1] * dtop;
      p[2] = H / 2.0f;
      return 0;
    }
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return 1;
  }
  int tries = 0;
  return column_normal(c, md, &tries, 1);
}

/* sensor plane reached between columns: reflect (specular, then
 * refl_bottom) or detect.  Returns 1 terminated, 0 alive. */
static int gap_sensor(ctx *c) {
  c->t->work[FDO_W_SENSOR]++;
  if (uni(c) < c->g->sensor_refl) {
    c->normal[0] = 0.0f;
    c->normal[1] = 0.0f;
    c->normal[2] = 1.0f;
    redirect(c, 0.0f, 0.0f, 0);
    return refl_bottom(c);
  }
  detect(c);
  return 1;
}

/* K:2421-2865 transmit with flagCCT = 0: flight between columns of length
 * U(dmin,dmax), with the detector's top (reflector/absorber), bottom
 * (sensor) and sides; then the next column's orientation.  1 = terminated. */
static int transmit_gap(ctx *c) {
  const fdo_geom *g = c->g;
  float *p = c->pos, *u = c->dcos, *n = c->normal;
  const float H = g->H;
  c->t->work[FDO_W_FLIGHT]++;
  float tp[3] = {p[0], p[1], p[2]};
  float md[3] = {-u[0], -u[1], -u[2]};
  int tries1 = 0, tries2 = 0;

  float d = uni(c) * (g->d_max - g->d_min) + g->d_min;
  p[0] = tp[0] + u[0] * d;
  p[1] = tp[1] + u[1] * d;
  p[2] = tp[2] + u[2] * d;
  float dtop = ((H / 2.0f) - tp[2]) / u[2];
  float dbot = ((-H / 2.0f) - tp[2]) / u[2];
  if ((double)p[0] < FDO_EPS || p[0] > g->xdet || (double)p[1] < FDO_EPS || p[1] > g->ydet) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return 1;
  }
  if (p[2] < -H / 2.0f || p[2] > H / 2.0f) {
    if (dtop < d && (double)dtop > FDO_EPS) {
      p[0] = tp[0] + u[0] * dtop;
      p[1] = tp[1] + u[1] * dtop;
      p[2] = H / 2.0f;
    } else if (dbot < d && (double)dbot > FDO_EPS) {
      p[0] = tp[0] + u[0] * dbot;
      p[1] = tp[1] + u[1] * dbot;
      p[2] = -H / 2.0f;
      if (gap_sensor(c)) return 1;
      if (at_top_up(c)) goto top;
    } else {
      count(c, FDO_C_LOST, FDO_FATE_LOST);
      return 1;
    }
  }
  if (column_normal(c, md, &tries1, 0)) return 1;

  if (fabs((double)(p[0] - 0.0f)) < FDO_EPS || fabs((double)(p[0] - g->xdet)) < FDO_EPS ||
      fabs((double)(p[1] - 0.0f)) < FDO_EPS || fabs((double)(p[1] - g->ydet)) < FDO_EPS) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return 1;
  }

top:  /* K:2646-2807: top reflector reached from the gap */
  if (++c->loops > FDO_MAX_LOOPS) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return 1;
  }
  if (at_top_up(c)) {
    c->t->work[FDO_W_TOP]++;
    n[0] = 0.0f;
    n[1] = 0.0f;
    n[2] = -1.0f;
    if (g->top_absfrac > 0.0f && uni(c) < g->top_absfrac) {
      count(c, FDO_C_ABS_TOP, FDO_FATE_ABS_TOP);
      return 1;
    }
    u[2] = (float)(-fabs((double)(uni(c) * 2.0f - 1.0f)));
    float rr = (float)sqrt((double)(1.0f - u[2] * u[2]));
    azimuth(c, rr, uni(c), &u[0], &u[1]);
    tp[0] = p[0];
    tp[1] = p[1];
    tp[2] = p[2];
    md[0] = -u[0];
    md[1] = -u[1];
    md[2] = -u[2];
    c->t->work[FDO_W_FLIGHT]++;
    d = uni(c) * (g->d_max - g->d_min) + g->d_min;
    dbot = ((-H / 2.0f) - tp[2]) / u[2];
    p[0] = tp[0] + u[0] * d;
    p[1] = tp[1] + u[1] * d;
    p[2] = tp[2] + u[2] * d;
    if ((double)p[0] < FDO_EPS || p[0] > g->xdet || (double)p[1] < FDO_EPS || p[1] > g->ydet) {
      count(c, FDO_C_LOST, FDO_FATE_LOST);
      return 1;
    }
    if (p[2] < -H / 2.0f || p[2] > H / 2.0f) {
      if (dbot < d && (double)dbot > FDO_EPS) {
        p[0] = tp[0] + u[0] * dbot;
        p[1] = tp[1] + u[1] * dbot;
        p[2] = -H / 2.0f;
        if (gap_sensor(c)) return 1;
        if (at_top_up(c)) goto top;
      } else {
        count(c, FDO_C_LOST, FDO_FATE_LOST);
        return 1;
      }
    }
    if (column_normal(c, md, &tries2, 0)) return 1;
  }

  /* K:2811-2859: sensor plane reached exactly */
  if (fabs((double)(p[2] - (-H / 2.0f))) < FDO_EPS) {
    if (gap_sensor(c)) return 1;
    if (at_top_up(c)) goto top;
  }
  return 0;
}

/* K:1811-1871: reflection/transmission probabilities.  In-column (inside):
 * TIR above asin(nIC/nC), else Snell sin t2 = (nC/nIC) sin t1.  From the
 * gap: only when nC/nIC >= 1.57 are Pr/Pt recomputed with sin t2 =
 * (nIC/nC) sin t1 (t1 ~ 0 uses nC/nIC); below 1.57 only the TIR branch may
 * set them and otherwise they keep their previous values. */
static void fresnel_pq(const ctx *c, float t1, float t2, float *pr, float *pt) {
  if (c->dev) { /* the kernel's form (fresnel_dev): Pr from sines/cosines */
    float s1, c1, s2, c2;
    fd2_sincos(t1, &s1, &c1);
    fd2_sincos(t2, &s2, &c2);
    *pr = fd2_fresnel_r(c1, s1, c2, s2);
    *pt = 1.0f - *pr;
    return;
  }
  if (c->emu) {
    fd2x_fresnel(t1, t2, pr, pt);
    return;
  }
  double dm = (double)(t1 - t2), dp = (double)(t1 + t2);
  *pr = (float)(0.5f * (pow(tan(dm), 2) / pow(tan(dp), 2) + pow(sin(dm), 2) / pow(sin(dp), 2)));
  double s12 = sin((double)(2.0f * t1)) * sin((double)(2.0f * t2));
  *pt = (float)(0.5f * (s12 / (pow(sin(dp), 2) * pow(cos(dm), 2)) + s12 / pow(sin(dp), 2)));
}

/* theta2 = asinf((float)((double)ratio * sin((double)t1))), K:1811-1865 */
static float snell_t2(const ctx *c, float ratio, float t1) {
  if (c->emu) return fd2x_asinf(fd2x_scaled_sin(ratio, t1));
  return asinf((float)((double)ratio * sin((double)t1)));
}

/* faithful mode (device mode: fresnel_dev) */
static void fresnel(ctx *c, int inside, float *t1, float *t2, float *pr, float *pt) {
  const float n1 = c->g->n_C, n2 = c->g->n_IC;
  if (inside) {
    if ((double)*t1 > asin((double)(n2 / n1))) {
      *pr = 1.0f;
      *pt = 0.0f;
    } else {
      if ((double)*t1 < FDO_EPS) *t1 = (float)0.00042;
      *t2 = snell_t2(c, n1 / n2, *t1);
      c->t->work[FDO_W_FRESNEL]++;
      fresnel_pq(c, *t1, *t2, pr, pt);
    }
  } else {
    if ((n1 / n2) < 1.57f) {
      if ((double)*t1 > asin((double)(n1 / n2))) {
        *pr = 1.0f;
        *pt = 0.0f;
      }
    } else if ((double)*t1 < FDO_EPS) {
      *t1 = (float)0.00042;
      *t2 = snell_t2(c, n1 / n2, *t1);
      c->t->work[FDO_W_FRESNEL]++;
      fresnel_pq(c, *t1, *t2, pr, pt);
    } else {
      *t2 = snell_t2(c, n2 / n1, *t1);
      c->t->work[FDO_W_FRESNEL]++;
      fresnel_pq(c, *t1, *t2, pr, pt);
    }
  }
}

/* The kernel's form of fresnel() (device mode): from c1 = cos(theta1) and
 * Snell's law without the angles -- sin(theta1) = sqrt(1 - c1^2),
 * sin(theta2) = (n_from/n_to) sin(theta1), cos(theta2) = sqrt(1 - sin^2),
 * Pr = fd2_fresnel_r, which needs no normalisation (Pr + Pt = 1).  Total
 * internal reflection as c1 below cos(asin(n2/n1)) (inside) or
 * cos(asin(n1/n2)) (gap); theta1 < eps (c1 == 1) becomes 0.00042 rad where
 * Pr is recomputed.  s2, c2 and pr keep their previous values where the
 * reference keeps theta2 and Pr/Pt. */
static float cos_asin_ge(double r) {
  const float f = (float)sqrt(1.0 - r * r);
  return (double)f < sqrt(1.0 - r * r) ? nextafterf(f, INFINITY) : f;
}

static void fresnel_dev(ctx *c, int inside, float c1, float *s1, float *s2, float *c2,
                        float *pr) {
  const float n1 = c->g->n_C, n2 = c->g->n_IC;
  *s1 = fd2_sqrt_unit(1.0f - c1 * c1);
  float snell = 0.0f;
  if (inside) {
    if (c1 < cos_asin_ge((double)(n2 / n1)))
      *pr = 1.0f;
    else
      snell = n1 / n2;
  } else if ((n1 / n2) < 1.57f) {
    if ((n1 / n2) <= 1.0f && c1 < cos_asin_ge((double)(n1 / n2))) *pr = 1.0f;
  } else {
    snell = c1 >= 1.0f ? n1 / n2 : n2 / n1;
  }
  if (snell != 0.0f) {
    if (c1 >= 1.0f) {
      c1 = FD2_COS_00042;
      *s1 = FD2_SIN_00042;
    }
    *s2 = snell * *s1;
    *c2 = fd2_sqrt_unit(fmaxf(1.0f - *s2 * *s2, 0.0f));
    c->t->work[FDO_W_FRESNEL]++;
    *pr = fd2_fresnel_r(c1, *s1, *c2, *s2);
  }
}

/* refraction (trans_dir_cos flag 1) in the kernel's form */
static void refract_dev(ctx *c, float s1, float s2, float c2) {
  float *u = c->dcos;
  const float *n = c->normal;
  float m[3] = {-u[0], -u[1], -u[2]};
  float ca = dot3(m, n);
  c->t->work[FDO_W_REFRACT]++;
  const float ratio = s2 / s1;
  u[0] = -n[0] * c2 - ratio * (u[0] + ca * n[0]);
  u[1] = -n[1] * c2 - ratio * (u[1] + ca * n[1]);
  u[2] = -n[2] * c2 - ratio * (u[2] + ca * n[2]);
  renorm(c, u);
}

/* K:1582-1958 boundary_analysis.  Returns 1 when the photon terminated. */
static int boundary_analysis(ctx *c) {
  const fdo_geom *g = c->g;
  float *p = c->pos, *u = c->dcos, *n = c->normal;
  const float H = g->H, R = g->R;

  /* K:1610-1634: top reflector reached inside a column */
  if (fabs((double)(p[2] - (float)(H / 2.0f))) < FDO_EPS && u[2] > 0.0f) {
    c->t->work[FDO_W_TOP]++;
    if (g->top_absfrac > 0.0f && uni(c) < g->top_absfrac) {
      count(c, FDO_C_ABS_TOP, FDO_FATE_ABS_TOP);
      return 1;
    }
    n[0] = 0.0f;
    n[1] = 0.0f;
    n[2] = -1.0f;
    u[2] = (float)(-fabs((double)(uni(c) * 2.0f - 1.0f)));
    float rr = (float)sqrt((double)(1.0f - u[2] * u[2]));
    azimuth(c, rr, uni(c), &u[0], &u[1]);
    return 0;
  }

  /* K:1639-1652: crosstalk probability, piecewise linear in depth */
  c->t->work[FDO_W_CCT]++;
  float depth20 = H * 0.2f;
  float cct = 0.0f;
  if (p[2] <= H / 2.0f && p[2] >= (H / 2.0f - depth20))
    cct = 1.0f;
  else if (p[2] < (H / 2.0f - depth20) && p[2] >= 0.0f)
    cct = c->emu ? p[2] / (2.0f * (H / 2.0f - depth20)) + 0.5f
                 : (float)((double)(p[2] / (2.0f * (H / 2.0f - depth20))) + 0.5);
  else if (p[2] < 0.0f && p[2] >= (-H / 2.0f))
    cct = c->emu ? (p[2] - (-H / 2.0f)) / (2.0f * (-H / 2.0f)) + 1.0f
                 : (float)((double)((p[2] - (-H / 2.0f)) / (2.0f * (-H / 2.0f))) + 1.0);

  if (uni(c) < cct) {
    /* K:2452-2488 transmit with flagCCT = 1: the neighbour column gets a
     * random orientation, the direction is unchanged. */
    float md[3] = {-u[0], -u[1], -u[2]};
    int tries = 0;
    if (column_normal(c, md, &tries, 0)) return 1;
    c->xc = (float)(p[0] + R * (-n[0]));
    c->yc = (float)(p[1] + R * (-n[1]));
    return 0;
  }

  int inside = 1;  /* flag_call_transmit */
  int theta_tries = 0, smooth_tries = 0, reperturb = 0, normal_tries = 0;
  float pr = 0.0f, pt = 0.0f, t1 = 0.0f, t2 = 0.0f;
  /* device mode: theta1 and theta2 as sines/cosines (theta2 = 0 at first) */
  float s1 = 0.0f, s2 = 0.0f, c2 = 1.0f;
  float md[3], on[3], od[3];

prpt:  /* K:1679 */
  if (++c->loops > FDO_MAX_LOOPS) {
    count(c, FDO_C_LOST, FDO_FATE_LOST);
    return 1;
  }
  if (inside) {
    c->t->work[FDO_W_WALLNORM]++;
    float mag = (float)sqrt((double)(((c->xc) - p[0]) * ((c->xc) - p[0]) +
                                     ((c->yc) - p[1]) * ((c->yc) - p[1])));
    if (c->dev) {
      const float ax = (c->xc) - p[0], ay = (c->yc) - p[1];
      const float inv = fd2_rsqrt(ax * ax + ay * ay);
      n[0] = ax * inv;
      n[1] = ay * inv;
    } else {
      n[0] = ((c->xc) - p[0]) / mag;
      n[1] = ((c->yc) - p[1]) / mag;
    }
    n[2] = 0.0f;
    if (g->beta > 0.0f) rough(c, n, g->beta);
  } else {
    c->xc = (float)(p[0] + R * (-n[0]));
    c->yc = (float)(p[1] + R * (-n[1]));
  }
  on[0] = n[0]; on[1] = n[1]; on[2] = n[2];
  od[0] = u[0]; od[1] = u[1]; od[2] = u[2];

reperturb:  /* K:1718 */
  n[0] = on[0]; n[1] = on[1]; n[2] = on[2];
  u[0] = od[0]; u[1] = od[1]; u[2] = od[2];
  md[0] = -u[0]; md[1] = -u[1]; md[2] = -u[2];
  if (reperturb != 0) {
    if (g->beta > 0.0f) rough(c, n, g->beta);
    if (!inside) {
      c->xc = (float)(p[0] + R * (-n[0]));
      c->yc = (float)(p[1] + R * (-n[1]));
    }
  }

no_perturbation:  /* K:1746 */
  c->t->work[FDO_W_THETA]++;
  t1 = dot3(md, n);
  if (t1 > 1.0f || t1 < 0.0f) {
    if (!inside) {
      if (column_normal(c, md, &normal_tries, 0)) return 1;
    }
    if (theta_tries < 100) {
      theta_tries++;
      goto prpt;
    }
    count(c, FDO_C_THETA1, FDO_FATE_THETA1);
    return 1;
  }
  if (c->dev) {
    fresnel_dev(c, inside, t1, &s1, &s2, &c2, &pr);
  } else {
    t1 = c->emu ? fd2x_acosf(t1) : acosf(t1);
    fresnel(c, inside, &t1, &t2, &pr, &pt);
    float tot = pr + pt;
    pr = pr / tot;
    pt = pt / tot;
  }

  if (uni(c) < pr) {  /* reflection, K:1874-1932 */
    redirect(c, t1, t2, 0);
    c->t->work[FDO_W_REFLCHK]++;
    float ang = c->dev ? fd2_acos(dot3(on, u)) : c->emu ? fd2x_acosf(dot3(on, u)) : acosf(dot3(on, u));
    if (ang > 1.57f) {
      reperturb++;
      if (reperturb < 4) goto reperturb;
      n[0] = on[0]; n[1] = on[1]; n[2] = on[2];
      u[0] = od[0]; u[1] = od[1]; u[2] = od[2];
      md[0] = -u[0]; md[1] = -u[1]; md[2] = -u[2];
      reperturb = 0;
      if (smooth_tries < 25) {
        smooth_tries++;
        goto no_perturbation;
      }
      count(c, FDO_C_THETA1, FDO_FATE_THETA1);
      return 1;
    }
    if (!inside) {
      if (transmit_gap(c)) return 1;
      goto prpt;
    }
    return 0;
  }
  /* transmission, K:1935-1949 */
  if (c->dev)
    refract_dev(c, s1, s2, c2);
  else
    redirect(c, t1, t2, 1);
  if (inside) {
    inside = 0;
    if (transmit_gap(c)) return 1;
    goto prpt;
  }
  return 0;
}

/* ------------------------------------------------------------ driver */

/* photons of one event, drv:240-299 with algo() K:553-655 unrolled:
 * each iteration is one algo() call. */
static void run_event(ctx *c, const fdo_event *ev, int64_t gid, uint32_t seed,
                      fdo_photon *rec) {
  const fdo_geom *g = c->g;
  c->ev = ev;
  if (c->rng_mode == FDO_RNG_RANECU) {
    fdo_init_prng((int)gid, 50000, (int)seed, c->ranecu);
  } else {
    c->key[0] = seed;
    c->key[1] = (uint32_t)gid;
    c->event_hi = (uint32_t)((uint64_t)gid >> 32);
  }
  int first_done = 0;
  for (int32_t k = 0; k < ev->N; ++k) {
    if (c->rng_mode == FDO_RNG_COUNTER) counter_photon(c, (uint32_t)k);
    if (!first_done) {
      emit_first_dir(c);               /* drv:257-277, normal stays 0 */
      c->normal[0] = c->normal[1] = c->normal[2] = 0.0f;
      first_done = 1;
    } else if (c->rng_mode == FDO_RNG_COUNTER) {
      emit_dir(c);                     /* K:578-602 drawn on this stream */
      c->normal[0] = c->normal[1] = c->normal[2] = 0.0f;
      if (g->beta > 0.0f) rough(c, c->normal, g->beta);
    }
    c->t->counters[FDO_C_GENERATED]++;
    c->pos[0] = ev->x;
    c->pos[1] = ev->y;
    c->pos[2] = ev->z;
    c->nreb = 0;
    c->fate = FDO_FATE_NONE;
    c->fpix = -1;
    c->fx = c->fy = 0.0f;
    int absorbed = 0, det = 0, bulk = 0;
    c->loops = 0;
    for (;;) {
      if (!absorbed && ++c->loops > FDO_MAX_LOOPS) {
        count(c, FDO_C_LOST, FDO_FATE_LOST);
        absorbed = 1;
      }
      if (!absorbed) {
        bulk = isotropic(c);
        if (!bulk) det = detection(c);
      }
      if (det || absorbed || bulk) {
        /* K:573-606: next photon's direction and normal are drawn here */
        if (c->rng_mode == FDO_RNG_RANECU) {
          emit_dir(c);
          c->normal[0] = c->normal[1] = c->normal[2] = 0.0f;
          if (g->beta > 0.0f) rough(c, c->normal, g->beta);
        }
        break;
      } else if (fabs((double)(c->dcos[2] - 0.0f)) < FDO_EPS) {
        /* K:609-644: trapped photon restarts from the event, Xc/Yc and the
         * rebound count are kept (C branch) */
        c->t->counters[FDO_C_REGEN]++;
        emit_dir(c);
        c->pos[0] = ev->x;
        c->pos[1] = ev->y;
        c->pos[2] = ev->z;
        c->normal[0] = c->normal[1] = c->normal[2] = 0.0f;
        if (g->beta > 0.0f) rough(c, c->normal, g->beta);
        absorbed = det = bulk = 0;
      } else {
        c->nreb++;
        absorbed = boundary_analysis(c);
      }
    }
    if (rec) {
      rec[k].fate = c->fate;
      rec[k].pix = c->fpix;
      rec[k].x = c->fx;
      rec[k].y = c->fy;
    }
  }
}

/* The reference C driver's per-event clears of its local xdim x ydim u64
 * image (drv:205-207 before the event, drv:248-250 after seeding), which the
 * transport never reads; on only for timing the driver as the reference
 * runs it.  The image is reachable from this global, so the clears are not
 * dead stores. */
static int g_image_rezero;
unsigned long long *fdo_rezero_image;

void fdo_set_image_rezero(int on) { g_image_rezero = on != 0; }

static void clear_image(unsigned long long *img, int xdim, int ydim) {
  for (int i = 0; i < xdim; i++)
    for (int j = 0; j < ydim; j++) img[(int64_t)i * ydim + j] = 0;
}

int fdo_run(const fdo_geom *g, const fdo_event *events, int64_t n,
            int64_t event_base, int rng_mode, int math_mode, uint32_t seed,
            fdo_tally *t) {
  if (!g || !t || (n > 0 && !events)) return -1;
  if (math_mode != FDO_MATH_FAITHFUL && math_mode != FDO_MATH_DEVICE &&
      math_mode != FDO_MATH_EMULATED)
    return -1;
  if (rng_mode != FDO_RNG_RANECU && rng_mode != FDO_RNG_COUNTER) return -1;
  ctx c;
  memset(&c, 0, sizeof c);
  c.g = g;
  c.t = t;
  c.rng_mode = rng_mode;
  c.dev = math_mode == FDO_MATH_DEVICE;
  c.emu = math_mode == FDO_MATH_EMULATED;
  fdo_photon *rec = t->photons;
  if (g_image_rezero) {
    fdo_rezero_image = (unsigned long long *)malloc((size_t)g->xdim * (size_t)g->ydim *
                                                   sizeof *fdo_rezero_image);
    if (!fdo_rezero_image) return -2;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (g_image_rezero) clear_image(fdo_rezero_image, g->xdim, g->ydim);
    if (g_image_rezero) clear_image(fdo_rezero_image, g->xdim, g->ydim);
    run_event(&c, &events[i], event_base + i, seed, rec);
    if (rec) rec += events[i].N > 0 ? events[i].N : 0;
  }
  if (g_image_rezero) {
    free(fdo_rezero_image);
    fdo_rezero_image = NULL;
  }
  return 0;
}

/* drv:343-366 (PHS).  Values past the last bin (possible when the bin size
 * does not divide max-min) go to the last bin instead of past the array. */
int fdo_phs(const uint32_t *hist_count, int32_t num_hist, int min_ph,
            int max_ph, int num_bins, int32_t *phs) {
  if (num_bins <= 0) return -1;
  int binsize = (max_ph - min_ph) / num_bins;
  if (binsize <= 0) return -1;
  int bincorr = min_ph / binsize;
  for (int32_t h = 0; h < num_hist; ++h) {
    int64_t v = hist_count[h];
    if (v <= 0) continue;
    if (v <= min_ph)
      phs[0]++;
    else if (v >= max_ph)
      phs[num_bins - 1]++;
    else {
      int64_t b = v / binsize - bincorr;
      if (b > num_bins - 1) b = num_bins - 1;
      if (b < 0) b = 0;
      phs[b]++;
    }
  }
  return 0;
}

/* ------------------------------------------------------ FP64 re-trace */

/* dist_to_surface (K:966-1185) of one step in double from a float state.
 * forced >= 0 replays the FP32 run's branch decisions (layout of
 * fd2_step.decisions: nudge moves/halvings and plane-vs-wall) instead of
 * re-deciding them in double. */
int fdo_retrace_step_f64(const float pin[3], const float dc[3], float xcf,
                         float ycf, const fdo_geom *g, int64_t forced,
                         double po[3]) {
  const double H = g->H, R = g->R, xc = xcf, yc = ycf;
  double p[3] = {pin[0], pin[1], pin[2]}, u[3] = {dc[0], dc[1], dc[2]};
  double qa = u[0] * u[0] + u[1] * u[1];
  double qb = 2.0 * ((p[0] - xc) * u[0] + (p[1] - yc) * u[1]);
  double qc = (p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc) - R * R;
  double sq = sqrt(qb * qb - 4.0 * qa * qc);
  double r1 = (-qb + sq) / (2.0 * qa), r2 = (-qb - sq) / (2.0 * qa);
  double dcyl = 999.0;
  if (r1 >= r2)
    dcyl = r1;
  else if (r2 > r1)
    dcyl = r2;
  double d, t2 = p[2];
  if (forced >= 0 && u[2] != 0.0) {
    const int plane = (int)((forced >> 25) & 1);
    const double zp = u[2] > 0.0 ? H / 2 : -H / 2;
    if (plane) {
      d = (zp - p[2]) / u[2];
      p[2] = zp;
    } else {
      d = dcyl;
      p[2] = t2 + d * u[2];
    }
  } else if (u[2] > 0.0) {
    if (fabs(u[0]) < FDO_EPS && fabs(u[1]) < FDO_EPS && fabs(u[2] - 1.0) < FDO_EPS) {
      d = (H / 2 - p[2]) / u[2];
      p[2] = H / 2;
    } else {
      double dpl = (H / 2 - p[2]) / u[2];
      if (dpl >= dcyl) { d = dcyl; p[2] = t2 + d * u[2]; }
      else { d = dpl; p[2] = H / 2; }
    }
  } else if (u[2] < 0.0) {
    if (fabs(u[0]) < FDO_EPS && fabs(u[1]) < FDO_EPS && fabs(u[2] + 1.0) < FDO_EPS) {
      d = (-H / 2 - p[2]) / u[2];
      p[2] = -H / 2;
    } else {
      double dpl = (-H / 2 - p[2]) / u[2];
      if (dpl >= dcyl) { d = dcyl; p[2] = t2 + d * u[2]; }
      else { d = dpl; p[2] = -H / 2; }
    }
  } else {
    d = dcyl;
    p[2] = t2 + d * u[2];
  }
  p[0] = pin[0] + d * u[0];
  p[1] = pin[1] + d * u[1];
#define OUT64(q) ((q)[0] < FDO_EPS || (q)[0] > g->xdet || (q)[1] < FDO_EPS || \
                  (q)[1] > g->ydet || (q)[2] < -H / 2 || (q)[2] > H / 2)
  if (OUT64(p)) { memcpy(po, p, sizeof p); return 1; }
  double rn = sqrt((p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc));
  int moved = 0, shrunk = 0;
  double step = 0.5;
  if (forced >= 0) {
    const int iters = (int)((forced >> 20) & 31);
    for (int it = 0; it < iters; ++it) {
      if ((forced >> it) & 1) {
        p[0] -= step * u[0];
        p[1] -= step * u[1];
      } else {
        step /= 2;
      }
    }
    memcpy(po, p, sizeof p);
    return 0;
  }
  while (rn > R - 1e-5 && moved < 10 && shrunk < 10) {
    double s0 = p[0], s1 = p[1];
    p[0] -= step * u[0];
    p[1] -= step * u[1];
    double r2n = sqrt((p[0] - xc) * (p[0] - xc) + (p[1] - yc) * (p[1] - yc));
    if (r2n > rn) { p[0] = s0; p[1] = s1; step /= 2; shrunk++; }
    else { rn = r2n; moved++; }
  }
  memcpy(po, p, sizeof p);
  if (rn > R - 1e-5) return 2;
  if (OUT64(p)) return 1;
#undef OUT64
  return 0;
}

/* ------------------------------------------------------- test hooks */

/* n uniforms of photon `photon`'s counter-mode stream */
void fdo_counter_stream(uint32_t seed, uint64_t event, uint32_t photon,
                        int64_t n, float *out) {
  fdo_tally t;
  memset(&t, 0, sizeof t);
  ctx c;
  memset(&c, 0, sizeof c);
  c.t = &t;
  c.rng_mode = FDO_RNG_COUNTER;
  c.key[0] = seed;
  c.key[1] = (uint32_t)event;
  c.event_hi = (uint32_t)(event >> 32);
  counter_photon(&c, photon);
  for (int64_t i = 0; i < n; ++i) out[i] = uni(&c);
}

/* fd2_math.h evaluated on the CPU: fn 0 sin, 1 cos, 2 log, 3 asin, 4 acos */
int fdo_math_eval(int fn, const float *x, float *y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: y[i] = fd2_sin(x[i]); break;
      case 1: y[i] = fd2_cos(x[i]); break;
      case 2: y[i] = fd2_log(x[i]); break;
      case 3: y[i] = fd2_asin(x[i]); break;
      case 4: y[i] = fd2_acos(x[i]); break;
      case 5: { float sn, cs; fd2_sincos_turns(x[i], &sn, &cs); y[i] = sn; break; }
      case 6: { float sn, cs; fd2_sincos_turns(x[i], &sn, &cs); y[i] = cs; break; }
      case 7: y[i] = fd2_rsqrt(x[i]); break;
      case 8: y[i] = fd2_sqrt_unit(x[i]); break;
      case 9: { /* the angle tests of K:2473 and K:1884 on fd2_acos */
        const float a = fd2_acos(x[i]);
        y[i] = (a < 0.0f || a > 1.57f) ? 1.0f : 0.0f;
        break;
      }
      /* the faithful arithmetic's pieces (fd2_ff.h) */
      case 10: y[i] = fd2x_acosf(x[i]); break;
      case 11: y[i] = fd2x_asinf(x[i]); break;
      case 12: y[i] = fd2x_sin(x[i]).hi; break;
      case 13: y[i] = fd2x_cos(x[i]).hi; break;
      case 14: y[i] = fd2x_log(x[i]).hi; break;
      default: return -1;
    }
  }
  return 0;
}

/* Fresnel probabilities (K:1811-1871 formula) for given angles */
void fdo_fresnel_probs(float t1, float t2, int math_mode, float *pr, float *pt) {
  ctx c;
  memset(&c, 0, sizeof c);
  c.dev = math_mode == FDO_MATH_DEVICE;
  c.emu = math_mode == FDO_MATH_EMULATED;
  fresnel_pq(&c, t1, t2, pr, pt);
}

/* n emission directions (K:578-592) from counter-mode streams of event 0,
 * photons 1..n (photon 0 would use the first-photon rejection) */
void fdo_emit_dirs(uint32_t seed, int math_mode, int64_t n, float *out3) {
  fdo_tally t;
  memset(&t, 0, sizeof t);
  ctx c;
  memset(&c, 0, sizeof c);
  c.t = &t;
  c.rng_mode = FDO_RNG_COUNTER;
  c.dev = math_mode == FDO_MATH_DEVICE;
  c.emu = math_mode == FDO_MATH_EMULATED;
  c.key[0] = seed;
  c.key[1] = 0;
  for (int64_t i = 0; i < n; ++i) {
    counter_photon(&c, (uint32_t)(i + 1));
    emit_dir(&c);
    out3[3 * i] = c.dcos[0];
    out3[3 * i + 1] = c.dcos[1];
    out3[3 * i + 2] = c.dcos[2];
  }
}

/* n RoughSurface perturbations (K:3261-3300) of the normal (nx,ny,nz) */
void fdo_rough_normals(uint32_t seed, int math_mode, float beta, const float n0[3],
                       int64_t n, float *out3) {
  fdo_tally t;
  memset(&t, 0, sizeof t);
  ctx c;
  memset(&c, 0, sizeof c);
  c.t = &t;
  c.rng_mode = FDO_RNG_COUNTER;
  c.dev = math_mode == FDO_MATH_DEVICE;
  c.emu = math_mode == FDO_MATH_EMULATED;
  c.key[0] = seed;
  c.key[1] = 1;
  for (int64_t i = 0; i < n; ++i) {
    counter_photon(&c, (uint32_t)i);
    float v[3] = {n0[0], n0[1], n0[2]};
    rough(&c, v, beta);
    out3[3 * i] = v[0];
    out3[3 * i + 1] = v[1];
    out3[3 * i + 2] = v[2];
  }
}
