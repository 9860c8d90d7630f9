// fd2_device.h -- device-side optical-photon transport for gfx950.
//
// The per-photon physics of fastDETECT2 as the reference C path defines it
// (main/kernel_cuda_c_ver1_0.cu, C branches), in FP32 with per-lane state in
// registers.  The control flow (including every random draw, resample loop
// and the stale-state quirks) follows the reference step for step because
// the draw order decides each photon's fate; the arithmetic is plain FP32
// (no double anywhere on the GPU).
//
// Geometry of the reference ("on-the-fly" columns): every column is an
// infinite vertical cylinder of radius R; the first one is centred on the
// event's (x,y); after crosstalk or a flight between columns the next column
// is centred at pos - R*n for a sampled inward normal n.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fd2.h"
#include "fd2_ff.h"
#include "fd2_math.h"

namespace fd2 {

// Arithmetic policies of the transport code below.
//   MathDevice    the production kernel's FP32 arithmetic (fd2_math.h):
//                 azimuths in turns, reciprocal square roots, the Fresnel
//                 branch from cosines and sines (Snell's law).
//   MathFaithful  the reference C path's own arithmetic: float state, each
//                 double <math.h> expression evaluated with float pairs
//                 (fd2_ff.h) and rounded to float as the reference rounds
//                 it, the C library's acosf/asinf restated, the angles
//                 themselves in the Fresnel branch.  Used by the trajectory
//                 test's kernel only; no FP64 instruction either way.
// The control flow, draw order and every float operation the reference
// writes in float are the same under both.
struct MathDevice {
  static constexpr bool kFaithful = false;
};
struct MathFaithful {
  static constexpr bool kFaithful = true;
};

// The reference's azimuths are u * twopipen (main/kernel_cuda_c_ver1_0.cu:56);
// here they are the turn u itself (fd2_sincos_turns).
// The reference compares floats against the double constant eps =
// 8.1929093e-6 (and 1 +- eps, R - 1e-5, asin(nIC/nC)).  Each such test is
// done here against the float threshold that gives the same answer for every
// float operand: x < c  <=>  x < lt(c), x > c  <=>  x > gt(c), with lt(c) the
// smallest float >= c and gt(c) the largest float <= c (host-computed, in
// GeomDev).
// Safety cap on one photon's loop passes (in-column steps plus the
// boundary_analysis / top-surface loops the reference leaves unbounded); a
// photon past it is ended as lost.  Typical photons need well under 10^3.
constexpr int32_t kMaxLoops = 1 << 20;

// Per-geometry transport constants, derived once on the host.
struct GeomDev {
  float xdet, ydet, H, R, n1, n2, top_absfrac, bulk, beta, d_min, d_max;
  float lbx, lby, ubx, uby, sensor_refl;
  float half_h;        // H/2
  float r_nudge;       // gt(R - 1e-5): out-of-column threshold
  float depth20;       // H*0.2 (crosstalk band)
  float neg_inv_bulk;  // -1/bulk
  // total internal reflection, theta1 > asin(n2/n1) inside a column and
  // theta1 > asin(n1/n2) from the gap, as cos(theta1) < threshold: the
  // smallest float >= cos(asin(.)) (NaN from the gap when n1/n2 > 1)
  float cos_tir_in, cos_tir_gap;
  float ratio12, ratio21;  // n1/n2, n2/n1
  float dspan;         // d_max - d_min
  float eps_lt, eps_gt;    // x < eps, x > eps
  float norm_lo, norm_hi;  // |v| < 1 - eps, |v| > 1 + eps
  int32_t pixel, xdim, ydim, num_hist;
  int32_t gap_fresnel; // (n1/n2) >= 1.57: gap side recomputes Pr/Pt
  float inv_pixel;     // 1/pixel when pixel is a power of two (x/pixel ==
                       // x*inv_pixel exactly), else 0
  // MathFaithful: total internal reflection as theta1 > the largest float
  // <= asin(n2/n1) (inside) or asin(n1/n2) (gap; NaN, never, when n1 > n2)
  float t1_tir_in, t1_tir_gap;
  unsigned long long* prf;   // xdim*ydim
  uint32_t* hist;            // num_hist
};

// ------------------------------------------------------------------ RNG

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0,
                                              uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Production stream: xoshiro128+ (Blackman & Vigna 2018) per photon, its
// 128-bit state = Philox4x32-10(counter (photon, event >> 32, 0, 0),
// key (seed, event mod 2^32)): distinct for all 2^64 event numbers, and the
// same as a 32-bit event key below 2^32.  The + scrambler (the sum of two
// state words) is the generator's authors' choice for floating-point
// output: its weak low bits never reach u(), which keeps the high 24.  The
// ** scrambler's two integer multiplies per draw cost 2% of the transport's
// time (mono25x50: 134.7 -> 132.0 ms per launch).
struct CounterRng {
  uint32_t s0, s1, s2, s3;
  __device__ __forceinline__ void seed(uint32_t key0, uint64_t event,
                                       uint32_t photon) {
    uint32_t c[4] = {photon, (uint32_t)(event >> 32), 0u, 0u};
    philox4x32_10(c, key0, (uint32_t)event);
    s0 = c[0]; s1 = c[1]; s2 = c[2]; s3 = c[3];
    if ((s0 | s1 | s2 | s3) == 0u) s0 = 0x9E3779B9u;
  }
  __device__ __forceinline__ uint32_t next() {
    const uint32_t out = s0 + s3;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = (s3 << 11) | (s3 >> 21);
    return out;
  }
  // 24 random bits -> (0,1], exact in float
  __device__ __forceinline__ float u() {
    return (float)((next() >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// Trajectory-test stream: RANECU (L'Ecuyer 1988) exactly as
// main/kernel_cuda_c_ver1_0.cu:3694-3711, float result in (0,1].
struct RanecuRng {
  int32_t s0, s1;
  __device__ __forceinline__ float u() {
    int32_t q1 = s0 / 53668;
    s0 = 40014 * (s0 - q1 * 53668) - q1 * 12211;
    int32_t q2 = s1 / 52774;
    s1 = 40692 * (s1 - q2 * 52774) - q2 * 3791;
    if (s0 < 0) s0 += 2147483563;
    if (s1 < 0) s1 += 2147483399;
    int32_t z = s0 - s1;
    if (z < 1) z += 2147483562;
    return (float)z * (float)(1.0 / 2147483563.0);
  }
};

// a^e mod m with 64-bit products (the reference's jump-ahead,
// main/kernel_cuda_c_ver1_0.cu:3516-3572)
__device__ __forceinline__ uint32_t powmod_u32(uint32_t a, uint64_t e,
                                               uint32_t m) {
  uint64_t r = 1, b = a % m;
  while (e) {
    if (e & 1u) r = (r * b) % m;
    b = (b * b) % m;
    e >>= 1;
  }
  return (uint32_t)r;
}

__device__ __forceinline__ void ranecu_init(RanecuRng& g, int batch,
                                            int32_t seed_input) {
  uint64_t leap = (uint64_t)(batch + 1) * (uint64_t)(50000 * 1000);
  uint64_t s = (uint64_t)(uint32_t)seed_input;
  g.s0 = (int32_t)((s * powmod_u32(40014u, leap, 2147483563u)) % 2147483563u);
  g.s1 = (int32_t)((s * powmod_u32(40692u, leap, 2147483399u)) % 2147483399u);
}

// ----------------------------------------------------------- tallies

// Tally atomics (PRF, history counts, radial bins) through
// global (address space 1) pointers.  The tally pointers are read from the
// geometry table, so the compiler cannot tell their address space and would
// issue flat atomics; a flat operation counts against lgkmcnt as well as
// vmcnt, and the wave's next wait for an LDS or scalar load would then wait
// for the atomic to reach L2 (on every detection, at the pencil beam's few
// hot pixels).  A global atomic without return counts against vmcnt only.
template <class T>
__device__ __forceinline__ void tally_add(T* p, T v) {
  __hip_atomic_fetch_add((__attribute__((address_space(1))) T*)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

enum Fate : int {
  kNone = 0, kDetect = 1, kDetectOut = 2, kAbsTop = 3, kAbsBulk = 4,
  kLost = 5, kOutOfCol = 6, kTheta1 = 7,
  kResume = 8,     // alive: reflected by the sensor, its next in-column step is due
  kNewColumn = 9   // alive: a crosstalk's neighbour-column sample is due
                   // (pix holds the samples rejected so far)
};

// photon state, all in registers.  The event's position and history stay
// in the event record; the caller passes them where they are needed.
struct Photon {
  float px, py, pz;   // position (the detection point once detected)
  float ux, uy, uz;   // direction cosines
  float nx, ny, nz;   // surface normal
  float xc, yc;       // current column axis (the event's first column at start)
  int32_t loops;      // loop passes so far (kMaxLoops cap)
  int fate;
  int32_t pix;        // detected pixel as row << 16 | column (pix_index)
};

// A new photon of event ev at its emission point.  The first column is
// centred on the event (x - col_dx, y): the reference re-centres at the
// first dist_to_surface of a photon (main/kernel_cuda_c_ver1_0.cu:985-989),
// and nothing reads the centre before that.
__device__ __forceinline__ void start_photon(Photon& p, const fd2_event& ev) {
  p.px = ev.x;
  p.py = ev.y;
  p.pz = ev.z;
  p.xc = ev.x - ev.col_dx;
  p.yc = ev.y;
  p.loops = 0;
  p.fate = kNone;
  p.pix = -1;
  p.nx = p.ny = p.nz = 0.0f;
}

// Detection: pixel floor((x-lb)/pitch), inside iff lb <= x <= ub
// (main/kernel_cuda_c_ver1_0.cu:3402-3419).  Records the fate and pixel
// (row << 16 | column, sides <= FD2_MAX_PRF_SIDE); the kernels add the PRF
// count and the per-history count (at hist-1) when the photon ends.
__device__ __forceinline__ void tally_detect(Photon& p, const GeomDev& g) {
  p.fate = kDetectOut;
  p.pix = -1;
  const float ax = p.px - g.lbx, ay = p.py - g.lby;
  float qx, qy;
  if (g.inv_pixel != 0.0f) {
    qx = ax * g.inv_pixel;
    qy = ay * g.inv_pixel;
  } else {
    qx = ax / (float)g.pixel;
    qy = ay / (float)g.pixel;
  }
  int i = (int)floorf(qx);
  int j = (int)floorf(qy);
  if (p.px <= g.ubx && p.py <= g.uby && p.px >= g.lbx && p.py >= g.lby &&
      i >= 0 && i < g.xdim && j >= 0 && j < g.ydim) {
    p.fate = kDetect;
    p.pix = (i << 16) | j;
  }
}

// PRF index (row-major, xdim x ydim) of a detected photon's pixel
__device__ __forceinline__ int32_t pix_index(int32_t pix, const GeomDev& g) {
  return (pix >> 16) * g.ydim + (pix & 0xFFFF);
}

// ------------------------------------------------------- vector helpers

template <class M = MathDevice>
__device__ __forceinline__ void renorm(float& x, float& y, float& z,
                                       const GeomDev& g) {
  if constexpr (M::kFaithful) {  // (float)sqrt((double)s) is sqrtf(s)
    const float nrm = sqrtf(x * x + y * y + z * z);
    if (nrm < g.norm_lo || nrm > g.norm_hi) {
      x = x / nrm;
      y = y / nrm;
      z = z / nrm;
    }
  } else {
    float nrm = fd2_sqrt_unit(x * x + y * y + z * z);
    if (nrm < g.norm_lo || nrm > g.norm_hi) {
      const float inv = 1.0f / nrm;
      x = x * inv;
      y = y * inv;
      z = z * inv;
    }
  }
}

// trans_dir_cos (main/kernel_cuda_c_ver1_0.cu:3169-3205)
template <class M = MathDevice>
__device__ __forceinline__ void reflect(Photon& p, const GeomDev& g) {
  float ca = -p.ux * p.nx + -p.uy * p.ny + -p.uz * p.nz;
  p.ux = 2.0f * ca * p.nx + p.ux;
  p.uy = 2.0f * ca * p.ny + p.uy;
  p.uz = 2.0f * ca * p.nz + p.uz;
  renorm<M>(p.ux, p.uy, p.uz, g);
}

// refraction with sin(theta1) = s1, sin(theta2) = s2, cos(theta2) = c2
__device__ __forceinline__ void refract(Photon& p, float s1, float s2, float c2,
                                        const GeomDev& g) {
  float ca = -p.ux * p.nx + -p.uy * p.ny + -p.uz * p.nz;
  float ratio = s2 / s1;
  p.ux = -p.nx * c2 - ratio * (p.ux + ca * p.nx);
  p.uy = -p.ny * c2 - ratio * (p.uy + ca * p.ny);
  p.uz = -p.nz * c2 - ratio * (p.uz + ca * p.nz);
  renorm(p.ux, p.uy, p.uz, g);
}

// refraction from the angles theta1, theta2, MathFaithful:
// d' = -n cos(t2) - (sin t2 / sin t1)(d + (-d.n) n), each component a double
// expression rounded to float
__device__ __forceinline__ void refract_angles(Photon& p, float t1, float t2,
                                               const GeomDev& g) {
  const float ca = -p.ux * p.nx + -p.uy * p.ny + -p.uz * p.nz;
  const fd2_ff ct = fd2x_cos(t2);
  const fd2_ff ratio = fd2x_div(fd2x_sin(t2), fd2x_sin(t1));
  p.ux = fd2x_refract_component(p.nx, p.ux + ca * p.nx, ct, ratio);
  p.uy = fd2x_refract_component(p.ny, p.uy + ca * p.ny, ct, ratio);
  p.uz = fd2x_refract_component(p.nz, p.uz + ca * p.nz, ct, ratio);
  renorm<MathFaithful>(p.ux, p.uy, p.uz, g);
}

// x, y direction cosines rr (cos, sin) of the azimuth 2 pi u: in turns
// (MathDevice), or theta = (float)(u * twopipen) and the double products
// (MathFaithful, main/kernel_cuda_c_ver1_0.cu:582-584)
template <class M>
__device__ __forceinline__ void azimuth(float rr, float u, float& x, float& y) {
  if constexpr (M::kFaithful) {
    fd2x_polar(rr, fd2x_mul_dconst(u, FD2X_TWOPI_D), &x, &y);
  } else {
    float sn, cs;
    fd2_sincos_turns(u, &sn, &cs);
    x = rr * cs;
    y = rr * sn;
  }
}

// RoughSurface (main/kernel_cuda_c_ver1_0.cu:3261-3300): perturbation with
// z = 2s-1 and radius sqrt(1-s^2) (s, not z), normalised, then
// n <- normalise(beta*p + n).  rough_offset draws the normalised p.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ void rough_offset(Rng& rng, float& qx, float& qy, float& qz) {
  float s = rng.u();
  qz = 2.0f * s - 1.0f;
  if constexpr (M::kFaithful) {
    // theta = s * 2.0f * pi with the double pi (K:3275), the double
    // cos/sin products and sqrt(pow + pow + pow) (K:3277-3285)
    const float rr = sqrtf(1.0f - s * s);
    s = rng.u();
    fd2x_polar(rr, fd2x_mul_dconst(s * 2.0f, FD2X_PI_D), &qx, &qy);
    const float nb = fd2x_norm3(qx, qy, qz);
    qx = qx / nb;
    qy = qy / nb;
    qz = qz / nb;
  } else {
    float rr = fd2_sqrt_unit(1.0f - s * s);
    s = rng.u();
    float sn, cs;
    fd2_sincos_turns(s, &sn, &cs);  // azimuth 2 pi s
    qx = rr * cs;
    qy = rr * sn;
    float inv = fd2_rsqrt(qx * qx + qy * qy + qz * qz);
    qx = qx * inv;
    qy = qy * inv;
    qz = qz * inv;
  }
}

template <class Rng, class M = MathDevice>
__device__ __forceinline__ void rough(Rng& rng, float& nx, float& ny, float& nz,
                                      float beta) {
  float qx, qy, qz;
  rough_offset<Rng, M>(rng, qx, qy, qz);
  float rx = beta * qx + nx, ry = beta * qy + ny, rz = beta * qz + nz;
  if constexpr (M::kFaithful) {
    const float nb = fd2x_norm3(rx, ry, rz);
    nx = rx / nb;
    ny = ry / nb;
    nz = rz / nb;
  } else {
    float inv = fd2_rsqrt(rx * rx + ry * ry + rz * rz);
    nx = rx * inv;
    ny = ry * inv;
    nz = rz * inv;
  }
}

// isotropic direction at photon termination / trapped restart
// (main/kernel_cuda_c_ver1_0.cu:578-592)
// (MathFaithful: the radius is sqrt(1.0 - r*r) in double, K:581)
template <class Rng, class M = MathDevice>
__device__ __forceinline__ void emit_dir(Rng& rng, Photon& p, const GeomDev& g) {
  float r = rng.u() * 2.0f - 1.0f;
  p.uz = r;
  float rr;
  if constexpr (M::kFaithful)
    rr = fd2x_sqrt_one_minus(r * r);
  else
    rr = fd2_sqrt_unit(1.0f - r * r);
  azimuth<M>(rr, rng.u(), p.ux, p.uy);
  renorm<M>(p.ux, p.uy, p.uz, g);
}

// first photon of an event: |2u-1| > 0.01 by rejection
// (main/hybridMANTIS_c_ver1_0.c:257-277)
template <class Rng, class M = MathDevice>
__device__ __forceinline__ void emit_first_dir(Rng& rng, Photon& p,
                                               const GeomDev& g) {
  float r = rng.u() * 2.0f - 1.0f;
  while (fabsf(r) <= 0.01f) r = rng.u() * 2.0f - 1.0f;
  p.uz = r;
  float rr = fd2_sqrt_unit(1.0f - r * r);
  azimuth<M>(rr, rng.u(), p.ux, p.uy);
  renorm<M>(p.ux, p.uy, p.uz, g);
}

// ---------------------------------------------------------- transport

__device__ __forceinline__ bool out_of_slab(const Photon& p, const GeomDev& g) {
  return p.px < g.eps_lt || p.px > g.xdet || p.py < g.eps_lt || p.py > g.ydet ||
         p.pz < -g.half_h || p.pz > g.half_h;
}


// Optional per-step trace (FP64 re-trace check): 14 floats per step.
struct StepSink {
  float* buf;
  int32_t* count;
  int32_t max_steps;
  int32_t photon;
  bool on;
};

// dist_to_surface (main/kernel_cuda_c_ver1_0.cu:966-1185): move to the
// nearer of the column wall (larger ray/cylinder root) and the plane
// z = +-H/2; lost check; out-of-column nudge (back 0.5 um along -d, halving
// on overshoot, at most 10+10 tries).  Returns the distance, or -99 when the
// photon was lost or left its column (fate set).
template <bool TRACE, class M = MathDevice>
__device__ __forceinline__ float dist_to_surface(Photon& p, const GeomDev& g,
                                                 StepSink* sink) {
  const float H2 = g.half_h, R = g.R;
  const float t0 = p.px, t1 = p.py, t2 = p.pz;
  const float xc = p.xc, yc = p.yc;
  const float dx0 = p.px - xc, dy0 = p.py - yc;
  const float qa = p.ux * p.ux + p.uy * p.uy;
  const float qb = 2.0f * (dx0 * p.ux + dy0 * p.uy);
  const float qc = dx0 * dx0 + dy0 * dy0 - R * R;
  float r1, r2;
  if constexpr (M::kFaithful) {  // double root and quotients (K:992-998)
    fd2x_roots(qb, qb * qb - 4.0f * qa * qc, 2.0f * qa, &r1, &r2);
  } else {
    const float sq = sqrtf(qb * qb - 4.0f * qa * qc);
    const float inv2a = 1.0f / (2.0f * qa);
    r1 = (-qb + sq) * inv2a;
    r2 = (-qb - sq) * inv2a;
  }
  float dcyl = 999.0f;
  if (r1 >= r2)
    dcyl = r1;
  else if (r2 > r1)
    dcyl = r2;
  // the plane z = +H/2 (going up) or -H/2 (going down) when it is nearer
  // than the wall, or when the direction is along the axis
  float d = dcyl;
  int plane = 0;  // traced: the plane was the nearer surface
  const bool up = p.uz > 0.0f;
  const float zpl = up ? H2 : -H2;
  if (up || p.uz < 0.0f) {
    const float dpl = (zpl - p.pz) / p.uz;
    const bool axial = fabsf(p.ux) < g.eps_lt && fabsf(p.uy) < g.eps_lt &&
                       fabsf(p.uz - (up ? 1.0f : -1.0f)) < g.eps_lt;
    if (axial || !(dpl >= dcyl)) {
      d = dpl;
      plane = 1;
    }
  }
  p.pz = plane ? zpl : t2 + d * p.uz;
  p.px = t0 + d * p.ux;
  p.py = t1 + d * p.uy;

  int code = 0;
  uint32_t decisions = 0, iters = 0;  // traced nudge moves (1) / halvings (0)
  if (out_of_slab(p, g)) {
    code = 1;
  } else {
    // Radii from the column axis.  fd2_sqrt_unit equals sqrtf for arguments
    // >= 2^-96; a smaller one has a root below 2^-48, and the radii are only
    // compared with r_nudge (~R) and, inside the loop, with an rn above it,
    // so every decision is the same.
    float dx = p.px - xc, dy = p.py - yc;
    float rn = fd2_sqrt_unit(dx * dx + dy * dy);
    int moved = 0, shrunk = 0;
    float step = 0.5f;
    while (rn > g.r_nudge && moved < 10 && shrunk < 10) {
      const float cx = p.px + step * (-p.ux), cy = p.py + step * (-p.uy);
      const float ex = cx - xc, ey = cy - yc;
      const float rc = fd2_sqrt_unit(ex * ex + ey * ey);
      if (rc > rn) {
        step = step / 2.0f;
        shrunk++;
      } else {
        p.px = cx;
        p.py = cy;
        rn = rc;
        moved++;
        if (TRACE) decisions |= 1u << iters;
      }
      if (TRACE) ++iters;
    }
    if (rn > g.r_nudge)
      code = 2;
    else if (out_of_slab(p, g))
      code = 1;
  }
  if (TRACE && sink->on) {
    const int slot = atomicAdd(sink->count, 1);
    if (slot < sink->max_steps) {
      float* o = sink->buf + (size_t)slot * 14;
      o[0] = t0; o[1] = t1; o[2] = t2;
      o[3] = p.ux; o[4] = p.uy; o[5] = p.uz;
      o[6] = xc; o[7] = yc;
      o[8] = p.px; o[9] = p.py; o[10] = p.pz;
      o[11] = __int_as_float(code);
      o[12] = __int_as_float(sink->photon);
      o[13] = __uint_as_float(decisions | (iters << 20) | ((uint32_t)plane << 25));
    }
  }
  if (code == 1) {
    p.fate = kLost;
    return -99.0f;
  }
  if (code == 2) {
    p.fate = kOutOfCol;
    return -99.0f;
  }
  return d;
}

// isotropic (main/kernel_cuda_c_ver1_0.cu:702-732): in-column step, then
// bulk absorption at -ln(u)/mu (drawn even when the step was lost).
// true = photon ended.
template <class Rng, bool TRACE, class M = MathDevice>
__device__ __forceinline__ bool isotropic(Rng& rng, Photon& p, const GeomDev& g,
                                          StepSink* sink) {
  const float ds = dist_to_surface<TRACE, M>(p, g, sink);
  float da = 999999.0f;
  if (M::kFaithful && g.bulk > 0.0f) {  // (float)((double)(-1/mu) * log(u)), K:712
    da = fd2x_scaled_log(g.neg_inv_bulk, rng.u());
  } else if (g.bulk > 0.0f) {
    const float u = rng.u();
    // -ln u >= 1 - u, so the absorption length da = -ln(u)/bulk exceeds a
    // step shorter than (1 - u)/bulk less a 2^-20 margin (which covers the
    // logarithm's error and the roundings, checked for every u the kernel
    // draws); only the remaining steps need the logarithm, and the outcome
    // is the same
    if (ds < ((1.0f - u) * -g.neg_inv_bulk) * 0.99999904632568359375f)
      da = 3.0e38f;
    else
      da = g.neg_inv_bulk * fd2_log(u);
  }
  if (fabsf(ds - (-99.0f)) < g.eps_lt) return true;
  if (ds < da && ds >= 0.0f) return false;
  if (ds >= da && ds >= 0.0f) {
    p.fate = kAbsBulk;
    return true;
  }
  return false;
}

// detection (main/kernel_cuda_c_ver1_0.cu:3372-3427): on the sensor plane,
// specular reflection with probability sensorRefl, else detected.  The
// reference follows the reflection with another in-column step inside the
// same pass; here the pass returns kSensorReflect and the caller runs that
// step as the start of the next pass (hot_step), so the step's code exists
// once.
enum SensorResult : int { kSensorNone = 0, kSensorDetect = 1, kSensorReflect = 2 };

template <class Rng, class M = MathDevice>
__device__ __forceinline__ int detection(Rng& rng, Photon& p, const GeomDev& g) {
  if (fabsf(p.pz - (-g.half_h)) < g.eps_lt) {
    if (rng.u() < g.sensor_refl) {
      p.nx = 0.0f;
      p.ny = 0.0f;
      p.nz = 1.0f;
      reflect<M>(p, g);
      return kSensorReflect;
    }
    tally_detect(p, g);
    return kSensorDetect;
  }
  return kSensorNone;
}

// fd2_acos(x) < 0 || fd2_acos(x) > 1.57 without branches.  The arc cosine
// crosses 1.57 at x = cos(1.57) = 7.9633e-4; over all floats x in [-1, 1],
// fd2_acos(x) >= 0, x <= 7.90e-4 gives an angle above 1.57 and x >= 8.03e-4
// one at or below it, and between the two fd2_acos is its |x| <= 1/2 branch
// (fd2_acos_small).  |x| > 1 and NaN give a NaN angle, which is neither.
__device__ __forceinline__ bool angle_over_157(float x) {
  const float near = fd2_acos_small(x);
  return (x >= -1.0f && x <= 7.90e-4f) || (x > 7.90e-4f && x < 8.03e-4f && near > 1.57f);
}

// acos(x) < 0 || acos(x) > 1.57 under either arithmetic (MathFaithful: the C
// library's acosf, K:2531, :1904)
template <class M>
__device__ __forceinline__ bool angle_over(float x) {
  if constexpr (M::kFaithful) {
    const float a = fd2x_acosf(x);
    return a < 0.0f || a > 1.57f;
  } else {
    return angle_over_157(x);
  }
}

// New column orientation (transmit/boundary_analysis/refl_bottom,
// main/kernel_cuda_c_ver1_0.cu:2452-2488, 2501-2538, 1769-1804, 3036-3070):
// n = -d perturbed with beta = 1, projected on xy and renormalised, then
// perturbed with beta (always for refl_bottom); accept when
// acos(-d.n) <= 1.57.  One sample; true = accepted (p.n is the sample
// either way).
template <class Rng, class M = MathDevice>
__device__ __forceinline__ bool column_normal_sample(Rng& rng, Photon& p, const GeomDev& g,
                                                     float mx, float my, float mz,
                                                     bool always_rough) {
  float nx, ny, nz = 0.0f;
  if constexpr (M::kFaithful) {
    // RoughSurface(beta = 1) of -d, then the xy projection renormalised
    nx = mx;
    ny = my;
    float tz = mz;
    rough<Rng, M>(rng, nx, ny, tz, 1.0f);
    const float den = sqrtf(nx * nx + ny * ny);
    nx = nx / den;
    ny = ny / den;
  } else {
    // -d perturbed with beta = 1 and projected on xy: the xy part of p + m
    // is normalised directly (the same direction as normalising p + m first)
    float qx, qy, qz;
    rough_offset(rng, qx, qy, qz);
    nx = qx + mx;
    ny = qy + my;
    const float inv = fd2_rsqrt(nx * nx + ny * ny);
    nx = nx * inv;
    ny = ny * inv;
  }
  if (always_rough || g.beta > 0.0f) rough<Rng, M>(rng, nx, ny, nz, g.beta);
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  return !angle_over<M>(mx * nx + my * ny + mz * nz);
}

// Samples until accepted, while the caller's counter < 100.  true = photon
// killed (theta1).
template <class Rng, class M = MathDevice>
__device__ __forceinline__ bool column_normal(Rng& rng, Photon& p,
                                              const GeomDev& g, float mx,
                                              float my, float mz, int& tries,
                                              bool always_rough) {
  while (!column_normal_sample<Rng, M>(rng, p, g, mx, my, mz, always_rough)) {
    if (tries >= 100) {
      p.fate = kTheta1;
      return true;
    }
    tries++;
  }
  return false;
}

__device__ __forceinline__ bool at_top_up(const Photon& p, const GeomDev& g) {
  return fabsf(p.pz - g.half_h) < g.eps_lt && p.uz > 0.0f;
}

__device__ __forceinline__ bool outside_xy(const Photon& p, const GeomDev& g) {
  return p.px < g.eps_lt || p.px > g.xdet || p.py < g.eps_lt || p.py > g.ydet;
}

// refl_bottom (main/kernel_cuda_c_ver1_0.cu:2985-3087).  true = ended.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ bool refl_bottom(Rng& rng, Photon& p,
                                            const GeomDev& g) {
  const float tx = p.px, ty = p.py, tz = p.pz;
  const float mx = -p.ux, my = -p.uy, mz = -p.uz;
  const float d = rng.u() * g.dspan + g.d_min;
  const float dtop = (g.half_h - tz) / p.uz;
  p.px = tx + p.ux * d;
  p.py = ty + p.uy * d;
  p.pz = tz + p.uz * d;
  if (outside_xy(p, g)) {
    p.fate = kLost;
    return true;
  }
  if (p.pz > g.half_h) {
    if (dtop < d && dtop > g.eps_gt) {
      p.px = tx + p.ux * dtop;
      p.py = ty + p.uy * dtop;
      p.pz = g.half_h;
      return false;
    }
    p.fate = kLost;
    return true;
  }
  int tries = 0;
  return column_normal<Rng, M>(rng, p, g, mx, my, mz, tries, true);
}

// sensor plane reached between columns: reflect + refl_bottom, or detect.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ bool gap_sensor(Rng& rng, Photon& p,
                                           const GeomDev& g) {
  if (rng.u() < g.sensor_refl) {
    p.nx = 0.0f;
    p.ny = 0.0f;
    p.nz = 1.0f;
    reflect<M>(p, g);
    return refl_bottom<Rng, M>(rng, p, g);
  }
  tally_detect(p, g);
  return true;
}

// transmit with flagCCT = 0 (main/kernel_cuda_c_ver1_0.cu:2421-2865):
// flight of U(dmin,dmax) between columns against the detector's top
// (reflector/absorber), bottom (sensor) and sides, then the next column's
// orientation.  true = ended.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ bool transmit_gap(Rng& rng, Photon& p,
                                          const GeomDev& g) {
  const float H2 = g.half_h;
  float tx = p.px, ty = p.py, tz = p.pz;
  float mx = -p.ux, my = -p.uy, mz = -p.uz;
  int tries1 = 0, tries2 = 0;
  float d = rng.u() * g.dspan + g.d_min;
  p.px = tx + p.ux * d;
  p.py = ty + p.uy * d;
  p.pz = tz + p.uz * d;
  float dtop = (H2 - tz) / p.uz;
  float dbot = (-H2 - tz) / p.uz;
  if (outside_xy(p, g)) {
    p.fate = kLost;
    return true;
  }
  if (p.pz < -H2 || p.pz > H2) {
    if (dtop < d && dtop > g.eps_gt) {
      p.px = tx + p.ux * dtop;
      p.py = ty + p.uy * dtop;
      p.pz = H2;
    } else if (dbot < d && dbot > g.eps_gt) {
      p.px = tx + p.ux * dbot;
      p.py = ty + p.uy * dbot;
      p.pz = -H2;
      if (gap_sensor<Rng, M>(rng, p, g)) return true;
      if (at_top_up(p, g)) goto top;
    } else {
      p.fate = kLost;
      return true;
    }
  }
  if (column_normal<Rng, M>(rng, p, g, mx, my, mz, tries1, false)) return true;
  if (fabsf(p.px) < g.eps_lt || fabsf(p.px - g.xdet) < g.eps_lt ||
      fabsf(p.py) < g.eps_lt || fabsf(p.py - g.ydet) < g.eps_lt) {
    p.fate = kLost;
    return true;
  }
top:
  if (++p.loops > kMaxLoops) {
    p.fate = kLost;
    return true;
  }
  if (at_top_up(p, g)) {
    p.nx = 0.0f;
    p.ny = 0.0f;
    p.nz = -1.0f;
    if (g.top_absfrac > 0.0f && rng.u() < g.top_absfrac) {
      p.fate = kAbsTop;
      return true;
    }
    p.uz = -fabsf(rng.u() * 2.0f - 1.0f);
    azimuth<M>(fd2_sqrt_unit(1.0f - p.uz * p.uz), rng.u(), p.ux, p.uy);
    tx = p.px;
    ty = p.py;
    tz = p.pz;
    mx = -p.ux;
    my = -p.uy;
    mz = -p.uz;
    d = rng.u() * g.dspan + g.d_min;
    dbot = (-H2 - tz) / p.uz;
    p.px = tx + p.ux * d;
    p.py = ty + p.uy * d;
    p.pz = tz + p.uz * d;
    if (outside_xy(p, g)) {
      p.fate = kLost;
      return true;
    }
    if (p.pz < -H2 || p.pz > H2) {
      if (dbot < d && dbot > g.eps_gt) {
        p.px = tx + p.ux * dbot;
        p.py = ty + p.uy * dbot;
        p.pz = -H2;
        if (gap_sensor<Rng, M>(rng, p, g)) return true;
        if (at_top_up(p, g)) goto top;
      } else {
        p.fate = kLost;
        return true;
      }
    }
    if (column_normal<Rng, M>(rng, p, g, mx, my, mz, tries2, false)) return true;
  }
  if (fabsf(p.pz - (-H2)) < g.eps_lt) {
    if (gap_sensor<Rng, M>(rng, p, g)) return true;
    if (at_top_up(p, g)) goto top;
  }
  return false;
}

// boundary_analysis (main/kernel_cuda_c_ver1_0.cu:1582-1958), split up.
// ba_entry is the frequent part -- the top reflector and the crosstalk
// decision (K:1610-1652); it returns kBaAlive, kBaEnded, kBaCross when
// crosstalk was drawn and the neighbour column's orientation follows
// (K:1655-1673, sampled by hot_step), or kBaPrpt when the wall's Fresnel
// branch must follow (K:1677-1955, ba_prpt below).  Running either later
// for a lane changes nothing for that photon: all its state and its random
// stream go with it.
enum BaResult : int { kBaAlive = 0, kBaEnded = 1, kBaPrpt = 2, kBaCross = 3 };

template <class Rng, class M = MathDevice>
__device__ __forceinline__ int ba_entry(Rng& rng, Photon& p, const GeomDev& g) {
  const float H2 = g.half_h;
  // top reflector reached inside a column (K:1610-1634)
  if (fabsf(p.pz - H2) < g.eps_lt && p.uz > 0.0f) {
    if (g.top_absfrac > 0.0f && rng.u() < g.top_absfrac) {
      p.fate = kAbsTop;
      return kBaEnded;
    }
    p.nx = 0.0f;
    p.ny = 0.0f;
    p.nz = -1.0f;
    p.uz = -fabsf(rng.u() * 2.0f - 1.0f);
    azimuth<M>(fd2_sqrt_unit(1.0f - p.uz * p.uz), rng.u(), p.ux, p.uy);
    return kBaAlive;
  }
  // crosstalk probability, piecewise linear in depth (K:1639-1652)
  const float top20 = H2 - g.depth20;
  float cct = 0.0f;
  if (p.pz <= H2 && p.pz >= top20) {
    cct = 1.0f;
  } else {
    // one division for both linear pieces (z in [0, top20) or [-H/2, 0))
    const bool upper = p.pz < top20 && p.pz >= 0.0f;
    if (upper || (p.pz < 0.0f && p.pz >= -H2)) {
      const float num = upper ? p.pz : p.pz - (-H2);
      const float den = upper ? 2.0f * top20 : 2.0f * (-H2);
      cct = num / den + (upper ? 0.5f : 1.0f);
    }
  }
  if (rng.u() < cct) return kBaCross;
  return kBaPrpt;
}

// Fresnel reflection/refraction at a column wall and the flights between
// columns it leads to (K:1677-1955).
//
// The reference runs this as one goto loop whose locals persist across its
// passes: the resample/retry counters, the stale Pr/Pt of the gap side
// without Fresnel recomputation, and the stale theta2 that a refraction
// then uses.  The Fresnel branch (K:1811-1871) is evaluated from cos(theta1)
// (the dot product the reference takes acos of) and Snell's law without the
// angles themselves: sin(theta1) = sqrt(1 - cos^2), sin(theta2) =
// (n_from/n_to) sin(theta1), cos(theta2) = sqrt(1 - sin^2), Pr =
// fd2_fresnel_r (Pr + Pt = 1, so no normalisation); theta2 persists as its
// sine and cosine.  PrptState holds them so that, with stop_at_gap, the loop stops
// before each flight through the gap (after the refraction out of a column,
// and after a reflection off the outside of the next column) and returns
// kPrGap; a later call with inside = false continues with that flight.  The
// photon's arithmetic and draws are the same whether or not it stops there.
struct PrptState {
  uint32_t tries;  // theta, smooth, reperturb, normal counters, 8 bits each
  float pr;        // reflection probability
  float s2, c2;    // sin, cos of theta2
};
// the reference starts with Pr = 0 and theta2 = 0
__device__ __forceinline__ PrptState prpt_initial() { return PrptState{0u, 0.0f, 0.0f, 1.0f}; }
enum PrptResult : int { kPrAlive = 0, kPrEnded = 1, kPrGap = 2 };

//
// MathFaithful keeps the reference's angles instead: theta1 = acosf of the
// dot product, theta2 from asinf, Pr and Pt from the angle sums and
// differences, normalised by their sum on every pass (stale values
// included, as the reference does).  Its state is local: that arithmetic
// runs only in the whole-branch form (ba_prpt), never with stop_at_gap.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ int prpt_loop(Rng& rng, Photon& p, const GeomDev& g, PrptState& st,
                                         bool inside, bool stop_at_gap) {
  const float R = g.R;
  int theta_tries = (int)(st.tries & 0xffu), smooth_tries = (int)((st.tries >> 8) & 0xffu);
  int reperturb = (int)((st.tries >> 16) & 0xffu), normal_tries = (int)(st.tries >> 24);
  float pr = st.pr, s2 = st.s2, c2 = st.c2, s1 = 0.0f;
  float pt = 0.0f, t1 = 0.0f, t2 = 0.0f;  // MathFaithful: Pt, theta1, theta2
  float onx, ony, onz, odx, ody, odz;
  float mx, my, mz;

  if (!inside) {
    if (transmit_gap<Rng, M>(rng, p, g)) return kPrEnded;
  }

prpt:
  if (++p.loops > kMaxLoops) {
    p.fate = kLost;
    return kPrEnded;
  }
  if (inside) {
    const float ax = p.xc - p.px, ay = p.yc - p.py;
    if constexpr (M::kFaithful) {  // K:1686-1690
      const float mag = sqrtf(ax * ax + ay * ay);
      p.nx = ax / mag;
      p.ny = ay / mag;
    } else {
      const float inv = fd2_rsqrt(ax * ax + ay * ay);
      p.nx = ax * inv;
      p.ny = ay * inv;
    }
    p.nz = 0.0f;
    if (g.beta > 0.0f) rough<Rng, M>(rng, p.nx, p.ny, p.nz, g.beta);
  } else {
    p.xc = p.px + R * (-p.nx);
    p.yc = p.py + R * (-p.ny);
  }
  onx = p.nx; ony = p.ny; onz = p.nz;
  odx = p.ux; ody = p.uy; odz = p.uz;

reperturb_point:
  p.nx = onx; p.ny = ony; p.nz = onz;
  p.ux = odx; p.uy = ody; p.uz = odz;
  mx = -p.ux; my = -p.uy; mz = -p.uz;
  if (reperturb != 0) {
    if (g.beta > 0.0f) rough<Rng, M>(rng, p.nx, p.ny, p.nz, g.beta);
    if (!inside) {
      p.xc = p.px + R * (-p.nx);
      p.yc = p.py + R * (-p.ny);
    }
  }

no_perturbation:
  {
    float c1 = mx * p.nx + my * p.ny + mz * p.nz;  // cos(theta1)
    if (c1 > 1.0f || c1 < 0.0f) {
      if (!inside) {
        if (column_normal<Rng, M>(rng, p, g, mx, my, mz, normal_tries, false))
          return kPrEnded;
      }
      if (theta_tries < 100) {
        theta_tries++;
        goto prpt;
      }
      p.fate = kTheta1;
      return kPrEnded;
    }
    if constexpr (M::kFaithful) {
      // K:1757-1871: theta1 = acosf; TIR against asin of the index ratio;
      // theta1 < eps becomes 0.00042; theta2 = asinf((float)(ratio *
      // sin(theta1))); Pr and Pt from the angles; then both normalised
      t1 = fd2x_acosf(c1);
      bool fresnel = false;
      float snell = 0.0f;
      if (inside) {
        if (t1 > g.t1_tir_in) {
          pr = 1.0f;
          pt = 0.0f;
        } else {
          if (t1 < g.eps_lt) t1 = 0.00042f;
          snell = g.ratio12;
          fresnel = true;
        }
      } else if (!g.gap_fresnel) {
        if (t1 > g.t1_tir_gap) {
          pr = 1.0f;
          pt = 0.0f;
        }
      } else {
        snell = t1 < g.eps_lt ? g.ratio12 : g.ratio21;
        if (t1 < g.eps_lt) t1 = 0.00042f;
        fresnel = true;
      }
      if (fresnel) {
        t2 = fd2x_asinf(fd2x_scaled_sin(snell, t1));
        fd2x_fresnel(t1, t2, &pr, &pt);
      }
      const float tot = pr + pt;
      pr = pr / tot;
      pt = pt / tot;
      goto fresnel_draw;
    }
    s1 = fd2_sqrt_unit(1.0f - c1 * c1);
    // theta1 < eps (c1 == 1) is replaced by 0.00042 rad where the reference
    // computes Pr from the angles
    float snell = 0.0f;  // n_from / n_to; 0: Pr is not recomputed
    if (inside) {
      if (c1 < g.cos_tir_in)
        pr = 1.0f;
      else
        snell = g.ratio12;
    } else if (!g.gap_fresnel) {
      if (c1 < g.cos_tir_gap) pr = 1.0f;
    } else {
      snell = c1 >= 1.0f ? g.ratio12 : g.ratio21;
    }
    if (snell != 0.0f) {
      if (c1 >= 1.0f) {
        c1 = FD2_COS_00042;
        s1 = FD2_SIN_00042;
      }
      s2 = snell * s1;
      c2 = fd2_sqrt_unit(fmaxf(1.0f - s2 * s2, 0.0f));
      pr = fd2_fresnel_r(c1, s1, c2, s2);
    }
  }

fresnel_draw:
  if (rng.u() < pr) {
    reflect<M>(p, g);
    if (angle_over<M>(onx * p.ux + ony * p.uy + onz * p.uz)) {
      reperturb++;
      if (reperturb < 4) goto reperturb_point;
      p.nx = onx; p.ny = ony; p.nz = onz;
      p.ux = odx; p.uy = ody; p.uz = odz;
      mx = -p.ux; my = -p.uy; mz = -p.uz;
      reperturb = 0;
      if (smooth_tries < 25) {
        smooth_tries++;
        goto no_perturbation;
      }
      p.fate = kTheta1;
      return kPrEnded;
    }
    if (!inside) goto gap_flight;
    return kPrAlive;
  }
  if constexpr (M::kFaithful)
    refract_angles(p, t1, t2, g);
  else
    refract(p, s1, s2, c2, g);
  if (!inside) return kPrAlive;
  inside = false;
gap_flight:
  if (stop_at_gap) {
    st.tries = (uint32_t)theta_tries | ((uint32_t)smooth_tries << 8) |
               ((uint32_t)reperturb << 16) | ((uint32_t)normal_tries << 24);
    st.pr = pr;
    st.s2 = s2;
    st.c2 = c2;
    return kPrGap;
  }
  if (transmit_gap<Rng, M>(rng, p, g)) return kPrEnded;
  goto prpt;
}

// The whole branch in one call.  true = ended.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ bool ba_prpt(Rng& rng, Photon& p, const GeomDev& g) {
  PrptState st = prpt_initial();
  return prpt_loop<Rng, M>(rng, p, g, st, true, false) == kPrEnded;
}

// One algo() pass of the reference (main/kernel_cuda_c_ver1_0.cu:553-655)
// fused with the boundary interaction that follows it: in-column step and
// sensor check, then (if alive and not trapped) boundary_analysis up to its
// Fresnel branch; kNeedPrpt asks the caller to run ba_prpt.
//
// Crosstalk moves the photon to a neighbour column whose orientation is
// sampled from -d until a sample passes the angle test, at most 101 samples
// (K:2452-2488).  A pass takes one sample: a rejected photon is left in
// state kNewColumn and its next pass samples again instead of stepping, so
// its draws and arithmetic are the same, and lanes that need one sample do
// not wait in a loop for lanes that need several.
enum StepResult : int { kAlive = 0, kEnded = 1, kTrapped = 2, kNeedPrpt = 3 };

template <class Rng, bool TRACE, class M = MathDevice>
__device__ __forceinline__ int hot_step(Rng& rng, Photon& p, const GeomDev& g,
                                        StepSink* sink) {
  if (p.fate != kNewColumn) {
    // a pass resumed after a sensor reflection is the rest of the previous
    // pass: no loop count, no second sensor check
    const bool resumed = p.fate == kResume;
    if (resumed) {
      p.fate = kNone;
    } else if (++p.loops > kMaxLoops) {
      p.fate = kLost;
      return kEnded;
    }
    if (isotropic<Rng, TRACE, M>(rng, p, g, sink)) return kEnded;
    if (!resumed) {
      const int d = detection<Rng, M>(rng, p, g);
      if (d == kSensorDetect) return kEnded;
      if (d == kSensorReflect) {
        p.fate = kResume;
        return kAlive;
      }
    }
    if (fabsf(p.uz) < g.eps_lt) return kTrapped;
    const int b = ba_entry<Rng, M>(rng, p, g);
    if (b == kBaEnded) return kEnded;
    if (b == kBaPrpt) return kNeedPrpt;
    if (b == kBaAlive) return kAlive;
    p.pix = 0;  // crosstalk: no sample rejected yet
  }
  // neighbour column with a random orientation, direction kept
  const int rejected = p.pix;
  p.pix = -1;
  p.fate = kNone;
  if (column_normal_sample<Rng, M>(rng, p, g, -p.ux, -p.uy, -p.uz, false)) {
    p.xc = p.px + g.R * (-p.nx);
    p.yc = p.py + g.R * (-p.ny);
    return kAlive;
  }
  if (rejected >= 100) {
    p.fate = kTheta1;
    return kEnded;
  }
  p.fate = kNewColumn;
  p.pix = rejected + 1;
  return kAlive;
}

// the whole pass, for kernels that do not defer the Fresnel branch
template <class Rng, bool TRACE, class M = MathDevice>
__device__ __forceinline__ int photon_step(Rng& rng, Photon& p,
                                           const GeomDev& g, StepSink* sink) {
  const int r = hot_step<Rng, TRACE, M>(rng, p, g, sink);
  if (r != kNeedPrpt) return r;
  return ba_prpt<Rng, M>(rng, p, g) ? kEnded : kAlive;
}

// trapped photon (|dcos.z| < eps): restart from the event with a new
// direction; the column centre and rebound count are kept (C branch,
// main/kernel_cuda_c_ver1_0.cu:609-644).  The normal drawn there is never
// read before being overwritten, so only its two draws are consumed.
template <class Rng, class M = MathDevice>
__device__ __forceinline__ void restart_trapped(Rng& rng, Photon& p,
                                                const GeomDev& g, const fd2_event& ev) {
  emit_dir<Rng, M>(rng, p, g);
  p.px = ev.x;
  p.py = ev.y;
  p.pz = ev.z;
  if (g.beta > 0.0f) {
    (void)rng.u();
    (void)rng.u();
  }
}

}  // namespace fd2
