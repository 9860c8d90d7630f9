// fd2_fortran.cpp -- the reference's gfortran-called optical entry points
// over the fd2 library (host only; built into libfd2_fortran.so).
//
//   gpuoptical_  main/hybridMANTIS_cuda_ver1_0.cu:107-447 (99/100/101 protocol)
//   cpuoptical_  main/hybridMANTIS_c_ver1_0.c:81-374
//   gpuoptlb_    main/hybridMANTIS_cuda_ver1_0_LB.cu:108-280
//   cpuoptlb_    main/hybridMANTIS_c_ver1_0_LB.c:86-290
//
// Event conversion follows the drivers: position = buffer (cm) * 1e4 in
// double, rounded to float (hybridMANTIS_cuda_ver1_0.cu:254-257); photon
// count N ~ Poisson(E * yield) with the yield read as float
// (hybridMANTIS_cuda_ver1_0.cu:262-263, yield declared float at :111).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "fd2.h"
#include "fd2_fortran.h"

// COMMON storage.  Weak so that an executable's own COMMON definition (a
// gfortran main program) is the one both sides use.
extern "C" {
__attribute__((weak)) fd2_params inputargs_;
__attribute__((weak)) struct fd2f_optical optical_;
__attribute__((weak)) struct fd2f_optstats optstats_;
__attribute__((weak)) struct fd2f_outputdetprim outputdetprim_;
__attribute__((weak)) struct fd2f_outputimage outputimage_;
}

namespace {

// ---- Poisson sampling -----------------------------------------------------
// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) streams, one per event,
// started from a mix of (seed, ordinal).
struct Stream {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

Stream event_stream(uint32_t seed, int64_t ordinal) {
  Stream st{((uint64_t)seed << 32) ^ 0x5851F42D4C957F2Dull};
  st.s ^= st.next() + (uint64_t)ordinal * 0xD1B54A32D192ED03ull;
  st.next();
  return st;
}

// Inversion by sequential search for small means; PTRS, the transformed
// rejection with squeeze of W. Hormann, "The transformed rejection method
// for generating Poisson random variables", Insurance: Mathematics and
// Economics 12 (1993) 39-45, for mu >= 10.
int64_t poisson(double mu, Stream& st) {
  if (mu <= 0.0) return 0;
  if (mu < 10.0) {
    double p = std::exp(-mu), f = p;
    const double u = st.uniform();
    int64_t k = 0;
    while (u > f && k < 1000) {
      ++k;
      p *= mu / (double)k;
      f += p;
    }
    return k;
  }
  const double slam = std::sqrt(mu), loglam = std::log(mu);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (;;) {
    const double u = st.uniform() - 0.5;
    const double v = st.uniform();
    const double us = 0.5 - std::fabs(u);
    const double k = std::floor((2.0 * a / us + b) * u + mu + 0.43);
    if (us >= 0.07 && v <= vr) return (int64_t)k;
    if (k < 0.0 || (us < 0.013 && v > us)) continue;
    if (std::log(v) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
        -mu + k * loglam - std::lgamma(k + 1.0))
      return (int64_t)k;
  }
}

// ---- shared state ------------------------------------------------------------
struct Run {
  fd2_ctx* ctx = nullptr;     // gpuoptical_'s context, 99 .. 101
  bool pending = false;       // a 100 submit whose kernel time is not yet read
  fd2_ctx* sync_ctx = nullptr;  // cpuoptical_ / gpuoptlb_ context
  int64_t run_events = 0;     // events of the current gpuoptical_ run
  int64_t cpu_calls = 0, lb_calls = 0;
  std::vector<fd2_event> ev;
};
Run g;

[[noreturn]] void die(const char* where, int rc) {
  std::fprintf(stderr, "\n\n fastDETECT2 (%s): %s (%d). program is aborting !!\n\n", where,
               fd2_strerror(rc), rc);
  std::fflush(stderr);
  std::abort();
}

void check(const char* where, int rc) {
  if (rc != FD2_OK) die(where, rc);
}

uint32_t base_seed() {
  const char* s = std::getenv("FD2_SEED");
  return s && *s ? (uint32_t)std::strtoul(s, nullptr, 0) : 271828182u;
}

int device() {
  const char* s = std::getenv("FD2_DEVICE");
  return s && *s ? std::atoi(s) : 0;
}

fd2_ctx* create(const char* where, uint32_t seed) {
  fd2_ctx* c = nullptr;
  check(where, fd2_create(&inputargs_, 1, device(), &c));
  check(where, fd2_set_seed(c, seed));
  int32_t xd = 0, yd = 0;
  check(where, fd2_prf_dims(c, 0, &xd, &yd));
  if (xd > 501 || yd > 501) {
    std::fprintf(stderr, "\n fastDETECT2 (%s): PRF of %d x %d pixels exceeds the 501 x 501 "
                 "outputimage COMMON block\n", where, xd, yd);
    fd2_destroy(c);
    die(where, FD2_E_INVAL);
  }
  if (inputargs_.mynumbins > 1000) {
    std::fprintf(stderr, "\n fastDETECT2 (%s): %d PHS bins exceed the 1000 of the "
                 "outputdetprim COMMON block\n", where, inputargs_.mynumbins);
    fd2_destroy(c);
    die(where, FD2_E_INVAL);
  }
  return c;
}

// Events [first, first + n) of /optical/ -> g.ev, with photon counts drawn
// from the streams (seed, number + i).
void load(const char* where, int64_t first, int64_t n, uint32_t seed, int64_t number) {
  if (n < 0 || first < 0 || first + n > FD2_OPTICAL_BUFSIZE) die(where, FD2_E_INVAL);
  g.ev.resize((size_t)n);
  const double yield = (double)(float)inputargs_.detyield;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t k = first + i;
    fd2_event& e = g.ev[(size_t)i];
    e.x = (float)(optical_.xbufopt[k] * 10000.0);
    e.y = (float)(optical_.ybufopt[k] * 10000.0);
    e.z = (float)(optical_.zbufopt[k] * 10000.0);
    e.E = (float)optical_.debufopt[k];
    e.hist = optical_.nbufopt[k];
    const double mu = optical_.debufopt[k] * yield;
    if (!(mu >= 0.0) || !std::isfinite(mu)) die(where, FD2_E_INVAL);
    Stream st = event_stream(seed, number + i);
    const int64_t N = poisson(mu, st);
    if (N > FD2_MAX_PHOTONS_PER_EVENT) {
      std::fprintf(stderr, "\n\n str_n exceeds max photons - %lld !! \n", (long long)N);
      die(where, FD2_E_TOO_MANY);
    }
    e.N = (int32_t)N;
    e.geom = 0;
    e.col_dx = 0.0f;
  }
}

// Add a context's PRF, PHS and counters into the COMMON blocks.
void accumulate(const char* where, fd2_ctx* c, double kernel_ms) {
  int32_t xd = 0, yd = 0;
  check(where, fd2_prf_dims(c, 0, &xd, &yd));
  std::vector<uint64_t> prf((size_t)xd * yd);
  check(where, fd2_get_prf(c, 0, prf.data()));
  for (int32_t i = 0; i < xd; ++i)
    for (int32_t j = 0; j < yd; ++j) outputimage_.newimageopt[i][j] += prf[(size_t)i * yd + j];
  const int32_t nb = inputargs_.mynumbins;
  std::vector<int32_t> phs((size_t)(nb > 0 ? nb : 1));
  check(where, fd2_get_phs(c, 0, phs.data()));
  for (int32_t b = 0; b < nb; ++b) outputdetprim_.gldetprimary[b] += phs[(size_t)b];
  uint64_t cnt[FD2_NCOUNT];
  check(where, fd2_get_counters(c, 0, cnt));
  optstats_.glgen += (double)cnt[FD2_C_GENERATED];
  optstats_.gldetect += (double)cnt[FD2_C_DETECT];
  optstats_.glabstop += (double)cnt[FD2_C_ABS_TOP];
  optstats_.glabsbulk += (double)cnt[FD2_C_ABS_BULK];
  optstats_.gllost += (double)cnt[FD2_C_LOST];
  optstats_.gloutofcol += (double)cnt[FD2_C_OUTOFCOL];
  optstats_.gltheta1 += (double)cnt[FD2_C_THETA1];
  optstats_.glgputime += kernel_ms;
}

double take_kernel_ms(fd2_ctx* c) {
  float ms = 0.0f;
  return fd2_last_kernel_ms(c, &ms) == FD2_OK ? (double)ms : 0.0;
}

double g_kernel_ms = 0.0;  // kernel time of the current gpuoptical_ run

// One synchronous pass over events [first, first + n) with a fresh context
// (the parameters may change between calls); returns the seconds from
// staging the events to the tallies added into the COMMON blocks.
double run_sync(const char* where, int64_t first, int64_t n, uint32_t seed) {
  if (g.sync_ctx) fd2_destroy(g.sync_ctx);
  g.sync_ctx = create(where, seed);
  const auto t0 = std::chrono::steady_clock::now();
  load(where, first, n, seed, 0);
  double ms = 0.0;
  if (n > 0) {
    check(where, fd2_submit(g.sync_ctx, g.ev.data(), n));
    check(where, fd2_sync(g.sync_ctx));
    ms = take_kernel_ms(g.sync_ctx);
  }
  accumulate(where, g.sync_ctx, ms);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int fd2f_poisson(const double* mu, int64_t n, uint32_t seed, int64_t first, int32_t* out) {
  if (n < 0 || (n > 0 && (!mu || !out))) return FD2_E_INVAL;
  for (int64_t i = 0; i < n; ++i) {
    if (!(mu[i] >= 0.0) || !std::isfinite(mu[i])) return FD2_E_INVAL;
    Stream st = event_stream(seed, first + i);
    const int64_t k = poisson(mu[i], st);
    out[i] = k > std::numeric_limits<int32_t>::max() ? std::numeric_limits<int32_t>::max()
                                                     : (int32_t)k;
  }
  return FD2_OK;
}

void gpuoptical_(int32_t* penctr, int32_t* myfactGPU) {
  static const char* where = "gpuoptical";
  if (!penctr || !myfactGPU) die(where, FD2_E_INVAL);
  const int32_t mode = *penctr;
  if (mode == 99) {
    if (g.ctx) fd2_destroy(g.ctx);
    g.ctx = create(where, base_seed());
    g.pending = false;
    g.run_events = 0;
    g_kernel_ms = 0.0;
    return;
  }
  if (mode != 100 && mode != 101) die(where, FD2_E_INVAL);
  if (!g.ctx) die(where, FD2_E_STATE);
  const int64_t n = *myfactGPU;
  load(where, 0, n, base_seed(), g.run_events);
  g.run_events += n;
  if (g.pending) {  // the previous buffer's kernel has to finish before staging
    check(where, fd2_sync(g.ctx));
    g_kernel_ms += take_kernel_ms(g.ctx);
    g.pending = false;
  }
  if (n > 0) {
    check(where, fd2_submit(g.ctx, g.ev.data(), n));
    g.pending = true;
  }
  if (mode == 101) {
    check(where, fd2_sync(g.ctx));
    if (g.pending) g_kernel_ms += take_kernel_ms(g.ctx);
    g.pending = false;
    accumulate(where, g.ctx, g_kernel_ms);
    fd2_destroy(g.ctx);
    g.ctx = nullptr;
    g_kernel_ms = 0.0;
  }
}

void cpuoptical_(int32_t* gpusize) {
  static const char* where = "cpuoptical";
  if (!gpusize) die(where, FD2_E_INVAL);
  const int64_t first = *gpusize;
  const int64_t n = (int64_t)optical_.myctropt - first;
  if (n < 0) die(where, FD2_E_INVAL);
  run_sync(where, first, n, base_seed() + 1u + (uint32_t)g.cpu_calls);
  ++g.cpu_calls;
}

void gpuoptlb_(double* gptime, int32_t* lbbuf) {
  static const char* where = "gpuoptlb";
  if (!gptime || !lbbuf) die(where, FD2_E_INVAL);
  *gptime = run_sync(where, 0, *lbbuf, base_seed() - 1u - (uint32_t)g.lb_calls);
  ++g.lb_calls;
}

void cpuoptlb_(double* cptime, int32_t* cpubufsize) {
  (void)cpubufsize;
  if (cptime) *cptime = std::numeric_limits<double>::infinity();
}

}  // extern "C"
