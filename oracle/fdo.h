/*
 * fdo.h -- CPU oracle for fastDETECT2 optical-photon transport.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * A plain-C restatement of the reference C path:
 *   transport  main/kernel_cuda_c_ver1_0.cu, C branches (#else of USING_CUDA)
 *   driver     main/hybridMANTIS_c_ver1_0.c:191-340 (event loop, first photon)
 *   PHS rule   main/hybridMANTIS_c_ver1_0.c:343-366
 * It reproduces the C path's mixed precision: float state, double <math.h>
 * calls (sqrt, sin, cos, tan, pow, log, asin) rounded back to float on
 * assignment, float acosf/asinf.  Each function cites the lines it follows.
 *
 * Two random-number modes:
 *   FDO_RNG_RANECU  one RANECU stream per event, seeded like the C driver
 *                   (init_PRNG(event_index, 50000, seed_input)); the photons
 *                   of an event consume it in order.
 *   FDO_RNG_COUNTER one stream per photon: xoshiro128+ whose 128-bit
 *                   state is Philox4x32-10(counter = (photon, event id
 *                   >> 32, 0, 0), key = (seed, event id mod 2^32)) -- the
 *                   production kernel's stream layout, independent of
 *                   scheduling, distinct for all 64-bit event ids.
 */
#ifndef FDO_H
#define FDO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FDO_RNG_RANECU = 0, FDO_RNG_COUNTER = 1 };

/* Two arithmetic modes over the same control flow:
 *   FDO_MATH_FAITHFUL  the C path's own mixed float/double evaluation with
 *                      the C library (bit-exact against the reference);
 *   FDO_MATH_DEVICE    every expression in FP32 with the elementary
 *                      functions of hybridmantis_amd/csrc/fd2_math.h, i.e.
 *                      the GPU kernel's arithmetic restated on the CPU.
 *                      Where the kernel computes a quantity another way
 *                      that is equal in exact arithmetic, so does this
 *                      mode: azimuths as turns, reciprocal square roots
 *                      for normalisations, the new-column normal's xy
 *                      projection taken before normalising, and the
 *                      Fresnel branch from the sines and cosines of the
 *                      angles (Snell's law) instead of the angles;
 *   FDO_MATH_EMULATED  FAITHFUL with each double expression evaluated by
 *                      hybridmantis_amd/csrc/fd2_ff.h (float pairs, no
 *                      double) and the C library's acosf/asinf by their
 *                      restatements there: the GPU's faithful trace mode
 *                      restated on the CPU. */
enum { FDO_MATH_FAITHFUL = 0, FDO_MATH_DEVICE = 1, FDO_MATH_EMULATED = 2 };

/* photon fates (exactly one per transported photon) */
enum {
  FDO_FATE_NONE = 0,
  FDO_FATE_DETECT = 1,      /* detected inside the PRF bounds            */
  FDO_FATE_DETECT_OUT = 2,  /* detected outside the PRF bounds           */
  FDO_FATE_ABS_TOP = 3,
  FDO_FATE_ABS_BULK = 4,
  FDO_FATE_LOST = 5,
  FDO_FATE_OUTOFCOL = 6,
  FDO_FATE_THETA1 = 7
};

/* counters, same order as the reference's optstats_ block plus extras */
enum {
  FDO_C_GENERATED = 0, FDO_C_DETECT, FDO_C_ABS_TOP, FDO_C_ABS_BULK,
  FDO_C_LOST, FDO_C_OUTOFCOL, FDO_C_THETA1, FDO_C_REGEN, FDO_NCOUNT
};

/* work counters for the FLOP model (one per primitive with a fixed cost) */
enum {
  FDO_W_DRAW = 0,    /* uniform draws                                  */
  FDO_W_INIT,        /* isotropic direction inits (emission / regen)    */
  FDO_W_DIST,        /* dist_to_surface calls                           */
  FDO_W_NUDGE,       /* out-of-column nudge iterations                  */
  FDO_W_ABSLOG,      /* bulk-absorption log draws                       */
  FDO_W_SENSOR,      /* sensor-plane decisions (detect / reflect)       */
  FDO_W_TOP,         /* top-surface reflector decisions                 */
  FDO_W_CCT,         /* crosstalk decisions                             */
  FDO_W_WALLNORM,    /* in-column wall normal computations (prpt)       */
  FDO_W_THETA,       /* incidence-angle evaluations                     */
  FDO_W_FRESNEL,     /* Fresnel Pr/Pt evaluations                       */
  FDO_W_REFLECT,     /* specular reflections (trans_dir_cos flag 0)     */
  FDO_W_REFRACT,     /* refractions (trans_dir_cos flag 1)              */
  FDO_W_ROUGH,       /* RoughSurface calls                              */
  FDO_W_NEWNORMAL,   /* new-column normal samples (xy projection+acos)  */
  FDO_W_FLIGHT,      /* inter-column flights                            */
  FDO_W_REFLCHK,     /* reflected-vs-smooth-normal angle checks         */
  FDO_NWORK
};

typedef struct {
  /* float copies of the 17 doubles of /inputargs/ that the transport uses */
  float xdet, ydet, H, R, n_C, n_IC, top_absfrac, bulk, beta, d_min, d_max;
  float lbx, lby, ubx, uby, sensor_refl;
  int pixel;          /* PRF pitch (um), an int as in the reference       */
  int xdim, ydim;     /* ceil((ub-lb)/pixel)                               */
} fdo_geom;

typedef struct {      /* one energy-deposition event, positions in um     */
  float x, y, z;
  float E;            /* eV (unused by the transport)                      */
  int32_t hist;       /* 1-based x-ray history number                      */
  int32_t N;          /* optical photons to transport                      */
  float col_dx;       /* event x - first column axis x (0 in the reference) */
} fdo_event;

typedef struct {      /* per-photon record (optional output)               */
  int32_t fate;
  int32_t pix;        /* i*ydim + j for FDO_FATE_DETECT, else -1           */
  float x, y;         /* detection point (FDO_FATE_DETECT[_OUT])           */
} fdo_photon;

typedef struct {
  uint64_t *prf;        /* xdim*ydim, [i][j] row-major, accumulated (+=)   */
  uint32_t *hist_count; /* num_hist, detections per history, accumulated   */
  int32_t num_hist;
  uint64_t counters[FDO_NCOUNT];
  uint64_t work[FDO_NWORK];
  fdo_photon *photons;  /* optional: Σ N records, event-major order        */
} fdo_tally;

/* Transport events [0, n).  event_base is the global id of events[0] (used
 * as the Philox key -- low word in the key, high word in counter word 1 --
 * and as the RANECU history_batch index). */
int fdo_run(const fdo_geom *g, const fdo_event *events, int64_t n,
            int64_t event_base, int rng_mode, int math_mode, uint32_t seed,
            fdo_tally *t);

/* Timing only: repeat the reference C driver's two per-event clears of an
 * unused xdim x ydim u64 image (main/hybridMANTIS_c_ver1_0.c:205-207,
 * 248-250) in fdo_run (process-wide setting, off by default). */
void fdo_set_image_rezero(int on);

/* Restated RANECU pieces, exported for known-answer tests. */
float fdo_ranecu(int32_t seed[2]);
void fdo_init_prng(int history_batch, int histories_per_thread, int seed_input,
                   int32_t seed[2]);
int32_t fdo_abmodm(int32_t m, int32_t a, int32_t s);
void fdo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* PHS binning rule of hybridMANTIS_c_ver1_0.c:343-366 (zero-detection
 * histories excluded); hist accumulates into phs[num_bins].  Returns 0, or
 * -1 when the bin size rounds to 0. */
int fdo_phs(const uint32_t *hist_count, int32_t num_hist, int min_ph,
            int max_ph, int num_bins, int32_t *phs);

/* FP64 re-trace of one in-column step (dist_to_surface) from a float state:
 * the position the step reaches, computed in double.  Returns the reference
 * code: 0 moved, 1 lost, 2 out of column.  forced >= 0 replays recorded
 * branch decisions (fd2_step.decisions layout; the code is then 0). */
int fdo_retrace_step_f64(const float pos_in[3], const float dcos[3], float xc,
                         float yc, const fdo_geom *g, int64_t forced,
                         double pos_out[3]);

/* Test hooks (closed-form and distribution checks). */
void fdo_counter_stream(uint32_t seed, uint64_t event, uint32_t photon,
                        int64_t n, float *out);
int fdo_math_eval(int fn, const float *x, float *y, int64_t n);
void fdo_fresnel_probs(float t1, float t2, int math_mode, float *pr, float *pt);
void fdo_emit_dirs(uint32_t seed, int math_mode, int64_t n, float *out3);
void fdo_rough_normals(uint32_t seed, int math_mode, float beta,
                       const float n0[3], int64_t n, float *out3);

#ifdef __cplusplus
}
#endif
#endif
