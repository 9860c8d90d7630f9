/*
 * fd2.h -- C ABI of the MI355X fastDETECT2 optical-photon transport library
 * (libfd2.so).  Plain pointers and sizes only.
 *
 * It replaces the reference's optical-transport entry points:
 *   gpuoptical_(int *penctr, int *myfactGPU)  main/hybridMANTIS_cuda_ver1_0.cu:107
 *   cpuoptical_(int *gpusize)                 main/hybridMANTIS_c_ver1_0.c:81
 *   gpuoptlb_(double *gptime, int *lbbuf)     main/hybridMANTIS_cuda_ver1_0_LB.cu:108
 *   cpuoptlb_(double *cptime, int *cpubufsize) main/hybridMANTIS_c_ver1_0_LB.c:86
 * Those four keep their gfortran ABI (fd2_fortran.h); this header is the
 * native API underneath them.  All functions return FD2_OK (0) or a negative
 * FD2_E* status; none calls exit().
 */
#ifndef FD2_H
#define FD2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FD2_OK 0
#define FD2_E_INVAL (-1)        /* bad argument                           */
#define FD2_E_NOMEM (-2)        /* host or device allocation failed      */
#define FD2_E_DEVICE (-3)       /* HIP runtime error                      */
#define FD2_E_TOO_MANY (-4)     /* event with N > FD2_MAX_PHOTONS_PER_EVENT */
#define FD2_E_NO_DEVICE (-5)    /* no usable GPU                          */
#define FD2_E_STATE (-6)        /* call not valid in the current state    */
#define FD2_E_IO (-7)           /* file could not be read or written      */

/* drv-C/drv-G's max_photon_per_EDE (main/hybridMANTIS_c_ver1_0.c:53) */
#define FD2_MAX_PHOTONS_PER_EVENT 900000
/* largest PRF the reference's COMMON /outputimage/ holds (501x501) is not a
 * limit here; a PRF may have up to this many pixels per side */
#define FD2_MAX_PRF_SIDE 4096
#define FD2_MAX_GEOMS 32
/* log-radial tally of detection offsets from the event: bin 0 = [0, 0.5) um,
 * bins 1..62 log-spaced over [0.5, 800) um, bin 63 = [800, inf) um (the
 * surrogate samples bin 63 over [800, 1000) um) */
#define FD2_RADIAL_BINS 64

/* Detector and tally parameters: the reference's COMMON /inputargs/ block,
 * same members, order and types (main/kernel_cuda_c_ver1_0.cu:178-205,
 * main/hybridMANTIS_tallyEnergyDepositionEvents.f:79-102), so a copy of
 * inputargs_ is a valid fd2_params.  Lengths in micrometres. */
typedef struct fd2_params {
  double detx, dety;        /* detector extent; x in (0,detx), y in (0,dety) */
  double detheight;         /* layer thickness H; z in (-H/2, H/2)          */
  double detradius;         /* column radius R                              */
  double detnC, detnIC;     /* refractive index of column / gap             */
  double dettop;            /* top-surface absorption fraction              */
  double detbulk;           /* bulk absorption coefficient (1/um)           */
  double detbeta;           /* wall roughness coefficient                   */
  double detdmin, detdmax;  /* inter-column flight range                    */
  double detlboundx, detlboundy, detuboundx, detuboundy;  /* PRF bounds      */
  double detyield;          /* light yield (1/eV), used by the shim only    */
  double detsensorRefl;     /* sensor reflectivity (probability)            */
  int32_t detpixel;         /* PRF pitch (um)                               */
  int32_t rungpu;           /* unused by the library                        */
  int32_t machinenum;       /* output file name tag                         */
  int32_t mynumhist;        /* number of x-ray histories (per-history tally) */
  int32_t minphotons, maxphotons, mynumbins;  /* PHS binning                */
} fd2_params;

/* One energy-deposition event (32 B).  Positions in um, E in eV, hist is
 * the 1-based x-ray history, N the optical photons to transport, geom the
 * index of its fd2_params in the context (0 for single-geometry runs).
 * Replaces struct start_info (main/kernel_cuda_c_ver1_0.cu:67-75): the
 * reference holds doubles but converts every use to float. */
typedef struct fd2_event {
  float x, y, z;
  float E;
  int32_t hist;
  int32_t N;
  uint32_t geom;
  /* x offset (um) of the event from the axis of its first column.  The
   * reference always centres the first column on the event (0); non-zero
   * values serve the surrogate's training data only. */
  float col_dx;
} fd2_event;

/* Per-photon record written by the recording entry points. */
typedef struct fd2_photon {
  int32_t fate;             /* FD2_FATE_*                                   */
  int32_t pix;              /* i*ydim + j for FD2_FATE_DETECT, else -1      */
  float x, y;               /* detection point                              */
} fd2_photon;

enum {
  FD2_FATE_NONE = 0, FD2_FATE_DETECT = 1, FD2_FATE_DETECT_OUT = 2,
  FD2_FATE_ABS_TOP = 3, FD2_FATE_ABS_BULK = 4, FD2_FATE_LOST = 5,
  FD2_FATE_OUTOFCOL = 6, FD2_FATE_THETA1 = 7
};

/* counters, in the order of the reference's /optstats/ block
 * (main/kernel_cuda_c_ver1_0.cu:111-121) plus trapped-photon restarts */
enum {
  FD2_C_GENERATED = 0, FD2_C_DETECT, FD2_C_ABS_TOP, FD2_C_ABS_BULK,
  FD2_C_LOST, FD2_C_OUTOFCOL, FD2_C_THETA1, FD2_C_REGEN,
  FD2_C_SURR_UNDETECTED,  /* photons of surrogate events not detected       */
  FD2_NCOUNT
};

/* Launch shape of the persistent transport kernel. 0 = library default. */
typedef struct fd2_tuning {
  int32_t chunk;            /* photons per work-queue item (power of 2)     */
  int32_t waves_per_cu;     /* resident waves per CU (workgroups of 4 waves;
                             * default 20)                                  */
  int32_t lds_tile;         /* side (pixels, <= 112) of the LDS PRF tiles,
                             * centred on each PRF image, of the first
                             * min(n_geoms, 4096 / side^2) geometries;
                             * 0 = default (none: the LDS holds the waves'
                             * wall queues), -1 = no tiles                 */
  int32_t park_lanes;       /* lanes of a wave that wait for the Fresnel /
                             * inter-column phase before it runs (1-64)     */
  int32_t refill_lanes;     /* free lanes of a wave before new photons are
                             * started (1-64)                               */
  int32_t wall_queue;       /* photons waiting for the wall-Fresnel branch:
                             * 1 = move to a per-wave queue (the branch runs
                             * with 64 lanes), -1 = park in their lanes
                             * (park_lanes applies); 0 = default (1)        */
  int64_t max_chunks;       /* work items (chunks) per launch, at most
                             * FD2_MAX_CHUNKS (the 32-bit queue cursor's
                             * range); 0 = default (FD2_MAX_CHUNKS).  A
                             * submit from host memory above it runs as
                             * several launches (same photons); a device
                             * submit or graph above it fails with
                             * FD2_E_INVAL at fd2_sync and tallies nothing  */
} fd2_tuning;
#define FD2_MAX_CHUNKS 4294901760LL /* 2^32 - 2^16 */

typedef struct fd2_ctx fd2_ctx;

/* Create a context on HIP device `device` for n_geoms parameter sets (each
 * gets its own PRF / per-history / counter tallies, all zeroed). */
int fd2_create(const fd2_params *params, int n_geoms, int device,
               fd2_ctx **out);
void fd2_destroy(fd2_ctx *ctx);

/* Seed of the per-photon streams (default 271828182, the reference's
 * seed_input, main/hybridMANTIS_c_ver1_0.c:141). */
int fd2_set_seed(fd2_ctx *ctx, uint32_t seed);
int fd2_set_tuning(fd2_ctx *ctx, const fd2_tuning *tuning);
int fd2_get_tuning(const fd2_ctx *ctx, fd2_tuning *tuning);

/* Transport n events.  fd2_submit copies the host array before returning
 * (the caller may reuse it at once) and returns without waiting for the GPU;
 * fd2_submit_device takes events already in device memory (they must stay
 * valid until fd2_sync).  Events are numbered consecutively across submits
 * since creation/reset; that number keys the photons' random streams, so
 * results do not depend on how the events are split into submits. */
int fd2_submit(fd2_ctx *ctx, const fd2_event *events, int64_t n);
int fd2_submit_device(fd2_ctx *ctx, const fd2_event *d_events, int64_t n);
int fd2_sync(fd2_ctx *ctx);

/* fd2_submit_device without any allocation, for callers that cannot call
 * hipMalloc (code under stream/graph capture, tight serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  The work
 * queue's scratch (cursor, error flag, block totals, per-event chunk and
 * photon offsets) lives in a static device array of the library sized for
 * FD2_NOALLOC_MAX_EVENTS events (12.6 MB per GPU, held until the process
 * ends); the transport runs in park-in-place mode (no per-wave photon
 * queues), which gives the same photons as fd2_submit_device.  The events
 * are in device memory (d_events, valid for the call); the tallies
 * accumulate in the context as for every submit, and the events take the
 * next n event numbers.  Synchronous: returns the launch's status
 * (FD2_E_INVAL for n > FD2_NOALLOC_MAX_EVENTS, an invalid event or more work
 * items than tune.max_chunks).  Every call on one GPU shares that GPU's
 * scratch array: calls that overlap on one GPU (from several host threads
 * or contexts) race and must be serialised by the caller. */
#define FD2_NOALLOC_MAX_EVENTS 1048576
int fd2_submit_device_noalloc(fd2_ctx *ctx, const fd2_event *d_events, int64_t n);

/* Same as fd2_submit but synchronous, and writes one record per photon
 * (Σ N records, events in order, photons in order). */
int fd2_submit_record(fd2_ctx *ctx, const fd2_event *events, int64_t n,
                      fd2_photon *records);

/* Same as fd2_submit but synchronous, and writes for each event the
 * log-radial histogram of its detections' offsets from the event position
 * (FD2_RADIAL_BINS u32 per event, events in order). */
int fd2_submit_radial(fd2_ctx *ctx, const fd2_event *events, int64_t n,
                      uint32_t *radial);

/* Trajectory-test mode: one GPU lane per event, photons in sequence, one
 * RANECU stream per event seeded as the reference C driver does
 * (init_PRNG(event_base + i, 50000, seed_input),
 * main/hybridMANTIS_c_ver1_0.c:237-238).  Synchronous; tallies accumulate
 * like fd2_submit; records may be NULL. */
int fd2_trace_ranecu(fd2_ctx *ctx, const fd2_event *events, int64_t n,
                     int64_t event_base, int32_t seed_input,
                     fd2_photon *records);

/* The same trajectory kernel in the reference C path's own arithmetic
 * (float state, each double <math.h> expression evaluated with pairs of
 * floats and rounded to float where the reference rounds it, the C
 * library's acosf/asinf restated; no FP64 on the GPU), so that its photons
 * follow the reference's own trajectories on the same streams
 * (main/kernel_cuda_c_ver1_0.cu C branch, main/hybridMANTIS_c_ver1_0.c:
 * 237-299).  A test path: slow, one lane per event. */
int fd2_trace_ranecu_faithful(fd2_ctx *ctx, const fd2_event *events, int64_t n,
                              int64_t event_base, int32_t seed_input,
                              fd2_photon *records);

/* Debug: the in-column steps (dist_to_surface) of the first `max_photons`
 * photons of events[0], FP32 state before and after each step; returns the
 * number of steps written (<= max_steps) in *n_steps.  Synchronous. */
typedef struct fd2_step {
  float pos_in[3], dcos[3], xc, yc, pos_out[3];
  int32_t code;             /* 0 moved, 1 lost, 2 out of column              */
  int32_t photon;
  /* branch decisions of the step: bits 0-19 out-of-column nudge moves (1)
   * or step halvings (0) in order, bits 20-24 their number, bit 25 the
   * plane z=+-H/2 was nearer than the wall */
  uint32_t decisions;
} fd2_step;
int fd2_debug_steps(fd2_ctx *ctx, const fd2_event *events, int64_t n,
                    int32_t max_photons, fd2_step *steps, int32_t max_steps,
                    int32_t *n_steps);

/* Debug: the transport's device arithmetic on n arguments, y[i] = f(x[i])
 * on GPU `device` (synchronous; no context).  Compared bit for bit with the
 * same functions on the CPU by the tests.  FD2_MATH_SQRT_UNIT is defined for
 * x = +-0, x >= 2^-96, +inf and NaN; FD2_MATH_ANGLE_OVER_157 gives 1 when
 * acos x < 0 or > 1.57 (the new-column and reflection angle tests), else 0.
 * The faithful arithmetic's pieces: FD2_MATH_ACOSF_LIBC / ASINF_LIBC the C
 * library's float arc functions as restated; FF_SIN / FF_COS / FF_LOG the
 * float-pair sin, cos (|x| <= 8) and log (x > 0) rounded to float. */
enum {
  FD2_MATH_SIN = 0, FD2_MATH_COS = 1, FD2_MATH_LOG = 2, FD2_MATH_ASIN = 3,
  FD2_MATH_ACOS = 4, FD2_MATH_SIN_TURNS = 5, FD2_MATH_COS_TURNS = 6,
  FD2_MATH_RSQRT = 7, FD2_MATH_SQRT_UNIT = 8, FD2_MATH_ANGLE_OVER_157 = 9,
  FD2_MATH_ACOSF_LIBC = 10, FD2_MATH_ASINF_LIBC = 11, FD2_MATH_FF_SIN = 12,
  FD2_MATH_FF_COS = 13, FD2_MATH_FF_LOG = 14,
  FD2_MATH_NFN = 15
};
int fd2_debug_math(int fn, const float *x, float *y, int64_t n, int device);

/* Tallies of geometry `geom` (synchronises first).  prf: xdim*ydim u64,
 * [i][j] row-major with i the x pixel; hist: mynumhist u32 detections per
 * history; phs: mynumbins bins with the reference's rule
 * (main/hybridMANTIS_cuda_ver1_0.cu:376-399); counters: FD2_NCOUNT u64. */
int fd2_prf_dims(const fd2_ctx *ctx, int geom, int32_t *xdim, int32_t *ydim);
int fd2_get_prf(fd2_ctx *ctx, int geom, uint64_t *prf);
int fd2_get_hist_counts(fd2_ctx *ctx, int geom, uint32_t *hist);
int fd2_get_phs(fd2_ctx *ctx, int geom, int32_t *phs);
int fd2_get_counters(fd2_ctx *ctx, int geom, uint64_t *counters);
/* zero all tallies and restart event numbering */
int fd2_reset(fd2_ctx *ctx);

/* Number the next submitted event `first` (>= 0; events are numbered
 * consecutively from there).  The number keys each photon's random stream
 * (64 bits, so streams never repeat within a context), so a job split
 * across contexts or GPUs gives the same photons as one context when each
 * part starts at its first global event number.  No counterpart in the
 * reference, whose CUDA branch reuses its streams in every buffer
 * (main/kernel_cuda_c_ver1_0.cu:420-424). */
int fd2_set_next_event(fd2_ctx *ctx, int64_t first);

/* Device memory helpers for callers that keep events resident in HBM
 * (fd2_submit_device) without a GPU runtime of their own. */
int fd2_device_alloc(fd2_ctx *ctx, size_t bytes, void **d_ptr);
int fd2_device_free(fd2_ctx *ctx, void *d_ptr);
int fd2_copy_to_device(fd2_ctx *ctx, void *d_dst, const void *h_src,
                       size_t bytes);

/* Elapsed GPU time (ms) of the last transport launch sequence (queue build
 * + persistent kernel), measured with HIP events on the library's stream. */
int fd2_last_kernel_ms(fd2_ctx *ctx, float *ms);

/* PHS from per-history counts, host side (same rule as fd2_get_phs);
 * accumulates into phs.  FD2_E_INVAL when (max-min)/num_bins rounds to 0. */
int fd2_phs_from_counts(const uint32_t *hist, int32_t num_hist, int32_t min_ph,
                        int32_t max_ph, int32_t num_bins, int32_t *phs);

const char *fd2_strerror(int status);

/* ---- surrogate network ------------------------------------------------- */
/* MLP 5 -> 256 -> 256 -> 256 -> 65 with GELU, BF16 weights and activations
 * on the matrix cores with FP32 accumulation, FP32 master weights, Adam.
 * Features (FD2_MLP_FEATURES per sample): depth below the top surface / H,
 * col_dx / R, H / 500 um, R / 5 um, top reflectivity (1 - dettop).
 * Outputs/targets (FD2_MLP_TARGETS): detection probability, then the
 * FD2_RADIAL_BINS-bin distribution of detection offsets.  No counterpart in
 * the reference (the surrogate is an addition to its transport). */
#define FD2_MLP_FEATURES 5
#define FD2_MLP_TARGETS 65
typedef struct fd2_mlp fd2_mlp;
size_t fd2_mlp_num_params(void);
int fd2_mlp_create(int device, uint32_t seed, fd2_mlp **out);
void fd2_mlp_destroy(fd2_mlp *mlp);
int fd2_mlp_set_params(fd2_mlp *mlp, const float *params, size_t n);
int fd2_mlp_get_params(fd2_mlp *mlp, float *params, size_t n);
/* probabilities for n host samples: out[n][FD2_MLP_TARGETS] */
int fd2_mlp_forward(fd2_mlp *mlp, const float *x, int64_t n, float *out);
/* the network's raw output logits for n host samples, logits[n][65]
 * (detection logit, then the 64 radial logits): layered = 0 runs the fused
 * inference forward (one kernel for the four layers, the path inference
 * and the surrogate transport use), layered = 1 the per-layer GEMMs that
 * training uses; the two give the same bits */
int fd2_mlp_logits(fd2_mlp *mlp, const float *x, int64_t n, float *logits, int layered);
/* gradient of the mean loss over `batch` samples (batch % 16 == 0), no
 * update: grads[fd2_mlp_num_params()], *loss */
int fd2_mlp_gradients(fd2_mlp *mlp, const float *x, const float *y, int batch,
                      float *grads, float *loss);
/* keep a training set resident on the GPU, then run Adam steps on random
 * minibatches of it; losses[steps] and *ms (GPU time) may be NULL */
int fd2_mlp_set_data(fd2_mlp *mlp, const float *x, const float *y, int64_t n);
int fd2_mlp_train(fd2_mlp *mlp, int steps, int batch, float lr, uint32_t seed,
                  float *losses, float *ms);
/* Measurement: the fused inference forward (4 layers in one kernel, BF16
 * MFMA, FP32 accumulate, bias + GELU epilogues) on a device-resident batch
 * of `batch` rows (multiple of 16, the inputs currently in the network's
 * batch buffer), `reps` times; *ms is the mean GPU time of one forward (no
 * host copies).  2 x 148,992 FLOP per row. */
int fd2_mlp_time_forward(fd2_mlp *mlp, int batch, int reps, float *ms);

/* Surrogate transport with plain-MC fallback.  Each event goes to the plain
 * Monte Carlo when it lies within surface_band of the top or bottom surface,
 * farther than center_band from the detector's centre (the network was
 * trained on events there: photons spread to the detector's sides), when
 * its geometry is outside the network's training domain -- thickness,
 * radius and top reflectivity outside the ranges below, a detector extent
 * other than det_extent, or another value of the parameters the network
 * does not take as input (n_C 1.8, n_IC 1.0, bulk 1e-4/um, beta 0.2,
 * d_min 1, d_max 280 um, sensor reflectivity 0) -- or (audit) with
 * probability audit_fraction.  All other events are transported by the
 * network: per event a binomial number of detected photons and, for each,
 * a radius drawn from the predicted radial distribution and a uniform
 * azimuth, tallied like the Monte Carlo's detections.  Audit events are
 * tallied from the Monte Carlo and also compared with the network's
 * prediction (fd2_surrogate_stats).  Synchronous. */
typedef struct fd2_surrogate_cfg {
  float surface_band;       /* um (default 5)                               */
  float audit_fraction;     /* (default 0.01)                               */
  float h_min, h_max;       /* training ranges of thickness,                */
  float r_min, r_max;       /* column radius                                */
  float refl_min, refl_max; /* and top reflectivity                         */
  uint32_t seed;
  float center_band;        /* um from the detector centre (default 25)     */
  float det_extent;         /* detx = dety of the training set (default 1001) */
  int32_t reserved;
} fd2_surrogate_cfg;
typedef struct fd2_surrogate_stats {
  int64_t n_events, n_mc, n_surrogate, n_audit;
  int64_t photons_mc, photons_surrogate;
  double audit_detected_mc;      /* detections of audit events (MC)         */
  double audit_detected_pred;    /* sum over audit events of N * p_det      */
  double audit_radial_l1;        /* L1 of pooled radial distributions       */
  /* wall time of the call; GPU time of the Monte Carlo launch (fallback and
   * audit events, with the queue build); GPU time of the network part
   * (upload, gather, forward, sampling) on the network's own stream, which
   * runs alongside the Monte Carlo launch and takes the compute units it
   * frees as it drains */
  double ms_total, ms_mc, ms_surrogate;
} fd2_surrogate_stats;
int fd2_surrogate_submit(fd2_ctx *ctx, fd2_mlp *mlp, const fd2_event *events,
                         int64_t n, const fd2_surrogate_cfg *cfg,
                         fd2_surrogate_stats *stats);

/* ---- batches as HIP graphs ----------------------------------------------- */

/* Capture the work-queue build and the persistent transport of n events in
 * device memory (valid for the graph's life) into a HIP graph with its own
 * queue buffers.  Each fd2_graph_launch replays it asynchronously on the
 * context's stream (fd2_sync waits): the events get the context's next n
 * event numbers, as an fd2_submit_device of them would, so every replay is
 * a new batch with its own random streams, tallied into the context.  The
 * launch shape is the context's tuning at capture.  No counterpart in the
 * reference (one kernel launch per buffer, drv-G:277). */
typedef struct fd2_graph fd2_graph;
int fd2_graph_create(fd2_ctx *ctx, const fd2_event *d_events, int64_t n,
                     fd2_graph **out);
int fd2_graph_launch(fd2_graph *graph);
void fd2_graph_destroy(fd2_graph *graph);

/* Launch-shape autotuner: times the transport of n device events on a
 * scratch context with the same parameters (the context's tallies are not
 * touched) for candidate shapes -- coordinate descent, two sweeps over chunk
 * {16..256}, waves per CU {12, 16, 20}, parked lanes {8..48}, LDS tile side
 * {none, 32..80}, refill lanes {1..32} and wall queue {off, on}, each the
 * best of `repeats` runs after a warm-up --
 * and sets the fastest on ctx.  Candidates and their times go to results
 * (up to max_results; *n_results = number written). */
typedef struct fd2_tune_result {
  fd2_tuning tuning;
  float ms;
} fd2_tune_result;
int fd2_autotune(fd2_ctx *ctx, const fd2_event *d_events, int64_t n, int32_t repeats,
                 fd2_tuning *best, fd2_tune_result *results, int32_t max_results,
                 int32_t *n_results);

/* ---- reference file formats -------------------------------------------- */

/* hybridMANTIS_input.in: 24 positional tokens, first token of every
 * non-comment, non-blank line (main/hybridMANTIS_penEasy.f:190-229).  The
 * ten tokens used in output file names are returned verbatim in `tokens`
 * (10 x 16 bytes: H, R, top, bulk, beta, dmax, sensorRefl, yield, pixel,
 * machine), may be NULL. */
int fd2_read_input(const char *path, fd2_params *out, char tokens[10][16]);
/* myimage_<tokens>.dat / detected_<tokens>.dat names
 * (main/hybridMANTIS_tallyEnergyDepositionEvents.f:421-431) */
int fd2_output_names(const char tokens[10][16], char *prf_name, size_t prf_len,
                     char *phs_name, size_t phs_len);
/* PRF writer (tally .f:418, 448-453): pixel blocks of (1x,F12.4) values of
 * count/n_hist, y outer and x inner, each block followed by a blank line,
 * int((ub-lb)/pitch) pixels per side.  Fails if the file exists. */
int fd2_write_prf(const char *path, const fd2_params *p, const uint64_t *prf,
                  int32_t xdim, int32_t ydim, double n_hist);
/* PHS writer (tally .f:456-458): mynumbins lines of (1x,I8). */
int fd2_write_phs(const char *path, const int32_t *phs, int32_t num_bins);
/* The console block "OPTICAL TRANSPORT STATISTICS" the reference tally
 * prints at the end of a run (tally .f:461-496), byte for byte, to `path`
 * (overwritten) or to standard output when path is NULL.  stats holds the
 * 8 doubles of COMMON /optstats/ in order (glgen, gldetect, glabstop,
 * glabsbulk, gllost, gloutofcol, gltheta1, glgputime in ms); rungpu = 1
 * names the GPU, else the CPU; penoptical_s is the optical wall time the
 * driver measured (COMMON /tpenoptical/, seconds).  Values print as
 * (1x,F12.2). */
int fd2_write_stats(const char *path, const double *stats, int rungpu, double penoptical_s);
/* EDE text records (3(1x,1pe16.9),1x,1pe12.4,i8): x y z (cm) de (eV)
 * history (tally .f:173-174).  A 72-character line is cut at the record's
 * columns; other lines are read as five blank-separated numbers.  Lines
 * starting with '#' and blank lines are skipped.  Reads up to max events
 * (FD2_E_INVAL past that), converts cm to um, sets N = 0 (the caller
 * samples the photon count); *n gets the count. */
int fd2_read_ede(const char *path, fd2_event *events, int64_t max, int64_t *n);

#ifdef __cplusplus
}
#endif
#endif
