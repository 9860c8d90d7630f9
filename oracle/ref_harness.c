/*
 * ref_harness.c -- drives the reference's own transport C branch.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile, target `ref`) with
 * main/kernel_cuda_c_ver1_0.cu compiled as C straight from the reference
 * tree; the result lives in oracle/_ref/ and is never committed.
 *
 * This file is the caller the reference expects: it defines the COMMON block
 * the C branch writes detections into (outputimage_, K:159-163) and repeats
 * the event loop of main/hybridMANTIS_c_ver1_0.c:191-299 with two changes
 * needed for a deterministic check: a fixed seed_input replaces the
 * wall-clock one (drv:237) and N is given instead of drawn from GSL's
 * Poisson sampler (drv:219-220).  It records which counter each photon
 * incremented (its fate) and, for detections, which PRF pixel changed.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fdo.h"

/* the reference's struct start_info layout (K:67-75) */
struct ref_event {
  double x, y, z, E;
  int hist;
  int N;
};

/* symbols defined by the reference C branch */
extern unsigned long long int num_generated, num_detect, num_abs_top,
    num_abs_bulk, num_lost, num_outofcol, local_counter, num_theta1;
extern int absorbed, detect, bulk_abs;
int algo(float *normal, float *old_pos, float *pos, float *dcos,
         unsigned long long int *num_rebound, int *seed, struct ref_event info,
         unsigned long long int *myimage, float xdetector, float ydetector,
         float R, float H, float n1, float n2, float top_absfrac,
         float bulk_abscoeff, float beta, float d_min, int pixelsize,
         float lbound_x, float lbound_y, float ubound_x, float ubound_y,
         float sensorRefl, float d_max, int ydim, int *h_num_detected_prim);
void init_PRNG(int history_batch, int histories_per_thread, int seed_input,
               int *seed);
int abMODm(int m, int a, int s);
float ranecu(int *seed);

/* COMMON /outputimage/ (Fortran integer*8 newimageopt(0:500,0:500)) */
struct {
  unsigned long long int newimageopt[501][501];
  unsigned long long int tempimageopt[501][501];
} outputimage_;

/* thin exports of the reference's RNG for known-answer comparisons */
float fdref_ranecu(int32_t seed[2]) { return ranecu(seed); }
void fdref_init_prng(int b, int h, int s, int32_t seed[2]) { init_PRNG(b, h, s, seed); }
int32_t fdref_abmodm(int32_t m, int32_t a, int32_t s) { return abMODm(m, a, s); }

/* Run events [0,n) through the reference C branch.  prf (xdim*ydim),
 * hist_count (num_hist) and counters (FDO_NCOUNT) are accumulated;
 * photons (optional) receives one record per photon.  rezero repeats the
 * reference driver's two per-event clears of its local xdim*ydim image
 * (drv:205-207, 248-250), which only matters for timing. */
int fdref_run(const fdo_geom *g, const fdo_event *events, int64_t n,
              int64_t event_base, uint32_t seed_input, uint64_t *prf,
              uint32_t *hist_count, int32_t num_hist, uint64_t *counters,
              fdo_photon *photons, int rezero) {
  if (g->xdim > 501 || g->ydim > 501 || g->xdim <= 0 || g->ydim <= 0) return -1;
  memset(&outputimage_, 0, sizeof outputimage_);
  int *hp = (int *)calloc((size_t)num_hist + 2, sizeof(int));
  unsigned long long int *shadow = NULL;
  if (photons) shadow = (unsigned long long int *)calloc(501 * 501, sizeof *shadow);
  unsigned long long int *local_img = NULL;
  size_t local_sz = (size_t)g->xdim * (size_t)g->ydim;
  if (rezero) local_img = (unsigned long long int *)malloc(local_sz * sizeof *local_img);
  if (!hp || (photons && !shadow) || (rezero && !local_img)) {
    free(hp); free(shadow); free(local_img);
    return -2;
  }
  fdo_photon *rec = photons;
  for (int64_t e = 0; e < n; ++e) {
    const fdo_event *fe = &events[e];
    struct ref_event info = {fe->x, fe->y, fe->z, fe->E, fe->hist, fe->N};
    num_generated = num_detect = num_abs_top = num_abs_bulk = 0;
    num_lost = num_outofcol = num_theta1 = 0;
    if (rezero) memset(local_img, 0, local_sz * sizeof *local_img);
    int N = info.N > 0 ? info.N : 0;
    unsigned long long int *nreb =
        (unsigned long long int *)calloc((size_t)N + 1, sizeof *nreb);
    int seed[2];
    init_PRNG((int)(event_base + e), 50000, (int)seed_input, seed);
    if (rezero) memset(local_img, 0, local_sz * sizeof *local_img);
    float pos[3] = {(float)info.x, (float)info.y, (float)info.z};
    float old_pos[3] = {pos[0], pos[1], pos[2]};
    float dcos[3] = {0.0f, 0.0f, 0.0f}, normal[3] = {0.0f, 0.0f, 0.0f};

    /* first photon's direction: drv:257-277 restated */
    float r = (ranecu(seed) * 2.0f) - 1.0f;
    while (r <= 0.01f && r >= -0.01f) r = (ranecu(seed) * 2.0f) - 1.0f;
    dcos[2] = r;
    float rr = (float)sqrt((double)(1.0f - r * r));
    float th = (float)((double)ranecu(seed) * 6.283185308);
    dcos[0] = (float)((double)rr * cos((double)th));
    dcos[1] = (float)((double)rr * sin((double)th));
    float nrm = (float)sqrt((double)(dcos[0] * dcos[0] + dcos[1] * dcos[1] + dcos[2] * dcos[2]));
    if ((double)nrm < 1.0 - 8.1929093e-6 || (double)nrm > 1.0 + 8.1929093e-6) {
      dcos[0] = dcos[0] / nrm;
      dcos[1] = dcos[1] / nrm;
      dcos[2] = dcos[2] / nrm;
    }

    local_counter = 0;
    int32_t k = 0;
    while (local_counter < (unsigned long long int)N) {
      absorbed = 0;
      detect = 0;
      bulk_abs = 0;
      pos[0] = (float)info.x; pos[1] = (float)info.y; pos[2] = (float)info.z;
      old_pos[0] = pos[0]; old_pos[1] = pos[1]; old_pos[2] = pos[2];
      num_generated++;
      unsigned long long int before[7] = {num_detect, num_abs_top, num_abs_bulk,
                                          num_lost, num_outofcol, num_theta1, 0};
      while (algo(normal, old_pos, pos, dcos, nreb, seed, info,
                  &outputimage_.newimageopt[0][0], g->xdet, g->ydet, g->R, g->H,
                  g->n_C, g->n_IC, g->top_absfrac, g->bulk, g->beta, g->d_min,
                  g->pixel, g->lbx, g->lby, g->ubx, g->uby, g->sensor_refl,
                  g->d_max, g->ydim, hp) == 0) {
      }
      if (rec) {
        int fate = FDO_FATE_NONE;
        int32_t pix = -1;
        if (num_detect != before[0]) {
          fate = FDO_FATE_DETECT_OUT;
          for (int i = 0; i < g->xdim && pix < 0; ++i)
            for (int j = 0; j < g->ydim; ++j)
              if (outputimage_.newimageopt[i][j] != shadow[i * 501 + j]) {
                shadow[i * 501 + j] = outputimage_.newimageopt[i][j];
                pix = i * g->ydim + j;
                fate = FDO_FATE_DETECT;
                break;
              }
        } else if (num_abs_top != before[1]) fate = FDO_FATE_ABS_TOP;
        else if (num_abs_bulk != before[2]) fate = FDO_FATE_ABS_BULK;
        else if (num_lost != before[3]) fate = FDO_FATE_LOST;
        else if (num_outofcol != before[4]) fate = FDO_FATE_OUTOFCOL;
        else if (num_theta1 != before[5]) fate = FDO_FATE_THETA1;
        rec[k].fate = fate;
        rec[k].pix = pix;
        rec[k].x = rec[k].y = 0.0f;
      }
      ++k;
    }
    free(nreb);
    counters[FDO_C_GENERATED] += num_generated;
    counters[FDO_C_DETECT] += num_detect;
    counters[FDO_C_ABS_TOP] += num_abs_top;
    counters[FDO_C_ABS_BULK] += num_abs_bulk;
    counters[FDO_C_LOST] += num_lost;
    counters[FDO_C_OUTOFCOL] += num_outofcol;
    counters[FDO_C_THETA1] += num_theta1;
    if (rec) rec += N;
  }
  for (int i = 0; i < g->xdim; ++i)
    for (int j = 0; j < g->ydim; ++j)
      prf[(int64_t)i * g->ydim + j] += outputimage_.newimageopt[i][j];
  /* the C branch indexes by the 1-based history number (K:3419) */
  for (int32_t h = 1; h <= num_hist; ++h) hist_count[h - 1] += (uint32_t)hp[h];
  free(hp);
  free(shadow);
  free(local_img);
  return 0;
}
