/* fd2_fortran.h -- the reference's four optical entry points, gfortran ABI.
 *
 * libfd2_fortran.so implements the entry points that penEasy's tally calls
 * (main/hybridMANTIS_tallyEnergyDepositionEvents.f:201-202, 257, 266, 285,
 * 304; main/hybridMANTIS_penEasy.f:436, 443) over the fd2 library
 * (fd2.h).  They read and accumulate into the same COMMON blocks as the
 * reference's drivers, with the reference's names and layouts, so a
 * gfortran-built penEasy links against this library unchanged:
 *
 *   gfortran penEasy.o ... -L<dir> -lfd2_fortran -lfd2 -Wl,-rpath,<dir>
 *
 * The blocks are also defined (weakly) in the library so that a C or Python
 * caller can fill them; an executable that declares the COMMON blocks
 * itself provides the storage both sides use.
 *
 * All optical transport runs on the GPU.  Photon counts N are drawn on the
 * host from Poisson(E * yield) (the reference uses GSL's default generator,
 * which is not reproduced: the draws here are counter-based, keyed by the
 * seed and the event's number, so they do not depend on how a run is split
 * into buffers).  Environment: FD2_SEED (default 271828182) seeds the
 * photon counts and the transport; FD2_DEVICE (default 0) selects the GPU.
 * Both are keyed by (seed, event number): a gpuoptical_ run (99 .. 101)
 * numbers its events 0, 1, ... across its buffers with seed FD2_SEED; the
 * k-th cpuoptical_ call (k = 0, 1, ...) uses seed FD2_SEED + 1 + k and the
 * k-th gpuoptlb_ call FD2_SEED - 1 - k, each numbering its events from 0.
 *
 * Errors print a message to stderr and abort, as the reference's
 * cutilSafeCall does (the reference's exit(0) on N > 900000 becomes an
 * abort too).  PRF images larger than 501 x 501 and PHS with more than 1000
 * bins do not fit the COMMON blocks and are rejected at initialisation.
 */
#ifndef FD2_FORTRAN_H
#define FD2_FORTRAN_H

#include <stdint.h>

#include "fd2.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Events per buffer: parameter maxvalopt,
 * main/hybridMANTIS_tallyEnergyDepositionEvents.f:55-56. */
#ifndef FD2_OPTICAL_BUFSIZE
#define FD2_OPTICAL_BUFSIZE 2304000
#endif

/* COMMON /optical/ (kernel_cuda_c_ver1_0.cu:86-97; tally .f:62-68):
 * positions in cm, deposited energy in eV, 1-based history numbers. */
struct fd2f_optical {
  double xbufopt[FD2_OPTICAL_BUFSIZE];
  double ybufopt[FD2_OPTICAL_BUFSIZE];
  double zbufopt[FD2_OPTICAL_BUFSIZE];
  double debufopt[FD2_OPTICAL_BUFSIZE];
  int32_t nbufopt[FD2_OPTICAL_BUFSIZE];
  int32_t myctropt;
  int32_t cpu_num_real;
};

/* COMMON /optstats/ (kernel_cuda_c_ver1_0.cu:111-123), accumulated. */
struct fd2f_optstats {
  double glgen, gldetect, glabstop, glabsbulk, gllost, gloutofcol, gltheta1;
  double glgputime; /* milliseconds */
};

/* COMMON /outputdetprim/ (kernel_cuda_c_ver1_0.cu:138-143): the PHS. */
struct fd2f_outputdetprim {
  int32_t gldetprimary[1000];
};

/* COMMON /outputimage/ (kernel_cuda_c_ver1_0.cu:151-157): PRF [x][y]. */
struct fd2f_outputimage {
  uint64_t newimageopt[501][501];
  uint64_t tempimageopt[501][501];
};

/* COMMON /inputargs/ (kernel_cuda_c_ver1_0.cu:178-203) has the layout of
 * fd2_params. */
extern fd2_params inputargs_;
extern struct fd2f_optical optical_;
extern struct fd2f_optstats optstats_;
extern struct fd2f_outputdetprim outputdetprim_;
extern struct fd2f_outputimage outputimage_;

/* main/hybridMANTIS_cuda_ver1_0.cu:107.  *penctr = 99: create the GPU
 * context for the run (inputargs_ must be set); 100: transport events
 * [0, *myfactGPU) of optical_ asynchronously (the buffer may be refilled on
 * return); 101: the same, then wait and add the run's PRF, PHS and
 * counters into outputimage_, outputdetprim_ and optstats_, and release the
 * context. */
void gpuoptical_(int32_t* penctr, int32_t* myfactGPU);

/* main/hybridMANTIS_c_ver1_0.c:81.  Transport events
 * [*gpusize, optical_.myctropt) synchronously and add their PRF, PHS (of
 * this call's events only, as the reference does) and counters into the
 * COMMON blocks.  Runs on the GPU. */
void cpuoptical_(int32_t* gpusize);

/* main/hybridMANTIS_cuda_ver1_0_LB.cu:108.  Load-balancing probe:
 * transport events [0, *lbbuf) synchronously, add the results into the
 * COMMON blocks, and return the elapsed time in seconds. */
void gpuoptlb_(double* gptime, int32_t* lbbuf);

/* main/hybridMANTIS_c_ver1_0_LB.c:86.  There is no CPU transport: returns
 * +inf seconds, which makes the tally's balancer (tally .f:209-221) send
 * every buffer to gpuoptical_ (factorLB = 1).  The reference's probe only
 * writes tempimageopt, which no output reads. */
void cpuoptlb_(double* cptime, int32_t* cpubufsize);

/* Photon counts of events [0, n): out[i] ~ Poisson(mu[i]) drawn from the
 * stream of (seed, first + i).  The entry points above use it with
 * mu = debufopt * (float)detyield and the seeds and event numbers stated
 * at the top of this header.  FD2_E_INVAL on a negative or non-finite mean. */
int fd2f_poisson(const double* mu, int64_t n, uint32_t seed, int64_t first,
                 int32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* FD2_FORTRAN_H */
