// fd2_kernels.hip -- HIP kernels of the MI355X fastDETECT2 transport.
//
//   k_queue_*     device work queue: exclusive scans of per-event chunk and
//                 photon counts (replaces the reference's host load
//                 balancer, main/hybridMANTIS_tallyEnergyDepositionEvents.f:
//                 195-250, and its one-thread-per-event launch,
//                 main/hybridMANTIS_cuda_ver1_0.cu:277)
//   k_transport   persistent kernel: each wave takes chunks of photons from
//                 the queue with one atomic per chunk and hands photons to
//                 its idle lanes, so events of very uneven size keep every
//                 lane busy
//   k_trace_ranecu one lane per event, photons in sequence, the reference's
//                 per-event RANECU stream (trajectory test)
//   k_debug_steps per-step FP32 trace of a few photons (FP64 re-trace check)
//   k_debug_math  the device arithmetic on given arguments
#include "fd2_internal.h"

namespace fd2 {

// ------------------------------------------------------------ helpers

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// exclusive scan over the block (kScanThreads threads); returns the
// exclusive prefix of `v` and the block total in `total`
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* s_wave, T& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int kWaves = kScanThreads / 64;
  T inc = wave_incl_scan(v);
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    T w = lane < kWaves ? s_wave[lane] : T(0);
    T wi = wave_incl_scan(w);
    if (lane < kWaves) s_wave[lane] = wi - w;
    if (lane == kWaves - 1) s_wave[kWaves] = wi;
  }
  __syncthreads();
  total = s_wave[kWaves];
  T r = s_wave[wid] + inc - v;
  __syncthreads();
  return r;
}

__device__ __forceinline__ void event_counts(const QueueArgs& q, int64_t e,
                                             uint32_t& chunks, uint32_t& photons) {
  chunks = 0;
  photons = 0;
  if (e >= q.n_events) return;
  const fd2_event ev = q.events[e];
  if (ev.N < 0 || ev.N > FD2_MAX_PHOTONS_PER_EVENT || ev.geom >= (uint32_t)q.n_geoms) {
    if (ev.N != 0) atomicOr(q.error_flag, 1u);
    return;
  }
  photons = (uint32_t)ev.N;
  chunks = (photons + (1u << q.chunk_log2) - 1u) >> q.chunk_log2;
}

// pass 1: per-block totals
__global__ void __launch_bounds__(kScanThreads) k_queue_totals(QueueArgs q) {
  __shared__ uint32_t s_c[kScanThreads / 64 + 1];
  __shared__ unsigned long long s_p[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t c = 0;
  unsigned long long ph = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint32_t cc, pp;
    event_counts(q, base + i, cc, pp);
    c += cc;
    ph += pp;
  }
  uint32_t tc;
  unsigned long long tp;
  block_excl_scan<uint32_t>(c, s_c, tc);
  block_excl_scan<unsigned long long>(ph, s_p, tp);
  if (threadIdx.x == 0) {
    q.block_chunks[blockIdx.x] = tc;
    q.block_photons[blockIdx.x] = tp;
  }
}

// pass 2: scan of the block totals (one block), grand totals, cursor reset.
// The work-item total is summed in 64 bits: a launch of more than
// q.max_chunks items (which the 32-bit offsets and cursor cannot number)
// sets error bit 4 and gets a total of 0, so the transport takes nothing.
__global__ void __launch_bounds__(kScanThreads) k_queue_blocks(QueueArgs q) {
  __shared__ uint32_t s_c[kScanThreads / 64 + 1];
  __shared__ unsigned long long s_p[kScanThreads / 64 + 1];
  uint64_t carry_c = 0;
  unsigned long long carry_p = 0;
  for (int b0 = 0; b0 < q.n_blocks; b0 += kScanThreads) {
    const int b = b0 + threadIdx.x;
    const uint32_t c = b < q.n_blocks ? q.block_chunks[b] : 0u;
    const unsigned long long p = b < q.n_blocks ? q.block_photons[b] : 0ull;
    uint32_t tc;
    unsigned long long tp;
    const uint32_t ec = block_excl_scan<uint32_t>(c, s_c, tc);
    const unsigned long long ep = block_excl_scan<unsigned long long>(p, s_p, tp);
    if (b < q.n_blocks) {
      q.block_chunks[b] = (uint32_t)(carry_c + ec);
      q.block_photons[b] = carry_p + ep;
    }
    carry_c += tc;
    carry_p += tp;
  }
  if (threadIdx.x == 0) {
    const bool over = carry_c > q.max_chunks;
    if (over) atomicOr(q.error_flag, 4u);
    q.chunk_off[q.n_events] = over ? 0u : (uint32_t)carry_c;
    q.photon_off[q.n_events] = carry_p;
    *q.cursor = 0u;
  }
}

// pass 3: per-event exclusive offsets
__global__ void __launch_bounds__(kScanThreads) k_queue_offsets(QueueArgs q) {
  __shared__ uint32_t s_c[kScanThreads / 64 + 1];
  __shared__ unsigned long long s_p[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t cc[kScanItems], pp[kScanItems];
  uint32_t c = 0;
  unsigned long long ph = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    event_counts(q, base + i, cc[i], pp[i]);
    c += cc[i];
    ph += pp[i];
  }
  uint32_t tc;
  unsigned long long tp;
  uint32_t oc = block_excl_scan<uint32_t>(c, s_c, tc) + q.block_chunks[blockIdx.x];
  unsigned long long op =
      block_excl_scan<unsigned long long>(ph, s_p, tp) + q.block_photons[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t e = base + i;
    if (e < q.n_events) {
      q.chunk_off[e] = oc;
      q.photon_off[e] = op;
    }
    oc += cc[i];
    op += pp[i];
  }
}

hipError_t launch_queue_build(const QueueArgs& q, hipStream_t s) {
  if (q.n_blocks <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_queue_totals, dim3(q.n_blocks), dim3(kScanThreads), 0, s, q);
  hipLaunchKernelGGL(k_queue_blocks, dim3(1), dim3(kScanThreads), 0, s, q);
  hipLaunchKernelGGL(k_queue_offsets, dim3(q.n_blocks), dim3(kScanThreads), 0, s, q);
  return hipGetLastError();
}

// ------------------------------------------------------------ tallies

__device__ __forceinline__ int fate_counter(int fate) {
  switch (fate) {
    case kDetect:
    case kDetectOut: return FD2_C_DETECT;
    case kAbsTop: return FD2_C_ABS_TOP;
    case kAbsBulk: return FD2_C_ABS_BULK;
    case kLost: return FD2_C_LOST;
    case kOutOfCol: return FD2_C_OUTOFCOL;
    case kTheta1: return FD2_C_THETA1;
    default: return FD2_C_REGEN;  // unreachable for an ended photon
  }
}

// Counter-mode initial direction of photon k: photon 0 takes the driver's
// first direction (|2u-1| > 0.01, main/hybridMANTIS_c_ver1_0.c:257-277),
// the others the direction algo() draws at the previous photon's end
// (main/kernel_cuda_c_ver1_0.cu:578-602) followed by the RoughSurface draws
// of the (never read) initial normal.
__device__ __forceinline__ void counter_initial_dir(CounterRng& rng, Photon& p,
                                                    uint32_t k, const GeomDev& g) {
  if (k == 0) {
    emit_first_dir(rng, p, g);
  } else {
    emit_dir(rng, p, g);
    if (g.beta > 0.0f) {
      (void)rng.u();
      (void)rng.u();
    }
  }
}

// ------------------------------------------------------- transport

// Register budget of the transport kernel: 5 waves per SIMD (96 VGPRs).  The
// wall branch's live set needs 114 VGPRs (66 without it), so 10 values are
// spilled to scratch at this budget; the fifth wave still wins: 149.8 ->
// 140.6 ms per mono25x50 launch on MI355X at 20 waves per CU (6 waves, 80
// VGPRs and 25 spilled values, lose).  FD2_WAVES_PER_EU=0 lets the compiler
// choose.
#ifndef FD2_WAVES_PER_EU
#define FD2_WAVES_PER_EU 5
#endif
// With several geometries per launch (the geometry table in LDS, read per
// lane) the kernel needs more registers and keeps 4 waves per SIMD.
#if FD2_WAVES_PER_EU > 0
#define FD2_TRANSPORT_ATTR \
  __attribute__((amdgpu_waves_per_eu(MULTI ? 4 : FD2_WAVES_PER_EU, MULTI ? 4 : FD2_WAVES_PER_EU)))
#else
#define FD2_TRANSPORT_ATTR
#endif

// log-radial bin of a detection offset (FD2_RADIAL_BINS layout)
__device__ __forceinline__ int radial_bin(float dx, float dy) {
  const float r = sqrtf(dx * dx + dy * dy);
  if (!(r >= 0.5f)) return 0;
  if (r >= 800.0f) return FD2_RADIAL_BINS - 1;
  // 1 + floor(62 * ln(r/0.5) / ln(1600))
  int b = 1 + (int)(fd2_log(r * 2.0f) * 8.40363595f);
  return b < 1 ? 1 : (b > FD2_RADIAL_BINS - 2 ? FD2_RADIAL_BINS - 2 : b);
}

// first pixel row/column of a geometry's LDS tile
__device__ __forceinline__ int tile_origin(int dim, int side) {
  return dim > side ? (dim - side) >> 1 : 0;
}

// Adds a lane's run of detections of one event to the event's history
// count (the event's history and geometry are read here, once per run, not
// once per detection).
template <bool MULTI>
__device__ __forceinline__ void flush_run(const GeomDev* __restrict__ geoms,
                                          const fd2_event* __restrict__ events, uint32_t run_ev,
                                          uint32_t run_cnt) {
  const fd2_event& ev = events[run_ev];
  const int32_t hist = ev.hist;
  const GeomDev& g = geoms[MULTI ? ev.geom : 0u];
  if (hist >= 1 && hist <= g.num_hist) tally_add(&g.hist[hist - 1], run_cnt);
}

// Tallies of an ended photon: fate counter, PRF pixel (LDS tile or HBM),
// per-lane run of the event's detections, radial bin, record.
template <bool MULTI, bool REC, bool RADIAL>
__device__ __forceinline__ void finish_photon(const TransportArgs& a,
                                              const GeomDev* __restrict__ geoms,
                                              const fd2_event* __restrict__ events,
                                              const Photon& p, uint32_t gi,
                                              uint32_t lane_ev, uint32_t lane_k, uint32_t* s_cnt,
                                              uint32_t* s_tile, int side, int tile_area,
                                              uint32_t* run) {
  atomicAdd(&s_cnt[gi * kCounters + fate_counter(p.fate)], 1u);
  const bool det = p.fate == kDetect || p.fate == kDetectOut;
  if (det) {
    const GeomDev& gt = geoms[MULTI ? gi : 0u];
    if (p.fate == kDetect) {
      bool in_tile = false;
      if ((int)gi < a.n_tiles) {
        const uint32_t ti = (uint32_t)((p.pix >> 16) - tile_origin(gt.xdim, side));
        const uint32_t tj = (uint32_t)((p.pix & 0xFFFF) - tile_origin(gt.ydim, side));
        if (ti < (uint32_t)side && tj < (uint32_t)side) {
          atomicAdd(&s_tile[gi * tile_area + ti * side + tj], 1u);
          in_tile = true;
        }
      }
      if (!in_tile) tally_add(&gt.prf[pix_index(p.pix, gt)], 1ull);
    }
    if (lane_ev == run[0]) {
      ++run[1];
    } else {
      if (run[1]) flush_run<MULTI>(geoms, events, run[0], run[1]);
      run[0] = lane_ev;
      run[1] = 1;
    }
  }
  if (RADIAL && det) {
    const fd2_event& ev = events[lane_ev];
    tally_add(&a.radial[(size_t)lane_ev * FD2_RADIAL_BINS + radial_bin(p.px - ev.x, p.py - ev.y)],
              1u);
  }
  if (REC) {
    fd2_photon rec;
    rec.fate = p.fate;
    rec.pix = p.fate == kDetect ? pix_index(p.pix, geoms[MULTI ? gi : 0u]) : -1;
    rec.x = det ? p.px : 0.0f;
    rec.y = det ? p.py : 0.0f;
    a.records[a.photon_off[lane_ev] + lane_k] = rec;
  }
}

// Per-wave photon queues in HBM (wall queue, gap queue, backlog: kQSlots
// entries of kQWords words).  Slots come in blocks of kQBlock = 32; word j
// of slot s is at [(s / 32) * 32 * kQWords + 32 j + s % 32], so the lanes of
// a wave write whole 128-byte lines and all words of a slot lie within
// 3 KB of its first (one address per lane, the rest immediate offsets).  An
// entry is everything a photon in flight carries: position, direction,
// column axis, loop count and flags, random stream, event, geometry and
// photon index (not the surface normal: every path writes it before reading
// it); gap-queue entries also the wall branch's loop state (PrptState).
__device__ __forceinline__ size_t q_base(uint32_t slot) {
  return (size_t)(slot / kQBlock) * (kQBlock * kQWords) + slot % kQBlock;
}

// An entry's words lie `S` words apart from its first: kQBlock in the HBM
// queues, kWallCap in a wave's LDS wall queue (word-major, one row of
// kWallCap slots per word).  The geometry index is stored only for several
// geometries per launch (MULTI; else 0) and the photon index only for
// per-photon records (REC).
template <bool MULTI, bool REC, int S = kQBlock>
__device__ __forceinline__ void q_put_at(uint32_t* w, const Photon& p, const CounterRng& r,
                                         uint32_t ev, uint32_t gi, uint32_t k, bool parked) {
  w[0 * S] = __float_as_uint(p.px);
  w[1 * S] = __float_as_uint(p.py);
  w[2 * S] = __float_as_uint(p.pz);
  w[3 * S] = __float_as_uint(p.ux);
  w[4 * S] = __float_as_uint(p.uy);
  w[5 * S] = __float_as_uint(p.uz);
  w[6 * S] = __float_as_uint(p.xc);
  w[7 * S] = __float_as_uint(p.yc);
  // loops (21 bits) | sensor resume (bit 22) | parked | new-column sample
  // due, with its rejected samples (<= 100) in bits 25-31
  w[8 * S] = (uint32_t)p.loops |
             (p.fate == kResume ? 1u << 22 : 0u) | (parked ? 1u << 23 : 0u) |
             (p.fate == kNewColumn ? (1u << 24) | ((uint32_t)p.pix << 25) : 0u);
  w[9 * S] = r.s0;
  w[10 * S] = r.s1;
  w[11 * S] = r.s2;
  w[12 * S] = r.s3;
  w[13 * S] = ev;
  if (MULTI) w[14 * S] = gi;
  if (REC) w[15 * S] = k;
}

template <bool MULTI, bool REC>
__device__ __forceinline__ void q_put(uint32_t* wq, uint32_t slot, const Photon& p,
                                      const CounterRng& r, uint32_t ev, uint32_t gi, uint32_t k,
                                      bool parked) {
  q_put_at<MULTI, REC>(wq + q_base(slot), p, r, ev, gi, k, parked);
}

__device__ __forceinline__ void q_put_state(uint32_t* wq, uint32_t slot, const PrptState& st) {
  uint32_t* w = wq + q_base(slot);
  w[16 * kQBlock] = st.tries;
  w[17 * kQBlock] = __float_as_uint(st.pr);
  w[18 * kQBlock] = __float_as_uint(st.s2);
  w[19 * kQBlock] = __float_as_uint(st.c2);
}

__device__ __forceinline__ void q_get_state(const uint32_t* wq, uint32_t slot, PrptState& st) {
  const uint32_t* w = wq + q_base(slot);
  st.tries = w[16 * kQBlock];
  st.pr = __uint_as_float(w[17 * kQBlock]);
  st.s2 = __uint_as_float(w[18 * kQBlock]);
  st.c2 = __uint_as_float(w[19 * kQBlock]);
}

template <bool MULTI, bool REC, int S = kQBlock>
__device__ __forceinline__ void q_get_at(const uint32_t* w, Photon& p, CounterRng& r,
                                         uint32_t& ev, uint32_t& gi, uint32_t& k, bool& parked) {
  p.px = __uint_as_float(w[0 * S]);
  p.py = __uint_as_float(w[1 * S]);
  p.pz = __uint_as_float(w[2 * S]);
  p.ux = __uint_as_float(w[3 * S]);
  p.uy = __uint_as_float(w[4 * S]);
  p.uz = __uint_as_float(w[5 * S]);
  p.nx = p.ny = p.nz = 0.0f;  // written before it is read
  p.xc = __uint_as_float(w[6 * S]);
  p.yc = __uint_as_float(w[7 * S]);
  const uint32_t f = w[8 * S];
  p.loops = (int32_t)(f & 0x1fffffu);
  p.fate = ((f >> 22) & 1u) ? kResume : ((f >> 24) & 1u) ? kNewColumn : kNone;
  parked = (f >> 23) & 1u;
  p.pix = ((f >> 24) & 1u) ? (int32_t)(f >> 25) : -1;
  r.s0 = w[9 * S];
  r.s1 = w[10 * S];
  r.s2 = w[11 * S];
  r.s3 = w[12 * S];
  ev = w[13 * S];
  gi = MULTI ? w[14 * S] : 0u;
  k = REC ? w[15 * S] : 0u;
}

template <bool MULTI, bool REC>
__device__ __forceinline__ void q_get(const uint32_t* wq, uint32_t slot, Photon& p, CounterRng& r,
                                      uint32_t& ev, uint32_t& gi, uint32_t& k, bool& parked) {
  q_get_at<MULTI, REC>(wq + q_base(slot), p, r, ev, gi, k, parked);
}

// number of set bits of a wave-wide mask below this lane (v_mbcnt)
__device__ __forceinline__ uint32_t rank_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Moves the wave's photons in flight (stepping or waiting for the wall
// queue) to the top of its backlog; false if the backlog would overflow,
// which the in-flight bound rules out.
template <bool MULTI, bool REC>
__device__ __forceinline__ bool swap_to_backlog(uint32_t* wq, uint32_t& n_back, bool& alive,
                                                bool& parked, const Photon& p,
                                                const CounterRng& rng, uint32_t lane_ev,
                                                uint32_t gi, uint32_t lane_k) {
  const unsigned long long am = __ballot(alive);
  const uint32_t m = (uint32_t)__popcll(am);
  if (n_back + m > (uint32_t)kBackCap) return false;
  if (alive) {
    q_put<MULTI, REC>(wq, kBackBase + n_back + rank_below(am), p, rng,
                      lane_ev, gi, lane_k, parked);
    alive = false;
    parked = false;
  }
  n_back += m;
  return true;
}

// QUEUE = false: a lane whose photon reaches the wall-Fresnel branch parks
// in place and the wave runs the branch for its parked lanes once
// park_lanes of them wait.  QUEUE = true: the photon moves to the wave's
// wall queue in HBM and the lane takes new work; when 64 photons wait, the
// wave moves its photons in flight to its backlog and runs the branch with
// all 64 lanes, up to a refraction out of the column; those photons wait in
// the gap queue and their flights through the gap run, again 64 at a time,
// one flight per photon and pass, in a phase of their own (a photon
// reflected off the next column's outside waits for another flight).  Free
// lanes refill from the backlog before the work queue, so no more than
// 64 + kWallCap + kGapCap photons are ever in flight per wave and the
// backlog (kBackCap) cannot overflow; the gap queue holds at most 127.
// Every photon keeps its own state and random stream, so the results are
// the same either way.
//
// `geoms_in`, `events` and `chunk_off` are a.geoms, a.events and
// a.chunk_off again, as parameters of their own: the kernel never writes
// these tables (only the tallies the geometries point to), and saying so
// (__restrict__) lets the compiler read them with scalar loads where they
// are uniform instead of reloading them with vector loads around every
// store.
template <bool MULTI, bool REC, bool RADIAL, bool QUEUE>
__global__ void __launch_bounds__(kTransportThreads) FD2_TRANSPORT_ATTR
k_transport(TransportArgs a, const GeomDev* __restrict__ geoms_in,
            const fd2_event* __restrict__ events, const uint32_t* __restrict__ chunk_off) {
  // With several geometries a lane's geometry is its own, so their constants
  // cannot be scalar loads: the table goes to LDS instead (160 B each).
  static_assert(sizeof(GeomDev) % 4 == 0, "GeomDev copied as words");
  __shared__ GeomDev s_geoms[MULTI ? FD2_MAX_GEOMS : 1];
  if (MULTI) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(geoms_in);
    uint32_t* dst = reinterpret_cast<uint32_t*>(s_geoms);
    const int nw = a.n_geoms * (int)(sizeof(GeomDev) / 4);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
  }
  const GeomDev* __restrict__ geoms = MULTI ? s_geoms : geoms_in;
  __shared__ uint32_t s_cnt[FD2_MAX_GEOMS * kCounters];
  // the next photons of the wave's chunk, started ahead with all lanes:
  // photon start_base + j has its first direction in words 0-2 of slot j
  // and its random stream (after the initial draws) in words 3-6
  __shared__ uint32_t s_start[kTransportThreads / 64][kStartWords][64];
  // the waves' wall queues (QUEUE): entries in the layout of q_put_at, word-major
  __shared__ uint32_t s_wall[QUEUE ? kTransportThreads / 64 : 1][kWallWords][kWallCap];
  extern __shared__ uint32_t s_tile[];  // n_tiles x tile_side^2
  for (int i = threadIdx.x; i < a.n_geoms * kCounters; i += blockDim.x) s_cnt[i] = 0;
  const int side = a.tile_side;
  const int tile_area = side * side;
  for (int i = threadIdx.x; i < a.n_tiles * tile_area; i += blockDim.x) s_tile[i] = 0;
  __syncthreads();

  // per-lane run of detections of one event (event, count), added to the
  // event's history count when the lane's detections move on to another
  // event (photons of a chunk share their event); in LDS, as it is touched
  // only when a photon ends
  __shared__ uint32_t s_run[kTransportThreads][2];
  uint32_t* run = s_run[threadIdx.x];
  run[0] = 0u;
  run[1] = 0u;
  const uint64_t event_base = a.event_base_dev ? *a.event_base_dev : a.event_base;

  const int lane = threadIdx.x & 63;
  // the wave's index in the block, uniform across the wave: the compiler
  // then keeps the wave's queue and LDS bases in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t chunk = 1u << a.chunk_log2;

  // wave-uniform work pool: photons [pool_next, pool_end) of event pool_ev;
  // those in [start_base, start_end) (up to 64) are started in s_start
  int64_t pool_ev = 0;
  int32_t pool_next = 0, pool_end = 0, start_base = 0, start_end = 0;
  uint32_t (*start)[64] = s_start[wave];
  bool pool_done = false;

  Photon p;
  CounterRng rng;
  bool alive = false;
  bool parked = false;  // waiting for the wall-Fresnel branch
  uint32_t gi = 0;
  uint32_t lane_ev = 0;  // the lane's event (index into events)
  uint32_t lane_k = 0;   // photon index within the event
  StepSink nosink{nullptr, nullptr, 0, 0, false};

  // wave queues (QUEUE): wall [0, n_wall) in LDS, gap [0, n_gap) and backlog
  // [0, n_back) in HBM
  uint32_t* wq = nullptr;
  uint32_t* wall = &s_wall[QUEUE ? wave : 0][0][0];
  uint32_t n_wall = 0, n_gap = 0, n_back = 0, idle = 0;
  if (QUEUE)
    wq = a.wave_queue + (size_t)(blockIdx.x * (kTransportThreads / 64) + wave) *
                            (kQWords * kQSlots);

  for (;;) {
    // Starting a photon (stream seeding, first direction) is long and would
    // run with one or two lanes if done whenever a photon ends: wait until
    // refill_lanes lanes are free, or no lane is left to step.
    unsigned long long need = __ballot(!alive);
    if ((int)__popcll(need) < a.refill_lanes && need != ~0ull) need = 0ull;
    if (QUEUE && need != 0ull && n_back != 0u) {
      __threadfence_block();  // the wave's queue stores are complete before its loads
      const uint32_t take = min((uint32_t)__popcll(need), n_back);
      const uint32_t rank = rank_below(need);
      if (!alive && rank < take) {
        q_get<MULTI, REC>(wq, kBackBase + n_back - 1u - rank, p, rng, lane_ev, gi, lane_k, parked);
        alive = true;
      }
      n_back -= take;
      need = __ballot(!alive);
    }
    while (need != 0ull && !pool_done) {
      if (pool_next >= pool_end) {
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(a.cursor, 1u);
        c = (uint32_t)__builtin_amdgcn_readfirstlane((int)c);
        if (c >= chunk_off[a.n_events]) {
          pool_done = true;
          break;
        }
        // largest e with chunk_off[e] <= c (skips events without photons):
        // each round the 64 lanes (all active here) probe 64 points of
        // [lo, hi) with one vector load, and the wave keeps the step after
        // the last probe at or below c -- 3 rounds for 2e5 events where a
        // binary search makes 18 dependent loads
        int64_t lo = 0, hi = a.n_events;
        while (hi - lo > 1) {
          const int64_t step = (hi - lo + 63) >> 6;
          const int64_t pos = lo + (int64_t)lane * step;
          const bool le = pos < hi && chunk_off[pos] <= c;
          const unsigned long long m = __ballot(le);
          lo += (int64_t)(63 - __clzll(m)) * step;
          hi = min(hi, lo + step);
        }
        pool_ev = lo;
        const int32_t n_ph = events[lo].N;
        pool_next = (int32_t)((c - chunk_off[lo]) << a.chunk_log2);
        pool_end = min(n_ph, pool_next + (int32_t)chunk);
        if (lane == 0) {
          const uint32_t gg = MULTI ? events[lo].geom : 0u;
          atomicAdd(&s_cnt[gg * kCounters + FD2_C_GENERATED], (uint32_t)(pool_end - pool_next));
        }
        start_end = pool_next;
      }
      if (pool_next >= start_end) {
        // Seeding a stream (Philox) and drawing the first direction are
        // long; every lane starts one of the next 64 photons of the chunk
        // here, at full width, instead of the few lanes that take photons
        // in a pass.
        start_base = pool_next;
        start_end = min(pool_end, pool_next + 64);
        if (start_base + lane < start_end) {
          const uint32_t k = (uint32_t)(start_base + lane);
          CounterRng r;
          Photon q;
          r.seed(a.seed, event_base + (uint64_t)pool_ev, k);
          counter_initial_dir(r, q, k, geoms[MULTI ? events[pool_ev].geom : 0u]);
          start[0][lane] = __float_as_uint(q.ux);
          start[1][lane] = __float_as_uint(q.uy);
          start[2][lane] = __float_as_uint(q.uz);
          start[3][lane] = r.s0;
          start[4][lane] = r.s1;
          start[5][lane] = r.s2;
          start[6][lane] = r.s3;
        }
        // lanes read other lanes' slots below; the wave's LDS accesses
        // complete in order
        __builtin_amdgcn_wave_barrier();
      }
      const int avail = start_end - pool_next;  // started, not yet taken
      const int rank = (int)rank_below(need);
      if (!alive && rank < avail) {
        const fd2_event ev = events[pool_ev];
        gi = MULTI ? ev.geom : 0u;
        const uint32_t k = (uint32_t)(pool_next + rank);
        const uint32_t j = k - (uint32_t)start_base;
        start_photon(p, ev);
        p.ux = __uint_as_float(start[0][j]);
        p.uy = __uint_as_float(start[1][j]);
        p.uz = __uint_as_float(start[2][j]);
        rng.s0 = start[3][j];
        rng.s1 = start[4][j];
        rng.s2 = start[5][j];
        rng.s3 = start[6][j];
        lane_ev = (uint32_t)pool_ev;
        lane_k = k;
        parked = false;
        alive = true;
      }
      pool_next += min((int)__popcll(need), avail);
      need = __ballot(!alive);
    }
    // with no lane alive the work queue and the backlog are empty
    if (__ballot(alive) == 0ull && (!QUEUE || (n_wall == 0u && n_gap == 0u))) break;
    bool ended = false;
    // every pass steps a photon or runs the wall branch; a wave that does
    // neither for many passes has a broken queue state: it stops and flags
    // the launch rather than spin
    bool progress = __ballot(alive && !parked) != 0ull;
    if (alive && !parked) {
      const GeomDev& g = geoms[MULTI ? gi : 0u];
      const int r = hot_step<CounterRng, false>(rng, p, g, &nosink);
      if (r == kEnded) {
        ended = true;
      } else if (r == kTrapped) {
        atomicAdd(&s_cnt[gi * kCounters + FD2_C_REGEN], 1u);
        restart_trapped(rng, p, g, events[lane_ev]);
      } else if (r == kNeedPrpt) {
        parked = true;
      }
    }
    if (QUEUE) {
      if (ended) {
        finish_photon<MULTI, REC, RADIAL>(a, geoms, events, p, gi, lane_ev, lane_k, s_cnt, s_tile, side,
                                          tile_area, run);
        alive = false;
        ended = false;
      }
      // parked photons move to the wall queue while it has room
      const unsigned long long pm = __ballot(alive && parked);
      if (pm != 0ull) {
        const uint32_t k = min((uint32_t)__popcll(pm), (uint32_t)kWallCap - n_wall);
        const uint32_t rank = rank_below(pm);
        if (alive && parked && rank < k) {
          q_put_at<MULTI, REC, kWallCap>(wall + n_wall + rank, p, rng, lane_ev, gi, lane_k, false);
          alive = false;
          parked = false;
        }
        n_wall += k;
      }
      // Once the chunks and the backlog are used up, the wall branch runs as
      // soon as at least as many photons wait for it as are stepping (the
      // stepping ones go to the backlog), not only when every lane is idle:
      // a wave's last photons would otherwise wait in the queue for its
      // slowest in-column photon on every wall hit, which made the launch's
      // drain about 12 ms of a 163-ms launch (3 ms this way).
      const bool draining = pool_done && n_back == 0u;
      const uint32_t n_alive = (uint32_t)__popcll(__ballot(alive));
      if (n_wall >= 64u || (draining && n_wall != 0u && n_wall >= n_alive)) {
        progress = true;
        if (!swap_to_backlog<MULTI, REC>(wq, n_back, alive, parked, p, rng, lane_ev, gi, lane_k)) {
          if (lane == 0) atomicOr(a.error_flag, 2u);
          break;
        }
        const uint32_t take = min(n_wall, 64u);
        // the wall entries are this wave's LDS writes, which complete in
        // order; the barrier keeps the compiler from moving the reads
        // above them
        __builtin_amdgcn_wave_barrier();
        if ((uint32_t)lane < take) {
          bool pk;
          q_get_at<MULTI, REC, kWallCap>(wall + n_wall - take + (uint32_t)lane, p, rng, lane_ev,
                                         gi, lane_k, pk);
          alive = true;
        }
        n_wall -= take;
        PrptState st = prpt_initial();
        int r = kPrAlive;
        if (alive) r = prpt_loop(rng, p, geoms[MULTI ? gi : 0u], st, true, true);
        if (alive && r == kPrEnded) {
          finish_photon<MULTI, REC, RADIAL>(a, geoms, events, p, gi, lane_ev, lane_k, s_cnt, s_tile, side,
                                            tile_area, run);
          alive = false;
        }
        // refracted out of the column: the gap flight waits in the gap queue
        // (it holds at most 63 + 64 entries here: the gap phase below runs
        // whenever 64 wait)
        const unsigned long long gm = __ballot(alive && r == kPrGap);
        if (gm != 0ull) {
          if (alive && r == kPrGap) {
            const uint32_t slot = kGapBase + n_gap + rank_below(gm);
            q_put<MULTI, REC>(wq, slot, p, rng, lane_ev, gi, lane_k, false);
            q_put_state(wq, slot, st);
            alive = false;
          }
          n_gap += (uint32_t)__popcll(gm);
        }
      }
      // Gap flights: a pass takes up to 64 waiting photons and runs one
      // flight each (to the next column and its Fresnel decision); a photon
      // reflected off the outside of that column goes back to the gap queue
      // for its next flight.  Passes repeat while 64 wait, so the queue
      // holds at most 63 entries after this loop and 127 during it.  While
      // draining (as for the wall queue) they also run once at least as
      // many flights wait as photons step.
      bool failed = false;
      while (n_gap >= 64u || (n_gap != 0u && n_wall == 0u && pool_done && n_back == 0u &&
                              n_gap >= (uint32_t)__popcll(__ballot(alive)))) {
        progress = true;
        if (!swap_to_backlog<MULTI, REC>(wq, n_back, alive, parked, p, rng, lane_ev, gi, lane_k)) {
          failed = true;
          break;
        }
        const uint32_t take = min(n_gap, 64u);
        PrptState st = prpt_initial();
        __threadfence_block();
        if ((uint32_t)lane < take) {
          bool pk;
          const uint32_t slot = kGapBase + n_gap - take + (uint32_t)lane;
          q_get<MULTI, REC>(wq, slot, p, rng, lane_ev, gi, lane_k, pk);
          q_get_state(wq, slot, st);
          alive = true;
        }
        n_gap -= take;
        int r = kPrAlive;
        if (alive) r = prpt_loop(rng, p, geoms[MULTI ? gi : 0u], st, false, true);
        if (alive && r == kPrEnded) {
          finish_photon<MULTI, REC, RADIAL>(a, geoms, events, p, gi, lane_ev, lane_k, s_cnt, s_tile, side,
                                            tile_area, run);
          alive = false;
        }
        // the slots just read are free again: [n_gap, n_gap + take)
        const unsigned long long gm = __ballot(alive && r == kPrGap);
        if (gm != 0ull) {
          if (alive && r == kPrGap) {
            const uint32_t slot = kGapBase + n_gap + rank_below(gm);
            q_put<MULTI, REC>(wq, slot, p, rng, lane_ev, gi, lane_k, false);
            q_put_state(wq, slot, st);
            alive = false;
          }
          n_gap += (uint32_t)__popcll(gm);
        }
      }
      if (failed) {
        if (lane == 0) atomicOr(a.error_flag, 2u);
        break;
      }
      if (progress) {
        idle = 0;
      } else if (++idle > kMaxIdlePasses) {
        if (lane == 0) atomicOr(a.error_flag, 2u);
        break;
      }
    } else {
      // The wall-Fresnel branch (with the gap flights it leads to) is long
      // and taken by few lanes at a time: run it once enough lanes wait for
      // it, or when nothing else can run.
      const unsigned long long parked_mask = __ballot(parked);
      if (parked_mask != 0ull &&
          ((int)__popcll(parked_mask) >= a.park_lanes || __ballot(alive && !parked) == 0ull)) {
        if (parked) {
          const GeomDev& g = geoms[MULTI ? gi : 0u];
          ended = ba_prpt(rng, p, g);
          parked = false;
        }
      }
      if (ended) {
        finish_photon<MULTI, REC, RADIAL>(a, geoms, events, p, gi, lane_ev, lane_k, s_cnt, s_tile, side,
                                          tile_area, run);
        alive = false;
      }
    }
  }

  if (run[1]) flush_run<MULTI>(geoms, events, run[0], run[1]);
  __syncthreads();
  for (int i = threadIdx.x; i < a.n_geoms * kCounters; i += blockDim.x)
    if (s_cnt[i]) tally_add(&a.counters[i], (unsigned long long)s_cnt[i]);
  for (int t = 0; t < a.n_tiles; ++t) {
    const GeomDev& gt = geoms[t];
    const int i0 = tile_origin(gt.xdim, side), j0 = tile_origin(gt.ydim, side);
    for (int e = threadIdx.x; e < tile_area; e += blockDim.x) {
      const uint32_t v = s_tile[t * tile_area + e];
      if (v) {
        const int ti = e / side;
        tally_add(&gt.prf[(size_t)(i0 + ti) * gt.ydim + (j0 + e - ti * side)],
                  (unsigned long long)v);
      }
    }
  }
}

template <bool MULTI, bool REC, bool RADIAL>
static void launch_one(const TransportArgs& a, dim3 grid, hipStream_t s) {
  const size_t smem = (size_t)a.n_tiles * a.tile_side * a.tile_side * sizeof(uint32_t);
  if (a.wave_queue)
    hipLaunchKernelGGL((k_transport<MULTI, REC, RADIAL, true>), grid, dim3(kTransportThreads),
                       smem, s, a, a.geoms, a.events, a.chunk_off);
  else
    hipLaunchKernelGGL((k_transport<MULTI, REC, RADIAL, false>), grid, dim3(kTransportThreads),
                       smem, s, a, a.geoms, a.events, a.chunk_off);
}

hipError_t launch_transport(const TransportArgs& a, bool multi, bool record,
                            bool radial, int n_blocks, hipStream_t s) {
  const dim3 grid(n_blocks);
  if (radial) {
    if (multi) launch_one<true, false, true>(a, grid, s);
    else launch_one<false, false, true>(a, grid, s);
  } else if (multi) {
    if (record) launch_one<true, true, false>(a, grid, s);
    else launch_one<true, false, false>(a, grid, s);
  } else {
    if (record) launch_one<false, true, false>(a, grid, s);
    else launch_one<false, false, false>(a, grid, s);
  }
  return hipGetLastError();
}

template <bool MULTI, bool REC, bool RADIAL>
static int occupancy_of(size_t smem, bool queue) {
  int nb = 0;
  const hipError_t e =
      queue ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &nb, k_transport<MULTI, REC, RADIAL, true>, kTransportThreads, smem)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &nb, k_transport<MULTI, REC, RADIAL, false>, kTransportThreads, smem);
  if (e != hipSuccess) return 1;
  return nb > 0 ? nb : 1;
}

template <bool MULTI, bool REC, bool RADIAL, bool QUEUE>
static size_t static_lds_of() {
  hipFuncAttributes at{};
  if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(
                                    &k_transport<MULTI, REC, RADIAL, QUEUE>)) != hipSuccess)
    return 0;
  return at.sharedSizeBytes;
}

size_t transport_static_lds() {
  // the several-geometry variants hold the geometry table: the largest
  const size_t a = static_lds_of<true, false, false, true>();
  const size_t b = static_lds_of<true, true, false, true>();
  const size_t c = static_lds_of<true, false, true, true>();
  const size_t d = static_lds_of<true, false, false, false>();
  size_t m = a > b ? a : b;
  m = m > c ? m : c;
  return m > d ? m : d;
}

int transport_max_blocks_per_cu(bool multi, bool record, bool radial, size_t smem, bool queue) {
  if (radial)
    return multi ? occupancy_of<true, false, true>(smem, queue)
                 : occupancy_of<false, false, true>(smem, queue);
  if (multi)
    return record ? occupancy_of<true, true, false>(smem, queue)
                  : occupancy_of<true, false, false>(smem, queue);
  return record ? occupancy_of<false, true, false>(smem, queue)
                : occupancy_of<false, false, false>(smem, queue);
}

// ------------------------------------------------ RANECU trajectory mode

template <class M>
__global__ void __launch_bounds__(64) k_trace_ranecu(
    const fd2_event* events, int64_t n, int64_t event_base, int32_t seed_input,
    const GeomDev* geoms, unsigned long long* counters,
    const unsigned long long* photon_off, fd2_photon* records) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fd2_event ev = events[i];
  const GeomDev& g = geoms[ev.geom];
  uint32_t cnt[kCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
  RanecuRng rng;
  ranecu_init(rng, (int)(event_base + i), seed_input);
  StepSink nosink{nullptr, nullptr, 0, 0, false};
  Photon p;
  start_photon(p, ev);
  if (ev.N > 0) emit_first_dir<RanecuRng, M>(rng, p, g);  // driver, normal stays 0
  for (int32_t k = 0; k < ev.N; ++k) {
    // direction and normal carry over from the previous photon's end
    const float ux = p.ux, uy = p.uy, uz = p.uz, nx = p.nx, ny = p.ny, nz = p.nz;
    start_photon(p, ev);
    p.ux = ux; p.uy = uy; p.uz = uz;
    p.nx = nx; p.ny = ny; p.nz = nz;
    cnt[FD2_C_GENERATED]++;
    for (;;) {
      const int r = photon_step<RanecuRng, false, M>(rng, p, g, &nosink);
      if (r == kEnded) break;
      if (r == kTrapped) {
        cnt[FD2_C_REGEN]++;
        restart_trapped<RanecuRng, M>(rng, p, g, ev);
      }
    }
    cnt[fate_counter(p.fate)]++;
    const bool det = p.fate == kDetect || p.fate == kDetectOut;
    if (p.fate == kDetect) atomicAdd(&g.prf[pix_index(p.pix, g)], 1ull);
    if (det && ev.hist >= 1 && ev.hist <= g.num_hist) atomicAdd(&g.hist[ev.hist - 1], 1u);
    if (records) {
      fd2_photon rec;
      rec.fate = p.fate;
      rec.pix = p.fate == kDetect ? pix_index(p.pix, g) : -1;
      rec.x = det ? p.px : 0.0f;
      rec.y = det ? p.py : 0.0f;
      records[photon_off[i] + k] = rec;
    }
    // next photon's direction and normal draws (main/kernel_cuda_c_ver1_0.cu:578-602)
    emit_dir<RanecuRng, M>(rng, p, g);
    if (g.beta > 0.0f) {
      (void)rng.u();
      (void)rng.u();
    }
  }
  unsigned long long* c = counters + (size_t)ev.geom * kCounters;
  for (int j = 0; j < kCounters; ++j)
    if (cnt[j]) atomicAdd(&c[j], (unsigned long long)cnt[j]);
}

hipError_t launch_trace_ranecu(const fd2_event* events, int64_t n,
                               int64_t event_base, int32_t seed_input,
                               const GeomDev* geoms, int32_t n_geoms,
                               unsigned long long* counters,
                               const unsigned long long* photon_off,
                               fd2_photon* records, bool faithful, hipStream_t s) {
  (void)n_geoms;
  if (n <= 0) return hipSuccess;
  const int64_t nb = (n + 63) / 64;
  if (faithful)
    hipLaunchKernelGGL(k_trace_ranecu<MathFaithful>, dim3((unsigned)nb), dim3(64), 0, s, events,
                       n, event_base, seed_input, geoms, counters, photon_off, records);
  else
    hipLaunchKernelGGL(k_trace_ranecu<MathDevice>, dim3((unsigned)nb), dim3(64), 0, s, events,
                       n, event_base, seed_input, geoms, counters, photon_off, records);
  return hipGetLastError();
}

// ----------------------------------------------------- step trace (debug)

__global__ void __launch_bounds__(64) k_debug_steps(
    const fd2_event* events, uint64_t event_base, uint32_t seed,
    const GeomDev* geoms, int32_t n_photons, float* steps, int32_t* n_steps,
    int32_t max_steps) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const fd2_event ev = events[0];
  if (k >= n_photons || k >= ev.N) return;
  const GeomDev& g = geoms[ev.geom];
  StepSink sink{steps, n_steps, max_steps, k, true};
  Photon p;
  CounterRng rng;
  start_photon(p, ev);
  rng.seed(seed, event_base, (uint32_t)k);
  counter_initial_dir(rng, p, (uint32_t)k, g);
  for (;;) {
    const int r = photon_step<CounterRng, true>(rng, p, g, &sink);
    if (r == kEnded) break;
    if (r == kTrapped) restart_trapped(rng, p, g, ev);
  }
}

hipError_t launch_debug_steps(const fd2_event* events, uint64_t event_base,
                              uint32_t seed, const GeomDev* geoms,
                              int32_t n_photons, float* steps,
                              int32_t* n_steps, int32_t max_steps,
                              hipStream_t s) {
  if (n_photons <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_steps, dim3((n_photons + 63) / 64), dim3(64), 0, s,
                     events, event_base, seed, geoms, n_photons, steps, n_steps,
                     max_steps);
  return hipGetLastError();
}

// Device arithmetic on its own: y[i] = f(x[i]) for the functions the
// transport uses, numbered as FD2_MATH_* (the oracle evaluates the same
// functions on the CPU, so the two can be compared bit for bit).
__global__ void k_debug_math(int fn, const float* x, float* y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  float r;
  float sn, cs;
  switch (fn) {
    case FD2_MATH_SIN: r = fd2_sin(v); break;
    case FD2_MATH_COS: r = fd2_cos(v); break;
    case FD2_MATH_LOG: r = fd2_log(v); break;
    case FD2_MATH_ASIN: r = fd2_asin(v); break;
    case FD2_MATH_ACOS: r = fd2_acos(v); break;
    case FD2_MATH_SIN_TURNS: fd2_sincos_turns(v, &sn, &cs); r = sn; break;
    case FD2_MATH_COS_TURNS: fd2_sincos_turns(v, &sn, &cs); r = cs; break;
    case FD2_MATH_RSQRT: r = fd2_rsqrt(v); break;
    case FD2_MATH_SQRT_UNIT: r = fd2_sqrt_unit(v); break;
    case FD2_MATH_ANGLE_OVER_157: r = angle_over_157(v) ? 1.0f : 0.0f; break;
    case FD2_MATH_ACOSF_LIBC: r = fd2x_acosf(v); break;
    case FD2_MATH_ASINF_LIBC: r = fd2x_asinf(v); break;
    case FD2_MATH_FF_SIN: r = fd2x_sin(v).hi; break;
    case FD2_MATH_FF_COS: r = fd2x_cos(v).hi; break;
    case FD2_MATH_FF_LOG: r = fd2x_log(v).hi; break;
    default: r = 0.0f; break;
  }
  y[i] = r;
}

hipError_t launch_debug_math(int fn, const float* x, float* y, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fn, x, y,
                     n);
  return hipGetLastError();
}

// ------------------------------------------------------------- fills

__global__ void k_fill_u64(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}
__global__ void k_fill_u32(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

hipError_t launch_fill_u64(unsigned long long* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t nb = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)nb), dim3(256), 0, s, p, n);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t nb = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)nb), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0) p[0] = v;
}

hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0) p[0] = v;
}

hipError_t launch_set_u64(unsigned long long* p, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

}  // namespace fd2
