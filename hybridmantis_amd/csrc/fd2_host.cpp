// fd2_host.cpp -- host side of libfd2: context, event staging, device work
// queue, kernel launches and tallies.  Replaces the state machine of
// gpuoptical_ (main/hybridMANTIS_cuda_ver1_0.cu:107-447): device memory is
// owned by a context instead of being stashed in COMMON /gpumemaddr/, and
// errors come back as status codes instead of aborting.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "fd2_internal.h"

using namespace fd2;

namespace {

#define FD2_HIP(call)                      \
  do {                                     \
    hipError_t e_ = (call);                \
    if (e_ != hipSuccess) return FD2_E_DEVICE; \
  } while (0)

int ilog2(int v) {
  int l = 0;
  while ((1 << (l + 1)) <= v) ++l;
  return l;
}

int prf_side(double lb, double ub, int pixel) {
  // xdim = ceil((ub - lb)/pixel) in float, as the reference drivers compute it
  // (main/hybridMANTIS_c_ver1_0.c:163-164)
  const float w = (float)ub - (float)lb;
  return (int)std::ceil(w / (float)pixel);
}

int check_params(const fd2_params& p) {
  if (!(p.detheight > 0) || !(p.detradius > 0) || !(p.detx > 0) || !(p.dety > 0))
    return FD2_E_INVAL;
  if (!(p.detnC > 0) || !(p.detnIC > 0) || p.detpixel <= 0) return FD2_E_INVAL;
  if (!(p.detdmax >= p.detdmin)) return FD2_E_INVAL;
  if (p.mynumhist < 0) return FD2_E_INVAL;
  const int xd = prf_side(p.detlboundx, p.detuboundx, p.detpixel);
  const int yd = prf_side(p.detlboundy, p.detuboundy, p.detpixel);
  if (xd <= 0 || yd <= 0 || xd > FD2_MAX_PRF_SIDE || yd > FD2_MAX_PRF_SIDE)
    return FD2_E_INVAL;
  return FD2_OK;
}

// smallest float >= c, largest float <= c
float float_ge(double c) {
  float f = (float)c;
  if ((double)f < c) f = std::nextafter(f, INFINITY);
  return f;
}
float float_le(double c) {
  float f = (float)c;
  if ((double)f > c) f = std::nextafter(f, -INFINITY);
  return f;
}

void derive(const fd2_params& p, GeomDev& g) {
  std::memset(&g, 0, sizeof g);
  g.xdet = (float)p.detx;
  g.ydet = (float)p.dety;
  g.H = (float)p.detheight;
  g.R = (float)p.detradius;
  g.n1 = (float)p.detnC;
  g.n2 = (float)p.detnIC;
  g.top_absfrac = (float)p.dettop;
  g.bulk = (float)p.detbulk;
  g.beta = (float)p.detbeta;
  g.d_min = (float)p.detdmin;
  g.d_max = (float)p.detdmax;
  g.lbx = (float)p.detlboundx;
  g.lby = (float)p.detlboundy;
  g.ubx = (float)p.detuboundx;
  g.uby = (float)p.detuboundy;
  g.sensor_refl = (float)p.detsensorRefl;
  g.half_h = g.H / 2.0f;
  g.depth20 = g.H * 0.2f;
  g.neg_inv_bulk = g.bulk > 0.0f ? -1.0f / g.bulk : 0.0f;
  g.ratio12 = g.n1 / g.n2;
  g.ratio21 = g.n2 / g.n1;
  g.gap_fresnel = (g.ratio12 < 1.57f) ? 0 : 1;
  g.dspan = g.d_max - g.d_min;
  // exact float thresholds for the reference's double comparisons
  const double eps = 8.1929093e-6;
  g.eps_lt = float_ge(eps);
  g.eps_gt = float_le(eps);
  g.norm_lo = float_ge(1.0 - eps);
  g.norm_hi = float_le(1.0 + eps);
  g.r_nudge = float_le((double)g.R - 1e-5);
  // cos(asin(r)) = sqrt(1 - r^2); c1 < C  <=>  c1 < float_ge(C) for float c1
  g.cos_tir_in = float_ge(std::sqrt(1.0 - (double)g.ratio21 * (double)g.ratio21));
  g.cos_tir_gap = g.ratio12 <= 1.0f
                      ? float_ge(std::sqrt(1.0 - (double)g.ratio12 * (double)g.ratio12))
                      : NAN;
  // theta1 > asin(r) (a double) <=> theta1 > the largest float <= asin(r)
  g.t1_tir_in = float_le(std::asin((double)g.ratio21));
  g.t1_tir_gap = g.ratio12 <= 1.0f ? float_le(std::asin((double)g.ratio12)) : NAN;
  g.pixel = p.detpixel;
  g.inv_pixel = (p.detpixel & (p.detpixel - 1)) == 0 ? 1.0f / (float)p.detpixel : 0.0f;
  g.xdim = prf_side(p.detlboundx, p.detuboundx, p.detpixel);
  g.ydim = prf_side(p.detlboundy, p.detuboundy, p.detpixel);
  g.num_hist = p.mynumhist;
}

int grow(fd2_ctx* c, size_t n_events) {
  if (n_events > c->cap_events) {
    size_t cap = n_events + n_events / 4 + 1024;
    if (c->d_events) (void)hipFree(c->d_events);
    c->d_events = nullptr;
    c->cap_events = 0;
    FD2_HIP(hipMalloc(&c->d_events, cap * sizeof(fd2_event)));
    c->cap_events = cap;
  }
  if (n_events + 1 > c->cap_off) {
    size_t cap = n_events + n_events / 4 + 1025;
    if (c->d_chunk_off) (void)hipFree(c->d_chunk_off);
    if (c->d_photon_off) (void)hipFree(c->d_photon_off);
    c->d_chunk_off = nullptr;
    c->d_photon_off = nullptr;
    c->cap_off = 0;
    FD2_HIP(hipMalloc(&c->d_chunk_off, cap * sizeof(uint32_t)));
    FD2_HIP(hipMalloc(&c->d_photon_off, cap * sizeof(unsigned long long)));
    c->cap_off = cap;
  }
  const size_t nb = (n_events + kScanTile - 1) / kScanTile + 1;
  if (nb > c->cap_blocks) {
    size_t cap = nb + 64;
    if (c->d_block_chunks) (void)hipFree(c->d_block_chunks);
    if (c->d_block_photons) (void)hipFree(c->d_block_photons);
    c->d_block_chunks = nullptr;
    c->d_block_photons = nullptr;
    c->cap_blocks = 0;
    FD2_HIP(hipMalloc(&c->d_block_chunks, cap * sizeof(uint32_t)));
    FD2_HIP(hipMalloc(&c->d_block_photons, cap * sizeof(unsigned long long)));
    c->cap_blocks = cap;
  }
  return FD2_OK;
}

int grow_stage(fd2_ctx* c, size_t n_events) {
  if (n_events <= c->cap_stage) return FD2_OK;
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  c->h_stage = nullptr;
  c->cap_stage = 0;
  const size_t cap = n_events + n_events / 4 + 1024;
  FD2_HIP(hipHostMalloc(&c->h_stage, cap * sizeof(fd2_event), hipHostMallocDefault));
  c->cap_stage = cap;
  return FD2_OK;
}

int validate_host(const fd2_ctx* c, const fd2_event* ev, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    if (ev[i].N > FD2_MAX_PHOTONS_PER_EVENT) return FD2_E_TOO_MANY;
    if (ev[i].N < 0) return FD2_E_INVAL;
    if (ev[i].geom >= (uint32_t)c->n_geoms) return FD2_E_INVAL;
  }
  return FD2_OK;
}

// LDS PRF tiles for a tuning value: -1 none, 0 the default side, else the
// side; as many geometries get a tile as fit the per-block budget.
void tile_shape(int lds_tile, int n_geoms, int32_t* side, int32_t* n_tiles) {
  const int s = lds_tile < 0 ? 0 : (lds_tile == 0 ? kTileDefault : lds_tile);
  *side = s;
  if (s == 0) {
    *n_tiles = 0;
    return;
  }
  const int fit = (s * s > kTileBudget) ? 1 : kTileBudget / (s * s);
  *n_tiles = n_geoms < fit ? n_geoms : fit;
}

// Work-queue buffers of one launch (the context's own, or a graph's).
struct QueueBufs {
  uint32_t* chunk_off;
  unsigned long long* photon_off;
  uint32_t* block_chunks;
  unsigned long long* block_photons;
};

// Arguments and grid of the queue build + transport of n device events.
struct Launch {
  QueueArgs q{};
  TransportArgs a{};
  int blocks = 0;
  bool multi = false, record = false, radial = false;
};

// Queue cursor and error flag of a launch other than the context's own,
// with the transport in park-in-place mode (no per-wave queues).
struct NoQueueScratch {
  uint32_t* cursor;
  uint32_t* error_flag;
};

Launch make_launch(const fd2_ctx* c, const QueueBufs& b, const fd2_event* d_ev, int64_t n,
                   fd2_photon* d_records, uint32_t* d_radial,
                   const NoQueueScratch* own = nullptr) {
  Launch L;
  QueueArgs& q = L.q;
  q.events = d_ev;
  q.chunk_off = b.chunk_off;
  q.photon_off = b.photon_off;
  q.block_chunks = b.block_chunks;
  q.block_photons = b.block_photons;
  q.cursor = own ? own->cursor : c->d_cursor;
  q.error_flag = own ? own->error_flag : c->d_error;
  q.n_events = n;
  q.n_geoms = c->n_geoms;
  q.chunk_log2 = ilog2(c->tune.chunk);
  q.n_blocks = (int)((n + kScanTile - 1) / kScanTile);
  q.max_chunks = (uint64_t)c->tune.max_chunks;

  TransportArgs& a = L.a;
  a.events = d_ev;
  a.chunk_off = b.chunk_off;
  a.photon_off = b.photon_off;
  a.geoms = c->d_geoms;
  a.counters = c->d_counters;
  a.cursor = q.cursor;
  a.records = d_records;
  a.radial = d_radial;
  a.event_base_dev = nullptr;
  const bool queue = !own && c->tune.wall_queue >= 0 && c->d_wave_queue != nullptr;
  a.wave_queue = queue ? c->d_wave_queue : nullptr;
  a.error_flag = q.error_flag;
  a.n_events = n;
  a.seed = c->seed;
  a.event_base = (uint64_t)c->next_event;
  a.n_geoms = c->n_geoms;
  a.chunk_log2 = q.chunk_log2;
  a.park_lanes = c->tune.park_lanes;
  a.refill_lanes = c->tune.refill_lanes;
  tile_shape(c->tune.lds_tile, c->n_geoms, &a.tile_side, &a.n_tiles);

  L.multi = c->n_geoms > 1;
  L.record = d_records != nullptr;
  L.radial = d_radial != nullptr;
  const int per_cu = c->tune.waves_per_cu / (kTransportThreads / 64);
  const size_t smem = (size_t)a.n_tiles * a.tile_side * a.tile_side * sizeof(uint32_t);
  const int occ = transport_max_blocks_per_cu(L.multi, L.record, L.radial, smem, queue);
  L.blocks = c->n_cu * (per_cu < 1 ? 1 : (per_cu > occ ? occ : per_cu));
  if (queue && L.blocks * (kTransportThreads / 64) > c->max_waves)
    L.blocks = c->max_waves / (kTransportThreads / 64);
  return L;
}

hipError_t issue(const Launch& L, hipStream_t s) {
  hipError_t e = launch_queue_build(L.q, s);
  if (e != hipSuccess) return e;
  return launch_transport(L.a, L.multi, L.record, L.radial, L.blocks, s);
}

// queue build + transport on events already in device memory.  With the
// events' host copy h_ev, the events are split into consecutive launches of
// at most tune.max_chunks work items each (each launch numbers its events
// on from the last, so the photons are those of one launch); without it the
// queue build rejects a launch above the limit (error bit 4).
int enqueue(fd2_ctx* c, const fd2_event* d_ev, int64_t n, bool record,
            fd2_photon* d_records, uint32_t* d_radial = nullptr,
            const fd2_event* h_ev = nullptr) {
  if (n <= 0) return FD2_OK;
  int rc = grow(c, (size_t)n);
  if (rc) return rc;
  const QueueBufs b{c->d_chunk_off, c->d_photon_off, c->d_block_chunks, c->d_block_photons};
  const uint64_t chunk = (uint64_t)c->tune.chunk, limit = (uint64_t)c->tune.max_chunks;
  FD2_HIP(hipEventRecord(c->ev_start, c->stream));
  uint64_t photons_before = 0;
  for (int64_t i0 = 0; i0 < n;) {
    int64_t i1 = n;
    uint64_t photons = 0;
    if (h_ev) {
      uint64_t items = 0;
      for (i1 = i0; i1 < n; ++i1) {
        const uint64_t k = ((uint64_t)h_ev[i1].N + chunk - 1) / chunk;
        if (items + k > limit && i1 > i0) break;
        items += k;
        photons += (uint64_t)h_ev[i1].N;
      }
    }
    const Launch L =
        make_launch(c, b, d_ev + i0, i1 - i0,
                    record ? d_records + photons_before : nullptr,
                    d_radial ? d_radial + (size_t)i0 * FD2_RADIAL_BINS : nullptr);
    FD2_HIP(issue(L, c->stream));
    c->next_event += i1 - i0;
    photons_before += photons;
    i0 = i1;
  }
  FD2_HIP(hipEventRecord(c->ev_stop, c->stream));
  c->timed = true;
  return FD2_OK;
}

// bit 0: an event failed validation in the queue build; bit 1: a wave of the
// persistent kernel hit its loop limit (results incomplete); bit 2: a launch
// had more work items than tune.max_chunks (nothing of it ran)
int check_error_flag(fd2_ctx* c) {
  uint32_t flag = 0;
  FD2_HIP(hipMemcpyAsync(&flag, c->d_error, sizeof flag, hipMemcpyDeviceToHost, c->stream));
  FD2_HIP(hipStreamSynchronize(c->stream));
  if (flag) {
    FD2_HIP(launch_fill_u32(c->d_error, 1, c->stream));
    FD2_HIP(hipStreamSynchronize(c->stream));
    return (flag & 2u) ? FD2_E_DEVICE : FD2_E_INVAL;
  }
  return FD2_OK;
}

// ------------------------------------------- allocation-free submit scratch
//
// fd2_submit_device_noalloc keeps its work-queue scratch -- cursor, error
// flag, block totals and the per-event chunk and photon offsets of at most
// FD2_NOALLOC_MAX_EVENTS events -- in this file-scope device array, 12.6 MB
// that each GPU the library runs on holds until the process ends.  Host
// code uses the array only through the address hipGetSymbolAddress gives
// on each device (its own copy there), looked up once per device at
// fd2_create with that device current.
constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr int64_t kNaEvents = FD2_NOALLOC_MAX_EVENTS;
constexpr int64_t kNaBlocks = (kNaEvents + kScanTile - 1) / kScanTile;
constexpr size_t kNaBlockChunks = 256;  // [0, 256): cursor, error flag
constexpr size_t kNaBlockPhotons = align256(kNaBlockChunks + sizeof(uint32_t) * kNaBlocks);
constexpr size_t kNaChunkOff = align256(kNaBlockPhotons + sizeof(unsigned long long) * kNaBlocks);
constexpr size_t kNaPhotonOff = align256(kNaChunkOff + sizeof(uint32_t) * (kNaEvents + 1));
constexpr size_t kNaBytes = align256(kNaPhotonOff + sizeof(unsigned long long) * (kNaEvents + 1));
constexpr int kNaMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNaBytes];

std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kNaMaxDevices] = {};

// the device address of `scratch` on `device` (current), looked up once
int scratch_setup(int device) {
  if (device < 0 || device >= kNaMaxDevices) return FD2_OK;  // that device: no noalloc
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return FD2_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FD2_E_DEVICE;
  g_scratch_dev[device] = (unsigned char*)p;
  return FD2_OK;
}

unsigned char* scratch_of(int device) {
  if (device < 0 || device >= kNaMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

}  // namespace

namespace fd2 {

int submit_radial_async(fd2_ctx* c, const fd2_event* events, int64_t n, uint32_t* d_radial) {
  if (!c || n < 0 || (n > 0 && (!events || !d_radial))) return FD2_E_INVAL;
  if (n == 0) return FD2_OK;
  int rc = validate_host(c, events, n);
  if (rc) return rc;
  FD2_HIP(hipSetDevice(c->device));
  // the previous submit may still read the staging buffer
  FD2_HIP(hipStreamSynchronize(c->stream));
  rc = grow_stage(c, (size_t)n);
  if (rc) return rc;
  rc = grow(c, (size_t)n);
  if (rc) return rc;
  std::memcpy(c->h_stage, events, (size_t)n * sizeof(fd2_event));
  FD2_HIP(hipMemcpyAsync(c->d_events, c->h_stage, (size_t)n * sizeof(fd2_event),
                         hipMemcpyHostToDevice, c->stream));
  FD2_HIP(launch_fill_u32(d_radial, (size_t)n * FD2_RADIAL_BINS, c->stream));
  return enqueue(c, c->d_events, n, false, nullptr, d_radial, c->h_stage);
}

}  // namespace fd2

extern "C" {

const char* fd2_strerror(int status) {
  switch (status) {
    case FD2_OK: return "ok";
    case FD2_E_INVAL: return "invalid argument";
    case FD2_E_NOMEM: return "out of memory";
    case FD2_E_DEVICE: return "HIP runtime error";
    case FD2_E_TOO_MANY: return "event has more than 900000 optical photons";
    case FD2_E_NO_DEVICE: return "no usable GPU";
    case FD2_E_STATE: return "invalid call order";
    case FD2_E_IO: return "file error";
    default: return "unknown error";
  }
}

int fd2_create(const fd2_params* params, int n_geoms, int device, fd2_ctx** out) {
  if (!params || !out || n_geoms < 1 || n_geoms > FD2_MAX_GEOMS) return FD2_E_INVAL;
  *out = nullptr;
  for (int i = 0; i < n_geoms; ++i)
    if (check_params(params[i])) return FD2_E_INVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FD2_E_NO_DEVICE;
  if (device < 0 || device >= ndev) return FD2_E_INVAL;
  fd2_ctx* c = new (std::nothrow) fd2_ctx();
  if (!c) return FD2_E_NOMEM;
  c->device = device;
  c->n_geoms = n_geoms;
  c->tune.chunk = 128;
  c->tune.waves_per_cu = 20;
  c->tune.lds_tile = 0;
  c->tune.park_lanes = 32;
  c->tune.refill_lanes = 4;
  c->tune.wall_queue = 0;
  c->tune.max_chunks = FD2_MAX_CHUNKS;
  int rc = FD2_OK;
  do {
    if (hipSetDevice(device) != hipSuccess) { rc = FD2_E_DEVICE; break; }
    if ((rc = scratch_setup(device)) != FD2_OK) break;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = FD2_E_DEVICE; break; }
    c->n_cu = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    c->lds_static = transport_static_lds();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { rc = FD2_E_DEVICE; break; }
    if (hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_stop) != hipSuccess) {
      rc = FD2_E_DEVICE;
      break;
    }
    for (int i = 0; i < n_geoms; ++i) {
      c->params[i] = params[i];
      derive(params[i], c->hgeom[i]);
      const size_t npix = (size_t)c->hgeom[i].xdim * c->hgeom[i].ydim;
      const size_t nh = params[i].mynumhist > 0 ? (size_t)params[i].mynumhist : 1;
      if (hipMalloc(&c->d_prf[i], npix * sizeof(unsigned long long)) != hipSuccess ||
          hipMalloc(&c->d_hist[i], nh * sizeof(uint32_t)) != hipSuccess) {
        rc = FD2_E_NOMEM;
        break;
      }
      c->hgeom[i].prf = c->d_prf[i];
      c->hgeom[i].hist = c->d_hist[i];
    }
    if (rc) break;
    if (hipMalloc(&c->d_geoms, sizeof(GeomDev) * n_geoms) != hipSuccess ||
        hipMalloc(&c->d_counters, sizeof(unsigned long long) * kCounters * n_geoms) != hipSuccess ||
        hipMalloc(&c->d_cursor, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&c->d_error, sizeof(uint32_t)) != hipSuccess) {
      rc = FD2_E_NOMEM;
      break;
    }
    if (hipMemcpy(c->d_geoms, c->hgeom, sizeof(GeomDev) * n_geoms, hipMemcpyHostToDevice) !=
        hipSuccess) {
      rc = FD2_E_DEVICE;
      break;
    }
    // per-wave photon queues for the largest resident grid of the queued kernel
    {
      int occ = 1;
      for (int v = 0; v < 5; ++v) {
        const int o = transport_max_blocks_per_cu(v == 1 || v == 3, v == 2 || v == 3, v == 4, 0,
                                                  true);
        if (o > occ) occ = o;
      }
      c->max_waves = c->n_cu * occ * (kTransportThreads / 64);
      const size_t qbytes = (size_t)c->max_waves * kQWords * kQSlots * sizeof(uint32_t);
      if (hipMalloc(&c->d_wave_queue, qbytes) != hipSuccess) { rc = FD2_E_NOMEM; break; }
    }
    if (launch_fill_u32(c->d_error, 1, c->stream) != hipSuccess) { rc = FD2_E_DEVICE; break; }
    if (launch_fill_u32(c->d_cursor, 1, c->stream) != hipSuccess) { rc = FD2_E_DEVICE; break; }
    rc = fd2_reset(c);
  } while (0);
  if (rc) {
    fd2_destroy(c);
    return rc;
  }
  *out = c;
  return FD2_OK;
}

void fd2_destroy(fd2_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < FD2_MAX_GEOMS; ++i) {
    if (c->d_prf[i]) (void)hipFree(c->d_prf[i]);
    if (c->d_hist[i]) (void)hipFree(c->d_hist[i]);
  }
  void* bufs[] = {c->d_geoms, c->d_wave_queue, c->d_counters, c->d_events, c->d_chunk_off,
                  c->d_photon_off, c->d_block_chunks, c->d_block_photons,
                  c->d_cursor, c->d_error};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->ev_start) (void)hipEventDestroy(c->ev_start);
  if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int fd2_set_seed(fd2_ctx* c, uint32_t seed) {
  if (!c) return FD2_E_INVAL;
  c->seed = seed;
  return FD2_OK;
}

int fd2_set_tuning(fd2_ctx* c, const fd2_tuning* t) {
  if (!c || !t) return FD2_E_INVAL;
  fd2_tuning n = c->tune;
  if (t->chunk) {
    if (t->chunk < 1 || t->chunk > (1 << 16) || (t->chunk & (t->chunk - 1))) return FD2_E_INVAL;
    n.chunk = t->chunk;
  }
  if (t->waves_per_cu) {
    if (t->waves_per_cu < 1 || t->waves_per_cu > 32) return FD2_E_INVAL;
    n.waves_per_cu = t->waves_per_cu;
  }
  if (t->lds_tile < -1 || t->lds_tile > kTileMax) return FD2_E_INVAL;
  {  // the tiles and the kernel's static LDS must fit one workgroup's LDS
    int32_t side = 0, n_tiles = 0;
    tile_shape(t->lds_tile, c->n_geoms, &side, &n_tiles);
    const size_t tiles = (size_t)n_tiles * side * side * sizeof(uint32_t);
    if (c->lds_per_block && tiles + c->lds_static > c->lds_per_block) return FD2_E_INVAL;
  }
  n.lds_tile = t->lds_tile;
  if (t->park_lanes) {
    if (t->park_lanes < 1 || t->park_lanes > 64) return FD2_E_INVAL;
    n.park_lanes = t->park_lanes;
  }
  if (t->refill_lanes) {
    if (t->refill_lanes < 1 || t->refill_lanes > 64) return FD2_E_INVAL;
    n.refill_lanes = t->refill_lanes;
  }
  if (t->wall_queue < -1 || t->wall_queue > 1) return FD2_E_INVAL;
  n.wall_queue = t->wall_queue;
  if (t->max_chunks) {
    if (t->max_chunks < 1 || t->max_chunks > FD2_MAX_CHUNKS) return FD2_E_INVAL;
    n.max_chunks = t->max_chunks;
  }
  c->tune = n;
  return FD2_OK;
}

int fd2_get_tuning(const fd2_ctx* c, fd2_tuning* t) {
  if (!c || !t) return FD2_E_INVAL;
  *t = c->tune;
  return FD2_OK;
}

int fd2_reset(fd2_ctx* c) {
  if (!c) return FD2_E_INVAL;
  FD2_HIP(hipSetDevice(c->device));
  for (int i = 0; i < c->n_geoms; ++i) {
    const size_t npix = (size_t)c->hgeom[i].xdim * c->hgeom[i].ydim;
    const size_t nh = c->params[i].mynumhist > 0 ? (size_t)c->params[i].mynumhist : 1;
    FD2_HIP(launch_fill_u64(c->d_prf[i], npix, c->stream));
    FD2_HIP(launch_fill_u32(c->d_hist[i], nh, c->stream));
  }
  FD2_HIP(launch_fill_u64(c->d_counters, (size_t)kCounters * c->n_geoms, c->stream));
  FD2_HIP(hipStreamSynchronize(c->stream));
  c->next_event = 0;
  return FD2_OK;
}

int fd2_set_next_event(fd2_ctx* c, int64_t first) {
  if (!c || first < 0) return FD2_E_INVAL;
  c->next_event = first;
  return FD2_OK;
}

int fd2_submit(fd2_ctx* c, const fd2_event* events, int64_t n) {
  if (!c || n < 0 || (n > 0 && !events)) return FD2_E_INVAL;
  if (n == 0) return FD2_OK;
  int rc = validate_host(c, events, n);
  if (rc) return rc;
  FD2_HIP(hipSetDevice(c->device));
  // the previous submit may still read the staging buffer
  FD2_HIP(hipStreamSynchronize(c->stream));
  rc = grow_stage(c, (size_t)n);
  if (rc) return rc;
  rc = grow(c, (size_t)n);
  if (rc) return rc;
  std::memcpy(c->h_stage, events, (size_t)n * sizeof(fd2_event));
  FD2_HIP(hipMemcpyAsync(c->d_events, c->h_stage, (size_t)n * sizeof(fd2_event),
                         hipMemcpyHostToDevice, c->stream));
  return enqueue(c, c->d_events, n, false, nullptr, nullptr, c->h_stage);
}

int fd2_submit_device(fd2_ctx* c, const fd2_event* d_events, int64_t n) {
  if (!c || n < 0 || (n > 0 && !d_events)) return FD2_E_INVAL;
  FD2_HIP(hipSetDevice(c->device));
  return enqueue(c, d_events, n, false, nullptr);
}

int fd2_submit_device_noalloc(fd2_ctx* c, const fd2_event* d_events, int64_t n) {
  if (!c || n < 0 || (n > 0 && !d_events)) return FD2_E_INVAL;
  if (n > FD2_NOALLOC_MAX_EVENTS) return FD2_E_INVAL;
  unsigned char* s = scratch_of(c->device);
  if (!s) return FD2_E_STATE;
  if (n == 0) return FD2_OK;
  FD2_HIP(hipSetDevice(c->device));
  const NoQueueScratch own{(uint32_t*)s, (uint32_t*)s + 1};
  const QueueBufs b{(uint32_t*)(s + kNaChunkOff), (unsigned long long*)(s + kNaPhotonOff),
                    (uint32_t*)(s + kNaBlockChunks), (unsigned long long*)(s + kNaBlockPhotons)};
  const Launch L = make_launch(c, b, d_events, n, nullptr, nullptr, &own);
  // the array keeps the last call's contents: the error flag starts at zero
  // here, the queue build zeroes the cursor, and the offsets are rewritten
  FD2_HIP(launch_fill_u32(own.error_flag, 1, c->stream));
  FD2_HIP(hipEventRecord(c->ev_start, c->stream));
  FD2_HIP(issue(L, c->stream));
  FD2_HIP(hipEventRecord(c->ev_stop, c->stream));
  c->timed = true;
  c->next_event += n;
  uint32_t flag = 0;
  FD2_HIP(hipMemcpyAsync(&flag, own.error_flag, sizeof flag, hipMemcpyDeviceToHost, c->stream));
  FD2_HIP(hipStreamSynchronize(c->stream));
  if (flag) return (flag & 2u) ? FD2_E_DEVICE : FD2_E_INVAL;
  return FD2_OK;
}


int fd2_sync(fd2_ctx* c) {
  if (!c) return FD2_E_INVAL;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  return check_error_flag(c);
}

int fd2_submit_record(fd2_ctx* c, const fd2_event* events, int64_t n,
                      fd2_photon* records) {
  if (!c || n < 0 || (n > 0 && (!events || !records))) return FD2_E_INVAL;
  if (n == 0) return FD2_OK;
  int rc = validate_host(c, events, n);
  if (rc) return rc;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  int64_t total = 0;
  for (int64_t i = 0; i < n; ++i) total += events[i].N;
  rc = grow(c, (size_t)n);
  if (rc) return rc;
  FD2_HIP(hipMemcpyAsync(c->d_events, events, (size_t)n * sizeof(fd2_event),
                         hipMemcpyHostToDevice, c->stream));
  fd2_photon* d_rec = nullptr;
  if (total > 0) FD2_HIP(hipMalloc(&d_rec, (size_t)total * sizeof(fd2_photon)));
  rc = enqueue(c, c->d_events, n, true, d_rec, nullptr, events);
  if (rc == FD2_OK && total > 0)
    rc = hipMemcpyAsync(records, d_rec, (size_t)total * sizeof(fd2_photon),
                        hipMemcpyDeviceToHost, c->stream) == hipSuccess ? FD2_OK : FD2_E_DEVICE;
  if (hipStreamSynchronize(c->stream) != hipSuccess) rc = FD2_E_DEVICE;
  if (d_rec) (void)hipFree(d_rec);
  if (rc) return rc;
  return check_error_flag(c);
}

int fd2_submit_radial(fd2_ctx* c, const fd2_event* events, int64_t n,
                      uint32_t* radial) {
  if (!c || n < 0 || (n > 0 && (!events || !radial))) return FD2_E_INVAL;
  if (n == 0) return FD2_OK;
  int rc = validate_host(c, events, n);
  if (rc) return rc;
  FD2_HIP(hipSetDevice(c->device));
  const size_t nbins = (size_t)n * FD2_RADIAL_BINS;
  uint32_t* d_rad = nullptr;
  FD2_HIP(hipMalloc(&d_rad, nbins * sizeof(uint32_t)));
  rc = submit_radial_async(c, events, n, d_rad);
  if (!rc && hipMemcpyAsync(radial, d_rad, nbins * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            c->stream) != hipSuccess)
    rc = FD2_E_DEVICE;
  if (hipStreamSynchronize(c->stream) != hipSuccess) rc = FD2_E_DEVICE;
  (void)hipFree(d_rad);
  if (rc) return rc;
  return check_error_flag(c);
}

static int trace_ranecu(fd2_ctx* c, const fd2_event* events, int64_t n,
                        int64_t event_base, int32_t seed_input, fd2_photon* records,
                        bool faithful) {
  if (!c || n < 0 || (n > 0 && !events)) return FD2_E_INVAL;
  if (n == 0) return FD2_OK;
  // the reference's init_PRNG takes the event index as an int
  if (event_base < 0 || event_base > (int64_t)INT32_MAX - n + 1) return FD2_E_INVAL;
  int rc = validate_host(c, events, n);
  if (rc) return rc;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  rc = grow(c, (size_t)n);
  if (rc) return rc;
  unsigned long long* h_off = (unsigned long long*)std::malloc((size_t)(n + 1) * sizeof(unsigned long long));
  if (!h_off) return FD2_E_NOMEM;
  unsigned long long acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    h_off[i] = acc;
    acc += (unsigned long long)events[i].N;
  }
  h_off[n] = acc;
  fd2_photon* d_rec = nullptr;
  rc = FD2_OK;
  if (hipMemcpyAsync(c->d_events, events, (size_t)n * sizeof(fd2_event), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->d_photon_off, h_off, (size_t)(n + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = FD2_E_DEVICE;
  if (!rc && records && acc > 0 && hipMalloc(&d_rec, acc * sizeof(fd2_photon)) != hipSuccess) rc = FD2_E_NOMEM;
  if (!rc && launch_trace_ranecu(c->d_events, n, event_base, seed_input, c->d_geoms, c->n_geoms,
                                 c->d_counters, c->d_photon_off, d_rec, faithful,
                                 c->stream) != hipSuccess)
    rc = FD2_E_DEVICE;
  if (!rc && d_rec &&
      hipMemcpyAsync(records, d_rec, acc * sizeof(fd2_photon), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = FD2_E_DEVICE;
  if (hipStreamSynchronize(c->stream) != hipSuccess) rc = FD2_E_DEVICE;
  if (d_rec) (void)hipFree(d_rec);
  std::free(h_off);
  return rc;
}

int fd2_trace_ranecu(fd2_ctx* c, const fd2_event* events, int64_t n,
                     int64_t event_base, int32_t seed_input, fd2_photon* records) {
  return trace_ranecu(c, events, n, event_base, seed_input, records, false);
}

int fd2_trace_ranecu_faithful(fd2_ctx* c, const fd2_event* events, int64_t n,
                              int64_t event_base, int32_t seed_input, fd2_photon* records) {
  return trace_ranecu(c, events, n, event_base, seed_input, records, true);
}

int fd2_debug_steps(fd2_ctx* c, const fd2_event* events, int64_t n, int32_t max_photons,
                    fd2_step* steps, int32_t max_steps, int32_t* n_steps) {
  if (!c || !events || n < 1 || !steps || max_steps < 1 || !n_steps || max_photons < 0)
    return FD2_E_INVAL;
  int rc = validate_host(c, events, 1);
  if (rc) return rc;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  rc = grow(c, 1);
  if (rc) return rc;
  float* d_steps = nullptr;
  int32_t* d_count = nullptr;
  static_assert(sizeof(fd2_step) == 14 * sizeof(float), "fd2_step layout");
  if (hipMalloc(&d_steps, (size_t)max_steps * sizeof(fd2_step)) != hipSuccess) return FD2_E_NOMEM;
  if (hipMalloc(&d_count, sizeof(int32_t)) != hipSuccess) {
    (void)hipFree(d_steps);
    return FD2_E_NOMEM;
  }
  rc = FD2_OK;
  int32_t cnt = 0;
  if (hipMemcpyAsync(c->d_events, events, sizeof(fd2_event), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      launch_fill_u32((uint32_t*)d_count, 1, c->stream) != hipSuccess ||
      launch_debug_steps(c->d_events, (uint64_t)c->next_event, c->seed, c->d_geoms, max_photons,
                         d_steps, d_count, max_steps, c->stream) != hipSuccess ||
      hipMemcpyAsync(&cnt, d_count, sizeof cnt, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    rc = FD2_E_DEVICE;
  if (!rc) {
    if (cnt > max_steps) cnt = max_steps;
    if (cnt > 0 && hipMemcpy(steps, d_steps, (size_t)cnt * sizeof(fd2_step), hipMemcpyDeviceToHost) != hipSuccess)
      rc = FD2_E_DEVICE;
    *n_steps = cnt;
  }
  (void)hipFree(d_steps);
  (void)hipFree(d_count);
  return rc;
}

int fd2_debug_math(int fn, const float* x, float* y, int64_t n, int device) {
  if (fn < 0 || fn >= FD2_MATH_NFN || !x || !y || n < 0) return FD2_E_INVAL;
  if (n == 0) return FD2_OK;
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) return FD2_E_NO_DEVICE;
  if (device < 0 || device >= n_dev) return FD2_E_INVAL;
  FD2_HIP(hipSetDevice(device));
  float* d = nullptr;
  const size_t bytes = (size_t)n * sizeof(float);
  if (hipMalloc(&d, 2 * bytes) != hipSuccess) return FD2_E_NOMEM;
  int rc = FD2_OK;
  if (hipMemcpy(d, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      launch_debug_math(fn, d, d + n, n, nullptr) != hipSuccess ||
      hipMemcpy(y, d + n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    rc = FD2_E_DEVICE;
  (void)hipFree(d);
  return rc;
}

int fd2_prf_dims(const fd2_ctx* c, int geom, int32_t* xdim, int32_t* ydim) {
  if (!c || geom < 0 || geom >= c->n_geoms || !xdim || !ydim) return FD2_E_INVAL;
  *xdim = c->hgeom[geom].xdim;
  *ydim = c->hgeom[geom].ydim;
  return FD2_OK;
}

int fd2_get_prf(fd2_ctx* c, int geom, uint64_t* prf) {
  if (!c || geom < 0 || geom >= c->n_geoms || !prf) return FD2_E_INVAL;
  int rc = fd2_sync(c);
  if (rc) return rc;
  const size_t npix = (size_t)c->hgeom[geom].xdim * c->hgeom[geom].ydim;
  FD2_HIP(hipMemcpy(prf, c->d_prf[geom], npix * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return FD2_OK;
}

int fd2_get_hist_counts(fd2_ctx* c, int geom, uint32_t* hist) {
  if (!c || geom < 0 || geom >= c->n_geoms || !hist) return FD2_E_INVAL;
  int rc = fd2_sync(c);
  if (rc) return rc;
  const int32_t nh = c->params[geom].mynumhist;
  if (nh > 0) FD2_HIP(hipMemcpy(hist, c->d_hist[geom], (size_t)nh * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return FD2_OK;
}

int fd2_get_counters(fd2_ctx* c, int geom, uint64_t* counters) {
  if (!c || geom < 0 || geom >= c->n_geoms || !counters) return FD2_E_INVAL;
  int rc = fd2_sync(c);
  if (rc) return rc;
  FD2_HIP(hipMemcpy(counters, c->d_counters + (size_t)geom * kCounters,
                    kCounters * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return FD2_OK;
}

// PHS rule of main/hybridMANTIS_cuda_ver1_0.cu:376-399 (same in
// main/hybridMANTIS_c_ver1_0.c:343-366): binsize = (max-min)/nbins in integer
// arithmetic; histories without detections are skipped; <= min -> bin 0,
// >= max -> last bin, else count/binsize - min/binsize.  A bin index past
// the last bin (bin size not dividing max-min) is clamped to the last bin.
int fd2_phs_from_counts(const uint32_t* hist, int32_t num_hist, int32_t min_ph,
                        int32_t max_ph, int32_t num_bins, int32_t* phs) {
  if (!hist || !phs || num_hist < 0 || num_bins <= 0) return FD2_E_INVAL;
  const int binsize = (max_ph - min_ph) / num_bins;
  if (binsize <= 0) return FD2_E_INVAL;
  const int bincorr = min_ph / binsize;
  for (int32_t h = 0; h < num_hist; ++h) {
    const int64_t v = hist[h];
    if (v <= 0) continue;
    if (v <= min_ph) {
      phs[0]++;
    } else if (v >= max_ph) {
      phs[num_bins - 1]++;
    } else {
      int64_t b = v / binsize - bincorr;
      if (b > num_bins - 1) b = num_bins - 1;
      if (b < 0) b = 0;
      phs[b]++;
    }
  }
  return FD2_OK;
}

int fd2_get_phs(fd2_ctx* c, int geom, int32_t* phs) {
  if (!c || geom < 0 || geom >= c->n_geoms || !phs) return FD2_E_INVAL;
  const fd2_params& p = c->params[geom];
  if (p.mynumbins <= 0) return FD2_E_INVAL;
  const int32_t nh = p.mynumhist;
  uint32_t* h = (uint32_t*)std::malloc((size_t)(nh > 0 ? nh : 1) * sizeof(uint32_t));
  if (!h) return FD2_E_NOMEM;
  int rc = fd2_get_hist_counts(c, geom, h);
  if (!rc) {
    std::memset(phs, 0, (size_t)p.mynumbins * sizeof(int32_t));
    rc = fd2_phs_from_counts(h, nh, p.minphotons, p.maxphotons, p.mynumbins, phs);
  }
  std::free(h);
  return rc;
}

int fd2_device_alloc(fd2_ctx* c, size_t bytes, void** d_ptr) {
  if (!c || !d_ptr || bytes == 0) return FD2_E_INVAL;
  FD2_HIP(hipSetDevice(c->device));
  if (hipMalloc(d_ptr, bytes) != hipSuccess) return FD2_E_NOMEM;
  return FD2_OK;
}

int fd2_device_free(fd2_ctx* c, void* d_ptr) {
  if (!c) return FD2_E_INVAL;
  if (!d_ptr) return FD2_OK;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  FD2_HIP(hipFree(d_ptr));
  return FD2_OK;
}

int fd2_copy_to_device(fd2_ctx* c, void* d_dst, const void* h_src, size_t bytes) {
  if (!c || !d_dst || (!h_src && bytes)) return FD2_E_INVAL;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  FD2_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return FD2_OK;
}

int fd2_last_kernel_ms(fd2_ctx* c, float* ms) {
  if (!c || !ms) return FD2_E_INVAL;
  if (!c->timed) return FD2_E_STATE;
  FD2_HIP(hipEventSynchronize(c->ev_stop));
  FD2_HIP(hipEventElapsedTime(ms, c->ev_start, c->ev_stop));
  return FD2_OK;
}

// ---- HIP graph of one batch --------------------------------------------------

}  // extern "C"

struct fd2_graph {
  fd2_ctx* c = nullptr;
  int64_t n = 0;
  uint32_t* d_chunk_off = nullptr;
  unsigned long long* d_photon_off = nullptr;
  uint32_t* d_block_chunks = nullptr;
  unsigned long long* d_block_photons = nullptr;
  unsigned long long* d_base = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

void fd2_graph_destroy(fd2_graph* g) {
  if (!g) return;
  if (g->c) {
    (void)hipSetDevice(g->c->device);
    (void)hipStreamSynchronize(g->c->stream);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  void* bufs[] = {g->d_chunk_off, g->d_photon_off, g->d_block_chunks, g->d_block_photons,
                  g->d_base};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete g;
}

int fd2_graph_create(fd2_ctx* c, const fd2_event* d_events, int64_t n, fd2_graph** out) {
  if (!c || !out || n <= 0 || !d_events || n >= (int64_t)UINT32_MAX) return FD2_E_INVAL;
  *out = nullptr;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipStreamSynchronize(c->stream));
  fd2_graph* g = new (std::nothrow) fd2_graph();
  if (!g) return FD2_E_NOMEM;
  g->c = c;
  g->n = n;
  const size_t nb = (size_t)((n + kScanTile - 1) / kScanTile) + 1;
  int rc = FD2_OK;
  if (hipMalloc(&g->d_chunk_off, (size_t)(n + 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&g->d_photon_off, (size_t)(n + 1) * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&g->d_block_chunks, nb * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&g->d_block_photons, nb * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&g->d_base, sizeof(unsigned long long)) != hipSuccess)
    rc = FD2_E_NOMEM;
  if (!rc) {
    const QueueBufs b{g->d_chunk_off, g->d_photon_off, g->d_block_chunks, g->d_block_photons};
    Launch L = make_launch(c, b, d_events, n, nullptr, nullptr);
    L.a.event_base_dev = g->d_base;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = FD2_E_DEVICE;
    } else {
      const hipError_t e = issue(L, c->stream);
      const hipError_t e2 = hipStreamEndCapture(c->stream, &g->graph);
      if (e != hipSuccess || e2 != hipSuccess ||
          hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = FD2_E_DEVICE;
    }
  }
  if (rc) {
    (void)hipGetLastError();
    fd2_graph_destroy(g);
    return rc;
  }
  *out = g;
  return FD2_OK;
}

int fd2_graph_launch(fd2_graph* g) {
  if (!g || !g->exec) return FD2_E_INVAL;
  fd2_ctx* c = g->c;
  if (g->n > INT64_MAX - c->next_event) return FD2_E_STATE;
  FD2_HIP(hipSetDevice(c->device));
  FD2_HIP(hipEventRecord(c->ev_start, c->stream));
  FD2_HIP(launch_set_u64(g->d_base, (unsigned long long)c->next_event, c->stream));
  FD2_HIP(hipGraphLaunch(g->exec, c->stream));
  FD2_HIP(hipEventRecord(c->ev_stop, c->stream));
  c->timed = true;
  c->next_event += g->n;
  return FD2_OK;
}

// ---- autotuner ------------------------------------------------------------------

int fd2_autotune(fd2_ctx* c, const fd2_event* d_events, int64_t n, int32_t repeats,
                 fd2_tuning* best, fd2_tune_result* results, int32_t max_results,
                 int32_t* n_results) {
  if (!c || !d_events || n <= 0 || (max_results > 0 && !results)) return FD2_E_INVAL;
  if (repeats < 1) repeats = 1;
  fd2_ctx* s = nullptr;
  int rc = fd2_create(c->params, c->n_geoms, c->device, &s);
  if (rc) return rc;
  fd2_set_seed(s, c->seed);
  int32_t count = 0;
  auto time_one = [&](const fd2_tuning& t, float* ms_out) -> int {
    int r = fd2_set_tuning(s, &t);
    if (r) return r;
    float best_ms = 0.0f;
    for (int k = 0; k <= repeats; ++k) {  // the first run warms up
      r = fd2_submit_device(s, d_events, n);
      if (!r) r = fd2_sync(s);
      float ms = 0.0f;
      if (!r) r = fd2_last_kernel_ms(s, &ms);
      if (r) return r;
      if (k == 1 || (k > 1 && ms < best_ms)) best_ms = ms;
    }
    if (count < max_results) {
      results[count].tuning = t;
      results[count].ms = best_ms;
    }
    ++count;
    *ms_out = best_ms;
    return fd2_reset(s);
  };
  // coordinate descent from the context's current shape, two sweeps over
  // chunk, waves per CU, parked lanes and LDS tile side
  static const int chunks[] = {16, 32, 64, 128, 256};
  static const int waves[] = {12, 16, 20};
  static const int parks[] = {8, 16, 32, 48};
  static const int tiles[] = {-1, 32, 48, 64, 80};
  static const int refills[] = {1, 4, 8, 16, 32};
  static const int queues[] = {-1, 1};
  fd2_tuning cur = c->tune;
  float cur_ms = 0.0f;
  rc = time_one(cur, &cur_ms);
  for (int sweep = 0; sweep < 2 && !rc; ++sweep) {
    for (int dim = 0; dim < 6 && !rc; ++dim) {
      const int* vals = dim == 0 ? chunks : dim == 1 ? waves : dim == 2 ? parks
                      : dim == 3 ? tiles : dim == 4 ? refills : queues;
      const int nv = dim == 1 ? 3 : dim == 2 ? 4 : dim == 5 ? 2 : 5;
      for (int i = 0; i < nv && !rc; ++i) {
        fd2_tuning t = cur;
        int32_t* f = dim == 0 ? &t.chunk : dim == 1 ? &t.waves_per_cu
                   : dim == 2 ? &t.park_lanes : dim == 3 ? &t.lds_tile
                   : dim == 4 ? &t.refill_lanes : &t.wall_queue;
        if (*f == vals[i] || (dim == 3 && *f == 0 && vals[i] == kTileDefault) ||
            (dim == 5 && *f == 0 && vals[i] == 1))
          continue;
        *f = vals[i];
        float ms = 0.0f;
        rc = time_one(t, &ms);
        if (!rc && ms < cur_ms) {
          cur = t;
          cur_ms = ms;
        }
      }
    }
  }
  fd2_destroy(s);
  if (rc) return rc;
  if (n_results) *n_results = count < max_results ? count : max_results;
  if (best) *best = cur;
  return fd2_set_tuning(c, &cur);
}

}  // extern "C"
