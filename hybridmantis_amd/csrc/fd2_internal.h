// fd2_internal.h -- declarations shared by the host library and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fd2.h"
#include "fd2_device.h"

namespace fd2 {

enum { kCounters = FD2_NCOUNT };

struct TransportArgs {
  const fd2_event* events;             // device
  const uint32_t* chunk_off;           // n_events + 1 (exclusive scan)
  const unsigned long long* photon_off;  // n_events + 1 (exclusive scan)
  const GeomDev* geoms;                // n_geoms
  unsigned long long* counters;        // n_geoms * FD2_NCOUNT
  uint32_t* cursor;                    // work-queue cursor (zeroed by the scan)
  fd2_photon* records;                 // per-photon records (recording only)
  uint32_t* radial;                    // n_events x FD2_RADIAL_BINS (radial mode)
  const unsigned long long* event_base_dev;  // if set, read the number of events[0] here
  uint32_t* wave_queue;                // per-wave photon queues (null: park in place)
  uint32_t* error_flag;                // bit 1: a wave's loop limit was reached
  int64_t n_events;
  uint32_t seed;
  uint64_t event_base;                 // global number of events[0]
  int32_t n_geoms;
  int32_t chunk_log2;
  int32_t park_lanes;                  // lanes waiting before a Fresnel phase
  int32_t refill_lanes;                // free lanes before new photons start
  int32_t tile_side;                   // LDS PRF tile side (pixels), 0 = none
  int32_t n_tiles;                     // geometries [0, n_tiles) have a tile
};

// LDS PRF tiles of the persistent kernel: geometry g < n_tiles counts the
// detections in a tile_side^2 pixel square centred on its PRF image in LDS
// (u32) and adds the block's counts to the PRF at the end.
constexpr int kTileDefault = 0;  // no tile: the LDS holds the wall queues
constexpr int kTileBudget = 64 * 64;   // u32 entries per block (16 KiB)
// started photons of a wave's chunk in LDS (direction + stream words)
constexpr int kStartWords = 7;
// per-wave photon queues of the persistent kernel: the wall queue in LDS
// (kWallCap slots of kWallWords words, word-major), the gap queue and the
// backlog in HBM (slots in that order)
constexpr int kWallCap = 64;
constexpr int kWallWords = 16;  // the photon words of a queue entry
constexpr int kGapCap = 128;
constexpr int kBackCap = 64 + kWallCap + kGapCap;
constexpr int kGapBase = 0;
constexpr int kBackBase = kGapCap;
constexpr int kQSlots = kGapCap + kBackCap;
constexpr int kQWords = 20;  // 16 photon words + 4 of wall-branch state
constexpr int kQBlock = 32;  // queue slots per block (word stride 128 bytes)
static_assert(kQSlots % kQBlock == 0 && kGapBase % kQBlock == 0 && kBackBase % kQBlock == 0,
              "queue regions are whole blocks");
constexpr uint32_t kMaxIdlePasses = 256;  // passes without a step or a wall branch
// largest LDS tile side: 49 KiB of tile, plus the kernel's static LDS (about
// 16 KiB with the several-geometry table); fd2_set_tuning checks the sum
// against the device's LDS per workgroup (160 KiB on gfx950)
constexpr int kTileMax = 112;

struct QueueArgs {
  const fd2_event* events;
  uint32_t* chunk_off;
  unsigned long long* photon_off;
  uint32_t* block_chunks;      // per scan block
  unsigned long long* block_photons;
  uint32_t* cursor;
  uint32_t* error_flag;        // set when an event is invalid
  int64_t n_events;
  int32_t n_geoms;
  int32_t chunk_log2;
  int32_t n_blocks;
  uint64_t max_chunks;         // more work items than this: error bit 4, nothing runs
};

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kTransportThreads = 256;

// launch wrappers (fd2_kernels.hip)
hipError_t launch_queue_build(const QueueArgs& q, hipStream_t s);
hipError_t launch_transport(const TransportArgs& a, bool multi, bool record,
                            bool radial, int n_blocks, hipStream_t s);
hipError_t launch_trace_ranecu(const fd2_event* events, int64_t n,
                               int64_t event_base, int32_t seed_input,
                               const GeomDev* geoms, int32_t n_geoms,
                               unsigned long long* counters,
                               const unsigned long long* photon_off,
                               fd2_photon* records, bool faithful, hipStream_t s);
hipError_t launch_debug_steps(const fd2_event* events, uint64_t event_base,
                              uint32_t seed, const GeomDev* geoms,
                              int32_t n_photons, float* steps,
                              int32_t* n_steps, int32_t max_steps,
                              hipStream_t s);
hipError_t launch_debug_math(int fn, const float* x, float* y, int64_t n, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, size_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, hipStream_t s);
hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
hipError_t launch_set_u64(unsigned long long* p, unsigned long long v, hipStream_t s);
int transport_max_blocks_per_cu(bool multi, bool record, bool radial, size_t smem, bool queue);
size_t transport_static_lds();  // largest static LDS of the transport kernels

}  // namespace fd2

// The transport context (host side).
struct fd2_ctx {
  int device = 0;
  int n_cu = 0;
  size_t lds_per_block = 0;   // the device's LDS per workgroup
  size_t lds_static = 0;      // the transport kernels' static LDS (largest)
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool timed = false;

  int n_geoms = 0;
  fd2_params params[FD2_MAX_GEOMS];
  fd2::GeomDev hgeom[FD2_MAX_GEOMS];
  fd2::GeomDev* d_geoms = nullptr;
  unsigned long long* d_prf[FD2_MAX_GEOMS] = {};
  uint32_t* d_hist[FD2_MAX_GEOMS] = {};
  unsigned long long* d_counters = nullptr;

  // staging and queue buffers, grown on demand
  fd2_event* h_stage = nullptr;   // pinned
  size_t cap_stage = 0;
  fd2_event* d_events = nullptr;
  size_t cap_events = 0;
  uint32_t* d_chunk_off = nullptr;
  unsigned long long* d_photon_off = nullptr;
  size_t cap_off = 0;
  uint32_t* d_block_chunks = nullptr;
  unsigned long long* d_block_photons = nullptr;
  size_t cap_blocks = 0;
  uint32_t* d_cursor = nullptr;
  uint32_t* d_error = nullptr;
  uint32_t* d_wave_queue = nullptr;    // kQWords x kQSlots words per resident wave
  int max_waves = 0;

  uint32_t seed = 271828182u;
  int64_t next_event = 0;
  fd2_tuning tune{};
};

namespace fd2 {

// Stage n events through the context's pinned buffer and enqueue their
// transport, with per-event radial tallies into d_radial (n x
// FD2_RADIAL_BINS u32, cleared here), on the context's stream without
// waiting for it: the events take the next n event numbers, as in
// fd2_submit_radial, and fd2_sync waits for them and checks the error flag.
int submit_radial_async(fd2_ctx* c, const fd2_event* events, int64_t n, uint32_t* d_radial);

}  // namespace fd2
