// psod_internal.h -- shared declarations of the MI355X SOD library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/psod.h"

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
int psod_set_error(int code, const char *fmt, ...);

#define PSOD_HIP(x)                                                          \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      return psod_set_error(PSOD_E_DEVICE, "%s: %s (%s:%d)", #x,             \
                            hipGetErrorString(e_), __FILE__, __LINE__);      \
  } while (0)

#define PSOD_TRY(x)                                                          \
  do {                                                                       \
    int r_ = (x);                                                            \
    if (r_ != PSOD_OK) return r_;                                            \
  } while (0)

#define PSOD_LAUNCH_CHECK() PSOD_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// context: one device, one stream, growable scratch slots
// ---------------------------------------------------------------------------
enum ScratchSlot {
  S_COUNT = 0,   // cell-sort histogram / fill counters
  S_SCAN,        // scan partials
  S_SCAN2,
  S_KEYS,        // radix sort keys / values (ping-pong)
  S_KEYS2,
  S_VALS,
  S_VALS2,
  S_HIST,        // radix / SOD histograms
  S_PLAN,        // SOD per-halo plan
  S_TASK,        // SOD task lists
  S_TASKRES,     // SOD task results
  S_COLLECT,     // SOD collected band particles
  S_BINS,        // SOD radial histograms
  S_MISC,
  S_HOST_A,      // buffers of the host-pointer entry points
  S_HOST_B,
  S_HOST_C,
  S_HOST_D,
  S_HOST_E,
  S_HOST_F,
  S_HOST_G,
  S_PLAIN,       // plain FP64 path
  S_PLAIN2,
  S_GREEDY,
  S_GREEDY2,
  S_GREEDY3,
  S_TILE,      // deposit tile counts and dispatch order
  S_CS_A,      // cell sort: records, block ids, labels (two buffers)
  S_CS_C,
  S_CS_H,      // cell sort: bucket tables and tile histograms
  S_PASS_CS,   // psod_pass: cell offsets, records, labels, halo arrays, grid
  S_PASS_REC,
  S_PASS_LAB,
  S_PASS_HALO,
  S_PASS_GRID,
  S_TUNE,      // autotuner: per-halo outputs
  S_GIANT,     // SOD: list and states of the giant halos (staged growth)
  S_GIANT_H,   // SOD: their histogram slices
  S_GIANT_P,   // SOD: their collection slices
  S_GIANT_SEG, // SOD: per-segment counts of their cubes, prefix sums
  S_GIANT_V,   // SOD: values of the device-wide sorts of their bands
  S_NSLOT
};

// tunables loaded from a tuning file, one entry per grid size
constexpr int PSOD_TUNE_ENTRIES = 16;
struct TuneEntry {
  int32_t nx;
  psod_tunables t;
};

// built-in defaults of the tunables
constexpr int32_t TUNE_SMALL_MAX = 1024, TUNE_BOUND_MIN = 1024,
                  TUNE_DENSE_TILE = 1 << 14, TUNE_SMALL_WG = 8,
                  TUNE_GIANT_MIN = 1 << 20;

struct psod_ctx {
  int device;
  psod_tunables tune_set;   // psod_set_tunables; 0 fields are unset
  TuneEntry tune_file[PSOD_TUNE_ENTRIES];
  int ntune_file;
  char device_name[256];
  hipStream_t stream;
  void *slot[S_NSLOT];
  size_t cap[S_NSLOT];
  int num_cus;
  int64_t sod_stat[13];  // see psod_sod_stats
  hipEvent_t ev[2];     // brackets the centre search (sod_stat[8])
  // second stream (default priority) for work that overlaps the main one,
  // and the events that fork to it and join from it
  hipStream_t side;
  hipEvent_t fork, join;
  // set while psod_pass_graph captures the pass: entry points queue their
  // work without waiting, device scratch may not grow, and the SOD reads
  // its work counts on the device (GraphCaps bounds them)
  int capturing;
  struct GraphCaps {
    int64_t tasks;  // centre-search tasks of one pass, at most
  } gcap;
  struct GraphState *graph;  // psod_pass_graph's captured pass (or NULL)
};

// Frees psod_pass_graph's captured pass of the context (psod_destroy).
void psod_graph_release(psod_ctx *ctx);

// End of an entry point: waits for the context's stream, except while the
// pass is being captured.
int psod_finish(psod_ctx *ctx);

// Resolves the device address of the fixed array that
// psod_cic_deposit_noalloc uses on `device` (which must be current).
int psod_static_scratch_init(int device);

// intermediates of the FoF-seeded pass (context scratch)
struct PassBuffers {
  int32_t *cell_start;  // [nx^3 + 1]
  float *rec;           // [4 npart]
  int32_t *lab_sorted;  // [npart]
  uint64_t *seed_key;   // [nhalo]
  int32_t *seed, *nmem; // [nhalo]
  float *grid;          // [ng^3]
};

// Tunables in force for grids of nx cells per side: the built-in defaults,
// overridden by a loaded tuning-file entry for nx, overridden by the fields
// set with psod_set_tunables; every field is > 0 and within its range.
psod_tunables psod_tunables_for(const psod_ctx *ctx, int32_t nx);

// Returns a device buffer of at least `bytes` in `slot` (contents not kept
// across growth) or nullptr with the error set.
void *psod_scratch(psod_ctx *ctx, int slot, size_t bytes);
int psod_use_device(psod_ctx *ctx);

// ---------------------------------------------------------------------------
// cells: 8^3 blocks, key = block*512 + (lz*8 + ly)*8 + lx
// ---------------------------------------------------------------------------
constexpr int BS = PSOD_BLOCK;
constexpr int BS3 = BS * BS * BS;

__host__ __device__ inline int32_t cell_key(int32_t nx, int32_t ix, int32_t iy,
                                            int32_t iz) {
  const int32_t nb = nx >> 3;
  const int32_t blk = ((iz >> 3) * nb + (iy >> 3)) * nb + (ix >> 3);
  return (blk << 9) | ((iz & 7) << 6) | ((iy & 7) << 3) | (ix & 7);
}

__host__ __device__ inline int32_t wrapi(int32_t i, int32_t n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Cell of a coordinate: x*nx is exact in float32 for a power-of-two nx.
__host__ __device__ inline int32_t cell_of(float x, int32_t nx) {
  const float xs = x * (float)nx;
  int32_t i = (int32_t)floorf(xs);
  if (i < 0) i += nx;
  if (i >= nx) i -= nx;
  return i;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
inline unsigned grid_for(int64_t n, int block, int64_t maxblocks = 1 << 30) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > maxblocks) g = maxblocks;
  return (unsigned)g;
}

// device-wide exclusive scan of int32 counts (n elements) into out (n
// elements); returns PSOD_OK.  `total` (optional, device) receives the sum.
int psod_scan_i32(psod_ctx *ctx, const int32_t *in, int32_t *out, int64_t n,
                  int32_t *total);

// radix sort of 64-bit keys with 32-bit values, ascending, stable.  With
// d_n not NULL only the first min(n, *d_n) elements are sorted (a count
// known on the device only; n is then the capacity).
int psod_radix_sort_u64(psod_ctx *ctx, int64_t n, uint64_t *keys,
                        uint32_t *vals, int key_bits,
                        const int32_t *d_n = nullptr);

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ inline int lane_id() { return __lane_id(); }

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline int32_t wave_incl_scan(int32_t v) {
  const int l = __lane_id();
  for (int o = 1; o < 64; o <<= 1) {
    int32_t t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }
