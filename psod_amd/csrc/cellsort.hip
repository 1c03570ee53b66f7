// cellsort.hip -- a5: cell bucketing as an integer counting sort.
//
// Replaces main file:1561-1604 (cell key, quick_sort of the keys as reals,
// run-length first_ngp/num_ngp).  Histogram with atomics, exclusive scan
// into cell_start, scatter with per-cell fill counters.  The output records
// hold the particle position relative to the origin of its 8^3-cell block:
// u = x*nx - 8*floor(floor(x*nx)/8), exact in float32 (a multiple-of-8
// subtraction from x*nx, itself exact for a power-of-two nx).
#include "psod_internal.h"

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, int64_t n, uint64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v) {
  if (n <= 0) return PSOD_OK;
  k_fill_i32<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(p, n, v);
  PSOD_LAUNCH_CHECK();
  return PSOD_OK;
}

int psod_fill_u64(psod_ctx *c, uint64_t *p, int64_t n, uint64_t v) {
  if (n <= 0) return PSOD_OK;
  k_fill_u64<<<grid_for(n, 256, 8192), 256, 0, c->stream>>>(p, n, v);
  PSOD_LAUNCH_CHECK();
  return PSOD_OK;
}

// scaled coordinate in [0, nx) and its cell
__device__ inline float scaled_wrap(float x, int32_t nx, int32_t *cell) {
  float xs = x * (float)nx;
  if (xs >= (float)nx) xs -= (float)nx;
  if (xs < 0.0f) xs += (float)nx;
  int32_t c = (int32_t)floorf(xs);
  if (c >= nx) { c -= nx; xs -= (float)nx; }
  *cell = c;
  return xs;
}

// Block of a particle (8^3 cells), its scaled coordinates and cell
struct PartCell {
  float xs, ys, zs;
  int32_t cx, cy, cz, blk;
};

__device__ inline PartCell part_cell(const float *pos, int64_t i, int32_t nx) {
  PartCell p;
  p.xs = scaled_wrap(pos[3 * i + 0], nx, &p.cx);
  p.ys = scaled_wrap(pos[3 * i + 1], nx, &p.cy);
  p.zs = scaled_wrap(pos[3 * i + 2], nx, &p.cz);
  const int32_t nb = nx >> 3;
  p.blk = ((p.cz >> 3) * nb + (p.cy >> 3)) * nb + (p.cx >> 3);
  return p;
}

// Wave-aggregated "atomicAdd(&cnt[key], 1)" returning each lane's slot:
// lanes with the leader's key share one atomic (particles of one clump are
// usually consecutive in the input).  After a few rounds the remaining
// lanes fall back to one atomic each.  Every lane of the wave must call it.
__device__ inline int32_t agg_slot(int32_t *cnt, int32_t key, bool valid,
                                   bool want_slot) {
  const int l = __lane_id();
  uint64_t act = __ballot(valid);
  int32_t slot = 0;
  for (int it = 0; it < 4 && act; it++) {
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int32_t kl = __shfl(key, leader, 64);
    const bool mine = valid && ((act >> l) & 1) && key == kl;
    const uint64_t m = __ballot(mine);
    int32_t base = 0;
    if (l == leader) {
      if (want_slot) base = atomicAdd(&cnt[kl], __popcll(m));
      else atomicAdd(&cnt[kl], __popcll(m));
    }
    base = __shfl(base, leader, 64);
    if (mine) {
      const uint64_t lower = (l == 0) ? 0ull : (~0ull >> (64 - l));
      slot = base + __popcll(m & lower);
    }
    act &= ~m;
  }
  if ((act >> l) & 1) {
    if (want_slot) slot = atomicAdd(&cnt[key], 1);
    else atomicAdd(&cnt[key], 1);
  }
  return slot;
}

__global__ void __launch_bounds__(256)
k_block_hist(const float *__restrict__ pos, int64_t n, int32_t nx,
             int32_t *__restrict__ bcount) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const bool ok = i < n;
    int32_t blk = 0;
    if (ok) blk = part_cell(pos, i, nx).blk;
    agg_slot(bcount, blk, ok, false);
  }
}

// records staged in their block's region, unsorted inside it
__global__ void __launch_bounds__(256)
k_block_scatter(const float *__restrict__ pos, const int32_t *__restrict__ lab,
                int64_t n, int32_t nx, const int32_t *__restrict__ bstart,
                int32_t *__restrict__ bfill, float4 *__restrict__ stage,
                int32_t *__restrict__ stage_lab) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const bool ok = i < n;
    PartCell p;
    p.blk = 0;
    if (ok) p = part_cell(pos, i, nx);
    const int32_t s = agg_slot(bfill, p.blk, ok, true);
    if (ok) {
      const int32_t slot = bstart[p.blk] + s;
      float4 r;
      r.x = p.xs - (float)(p.cx & ~7);
      r.y = p.ys - (float)(p.cy & ~7);
      r.z = p.zs - (float)(p.cz & ~7);
      r.w = __int_as_float((int32_t)i);
      stage[slot] = r;
      if (stage_lab) stage_lab[slot] = lab ? lab[i] : -1;
    }
  }
}

// one workgroup per block: counting sort of its records by local cell
// (LDS), final records and the block's 512 cell offsets
__global__ void __launch_bounds__(256)
k_block_sort(const int32_t *__restrict__ bstart, int64_t nblk,
             const float4 *__restrict__ stage,
             const int32_t *__restrict__ stage_lab,
             int32_t *__restrict__ cell_start, float4 *__restrict__ rec,
             int32_t *__restrict__ perm, int32_t *__restrict__ lab_sorted) {
  __shared__ int32_t cnt[512];
  for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int32_t a = bstart[b], n = bstart[b + 1] - a;
    for (int i = threadIdx.x; i < 512; i += 256) cnt[i] = 0;
    __syncthreads();
    for (int32_t t = threadIdx.x; t < n; t += 256) {
      const float4 r = stage[a + t];
      const int32_t l = ((int32_t)r.z * 8 + (int32_t)r.y) * 8 + (int32_t)r.x;
      atomicAdd(&cnt[l], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of 512 counters by one wave
      int32_t v[8], s = 0;
      for (int k = 0; k < 8; k++) { v[k] = cnt[threadIdx.x * 8 + k]; s += v[k]; }
      const int32_t inc = wave_incl_scan(s);
      int32_t e = inc - s;
      for (int k = 0; k < 8; k++) {
        cnt[threadIdx.x * 8 + k] = e;
        cell_start[b * 512 + threadIdx.x * 8 + k] = a + e;
        e += v[k];
      }
    }
    __syncthreads();
    for (int32_t t = threadIdx.x; t < n; t += 256) {
      const float4 r = stage[a + t];
      const int32_t l = ((int32_t)r.z * 8 + (int32_t)r.y) * 8 + (int32_t)r.x;
      const int32_t d = a + atomicAdd(&cnt[l], 1);
      rec[d] = r;
      if (perm) perm[d] = __float_as_int(r.w);
      if (lab_sorted) lab_sorted[d] = stage_lab[a + t];
    }
    __syncthreads();
  }
}

__global__ void k_set_last(int32_t *__restrict__ p, int64_t i, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[i] = v;
}

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);

// ---------------------------------------------------------------------------
// Dense cells (>= 64 particles) are reordered by the Morton key of the
// particle's 1/8-cell sub-cell, so that any 64 consecutive particles of a
// dense cell occupy a compact box.  The order inside a cell carries no
// meaning elsewhere; the centre search uses the compact chunks to skip
// neighbour chunks beyond its 1-cell reach.
// ---------------------------------------------------------------------------
constexpr int DENSE = 64;
constexpr int SUB_T = 256, SUB_V = 16, SUB_REG = SUB_T * SUB_V;
constexpr int64_t SUB_TMP = 1 << 18;  // records per workgroup slice

__device__ inline uint32_t spread3(uint32_t v) {
  return (v & 1) | ((v & 2) << 2) | ((v & 4) << 4);
}
__device__ inline uint32_t morton9(const float4 &r) {
  const uint32_t qx = (uint32_t)((r.x - floorf(r.x)) * 8.0f) & 7;
  const uint32_t qy = (uint32_t)((r.y - floorf(r.y)) * 8.0f) & 7;
  const uint32_t qz = (uint32_t)((r.z - floorf(r.z)) * 8.0f) & 7;
  return spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
}

__global__ void k_dense_cells(const int32_t *__restrict__ cs, int64_t ncell,
                              int32_t *__restrict__ list,
                              int32_t *__restrict__ cnt, int32_t cap) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += (int64_t)gridDim.x * blockDim.x)
    if (cs[c + 1] - cs[c] >= DENSE) {
      const int32_t k = atomicAdd(cnt, 1);
      if (k < cap) list[k] = (int32_t)c;
    }
}

__global__ void __launch_bounds__(SUB_T)
k_subsort(const int32_t *__restrict__ cs, const int32_t *__restrict__ list,
          const int32_t *__restrict__ nlist, int32_t cap,
          int32_t *__restrict__ queue,
          float4 *__restrict__ rec, int32_t *__restrict__ perm,
          int32_t *__restrict__ lab, float4 *__restrict__ trec,
          int32_t *__restrict__ tperm, int32_t *__restrict__ tlab) {
  __shared__ int32_t hist[512];
  __shared__ int32_t item;
  const int32_t nl = min(*nlist, cap);
  float4 *tr = trec + (int64_t)blockIdx.x * SUB_TMP;
  int32_t *tp = tperm + (int64_t)blockIdx.x * SUB_TMP;
  int32_t *tl = tlab + (int64_t)blockIdx.x * SUB_TMP;
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(queue, 1);
    for (int i = threadIdx.x; i < 512; i += SUB_T) hist[i] = 0;
    __syncthreads();
    const int32_t it = item;
    if (it >= nl) break;
    const int32_t c = list[it];
    const int32_t a = cs[c], n = cs[c + 1] - a;
    if (n > SUB_TMP) {  // would not fit the slice: keep the scatter order
      __syncthreads();
      continue;
    }
    for (int32_t t = threadIdx.x; t < n; t += SUB_T)
      atomicAdd(&hist[morton9(rec[a + t])], 1);
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of 512 counters by one wave
      int32_t v[8], s = 0;
      for (int k = 0; k < 8; k++) { v[k] = hist[threadIdx.x * 8 + k]; s += v[k]; }
      const int32_t inc = wave_incl_scan(s);
      int32_t e = inc - s;
      for (int k = 0; k < 8; k++) { hist[threadIdx.x * 8 + k] = e; e += v[k]; }
    }
    __syncthreads();
    if (n <= SUB_REG) {
      // in place: everything is read into registers before any write
      float4 rr[SUB_V];
      int32_t pp[SUB_V], ll[SUB_V], dd[SUB_V];
#pragma unroll
      for (int k = 0; k < SUB_V; k++) {
        const int32_t t = threadIdx.x + k * SUB_T;
        if (t < n) {
          rr[k] = rec[a + t];
          pp[k] = perm ? perm[a + t] : 0;
          ll[k] = lab ? lab[a + t] : 0;
          dd[k] = atomicAdd(&hist[morton9(rr[k])], 1);
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < SUB_V; k++) {
        const int32_t t = threadIdx.x + k * SUB_T;
        if (t < n) {
          rec[a + dd[k]] = rr[k];
          if (perm) perm[a + dd[k]] = pp[k];
          if (lab) lab[a + dd[k]] = ll[k];
        }
      }
    } else {
      for (int32_t t = threadIdx.x; t < n; t += SUB_T) {
        const float4 r = rec[a + t];
        const int32_t d = atomicAdd(&hist[morton9(r)], 1);
        tr[d] = r;
        if (perm) tp[d] = perm[a + t];
        if (lab) tl[d] = lab[a + t];
      }
      __threadfence_block();
      __syncthreads();
      for (int32_t t = threadIdx.x; t < n; t += SUB_T) {
        rec[a + t] = tr[t];
        if (perm) perm[a + t] = tp[t];
        if (lab) lab[a + t] = tl[t];
      }
    }
    __syncthreads();
  }
}

extern "C" int psod_cell_sort(psod_ctx *ctx, int64_t npart, const float *d_pos,
                              const int32_t *d_labels, int32_t nx,
                              int32_t *d_cell_start, int32_t *d_perm,
                              float *d_rec, int32_t *d_lab_sorted) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart < 0 || npart >= (1ll << 31) || !d_pos || !d_cell_start || !d_rec)
    return psod_set_error(PSOD_E_ARG, "psod_cell_sort: bad arguments");
  if (nx < 8 || nx > 1024 || (nx & (nx - 1)))
    return psod_set_error(PSOD_E_ARG,
                          "psod_cell_sort: nx=%d must be a power of two in "
                          "[8, 1024]", nx);
  const int64_t ncell = (int64_t)nx * nx * nx;
  const int64_t nblk = ncell / BS3;
  hipStream_t st = ctx->stream;
  // block counters (+1 for the total), their scan and the fill counters
  int32_t *bc = (int32_t *)psod_scratch(ctx, S_COUNT,
                                        (size_t)(3 * (nblk + 1) + 64) * 4);
  float4 *stage = (float4 *)psod_scratch(ctx, S_KEYS2,
                                         (size_t)(npart > 0 ? npart : 1) * 16);
  int32_t *slab = d_lab_sorted
                      ? (int32_t *)psod_scratch(ctx, S_VALS2,
                                                (size_t)(npart > 0 ? npart : 1) * 4)
                      : nullptr;
  if (!bc || !stage || (d_lab_sorted && !slab)) return PSOD_E_NOMEM;
  int32_t *bstart = bc + nblk + 16, *bfill = bstart + nblk + 16;
  PSOD_TRY(psod_fill_i32(ctx, bc, nblk + 1, 0));
  PSOD_TRY(psod_fill_i32(ctx, bfill, nblk, 0));
  const unsigned g = grid_for(npart, 256, (int64_t)ctx->num_cus * 64);
  if (npart > 0) {
    k_block_hist<<<g, 256, 0, st>>>(d_pos, npart, nx, bc);
    PSOD_LAUNCH_CHECK();
  }
  PSOD_TRY(psod_scan_i32(ctx, bc, bstart, nblk + 1, nullptr));
  if (npart > 0) {
    k_block_scatter<<<g, 256, 0, st>>>(d_pos, d_labels, npart, nx, bstart,
                                       bfill, stage, slab);
    PSOD_LAUNCH_CHECK();
  }
  k_block_sort<<<grid_for(nblk, 1, (int64_t)ctx->num_cus * 32), 256, 0, st>>>(
      bstart, nblk, stage, slab, d_cell_start, (float4 *)d_rec, d_perm,
      d_lab_sorted);
  PSOD_LAUNCH_CHECK();
  k_set_last<<<1, 64, 0, st>>>(d_cell_start, ncell, (int32_t)npart);
  PSOD_LAUNCH_CHECK();
  if (npart > 0) {
    // Morton order of the 1/8-cell sub-cells inside dense cells
    int32_t *list = (int32_t *)psod_scratch(ctx, S_HIST,
                                            (size_t)(ncell / DENSE + 64) * 4);
    int32_t *ctl = (int32_t *)psod_scratch(ctx, S_MISC, 64);
    if (!list || !ctl) return PSOD_E_NOMEM;
    PSOD_TRY(psod_fill_i32(ctx, ctl, 2, 0));
    k_dense_cells<<<grid_for(ncell, 256, (int64_t)ctx->num_cus * 64), 256, 0,
                    st>>>(d_cell_start, ncell, list, ctl,
                          (int32_t)(ncell / DENSE + 32));
    PSOD_LAUNCH_CHECK();
    const unsigned nsub = (unsigned)ctx->num_cus;
    char *tmp = (char *)psod_scratch(ctx, S_KEYS,
                                     (size_t)nsub * SUB_TMP * 24);
    if (!tmp) return PSOD_E_NOMEM;
    k_subsort<<<nsub, SUB_T, 0, st>>>(
        d_cell_start, list, ctl, (int32_t)(ncell / DENSE + 32), ctl + 1,
        (float4 *)d_rec, d_perm,
        d_lab_sorted, (float4 *)tmp,
        (int32_t *)(tmp + (size_t)nsub * SUB_TMP * 16),
        (int32_t *)(tmp + (size_t)nsub * SUB_TMP * 20));
    PSOD_LAUNCH_CHECK();
  }
  PSOD_HIP(hipStreamSynchronize(st));
  return PSOD_OK;
}
