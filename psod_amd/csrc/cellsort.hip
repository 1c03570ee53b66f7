// cellsort.hip -- a5: cell bucketing as an integer counting sort.
//
// Replaces main file:1561-1604 (cell key, quick_sort of the keys as reals,
// run-length first_ngp/num_ngp).  Histogram with atomics, exclusive scan
// into cell_start, scatter with per-cell fill counters.  The output records
// hold the particle position relative to the origin of its 8^3-cell block:
// u = x*nx - 8*floor(floor(x*nx)/8), exact in float32 (a multiple-of-8
// subtraction from x*nx, itself exact for a power-of-two nx).
#include "psod_internal.h"

#include <algorithm>

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

// ---------------------------------------------------------------------------
// Records into block order without global atomics: an MSD radix sort of the
// block id (3 log2(nx/8) bits) in passes of at most 7 bits, each pass
// splitting every bucket of the previous one into at most 128.  A pass works
// on tiles of CS_T elements inside one bucket.  A workgroup counts its tile
// per digit in LDS; the counts of all tiles are scanned (bucket by bucket,
// digit-major, so that the scan gives each (digit, tile) its first slot);
// the workgroup then places its tile in chunks of CS_C elements: a counting
// sort of the chunk in LDS followed by writes in digit order, so that
// consecutive lanes write consecutive slots.  A chunk's run of one digit
// averages CS_C / 128 = 64 elements, so its writes are mostly whole 128-byte
// lines (runs of 16 left partial lines, which cost a read-modify-write).
// The block ids, records and labels of a chunk pass through one LDS buffer
// in turn.  The first pass reads the positions and builds the records.  The
// order inside a block is not kept: the block sort below orders it by cell.
// ---------------------------------------------------------------------------
constexpr int32_t CS_T = 1 << 16;  // elements per tile
constexpr int CS_TH = 1024;        // threads of the scatter
constexpr int CS_C = 8192;         // elements per LDS chunk (128 KB of records)
constexpr int CS_U = CS_C / CS_TH; // per thread
constexpr int CS_DB = 7;           // digit bits per pass

// tile t of a pass: bucket b, its j-th tile, elements [k0, k1)
__device__ inline bool cs_tile(int32_t t, const int32_t *tb,
                               const int32_t *tmap, const int32_t *bs,
                               int32_t nbk, int32_t *b, int32_t *j,
                               int32_t *k0, int32_t *k1) {
  if (t >= tb[nbk]) return false;
  *b = tmap[t];
  *j = t - tb[*b];
  *k0 = bs[*b] + *j * CS_T;
  *k1 = min(bs[*b + 1], *k0 + CS_T);
  return true;
}

// element k of a pass: its record, block and label
template <bool FROM_POS>
__device__ inline void cs_load(int64_t k, const float *pos, const int32_t *lab,
                               const float4 *sin, const int32_t *bin,
                               const int32_t *lin, int32_t nx, float4 *r,
                               int32_t *blk, int32_t *lb) {
  if (FROM_POS) {
    const PartCell p = part_cell(pos, k, nx);
    r->x = p.xs - (float)(p.cx & ~7);
    r->y = p.ys - (float)(p.cy & ~7);
    r->z = p.zs - (float)(p.cz & ~7);
    r->w = __int_as_float((int32_t)k);
    *blk = p.blk;
    *lb = lab ? lab[k] : -1;
  } else {
    // component-wise, so that the chunk's records stay in registers
    const float4 v = sin[k];
    r->x = v.x;
    r->y = v.y;
    r->z = v.z;
    r->w = v.w;
    *blk = bin[k];
    *lb = lin ? lin[k] : -1;
  }
}

template <bool FROM_POS>
__global__ void __launch_bounds__(256)
k_cs_hist(const float *__restrict__ pos, const int32_t *__restrict__ bin,
          int32_t nx, int shift, int dbits, int32_t nbk,
          const int32_t *__restrict__ bs, const int32_t *__restrict__ ntb,
          const int32_t *__restrict__ tb, const int32_t *__restrict__ tmap,
          int32_t *__restrict__ H) {
  __shared__ int32_t hist[1 << CS_DB];
  int32_t b, j, k0, k1;
  if (!cs_tile(blockIdx.x, tb, tmap, bs, nbk, &b, &j, &k0, &k1)) return;
  const int32_t nd = 1 << dbits;
  for (int d = threadIdx.x; d < nd; d += 256) hist[d] = 0;
  __syncthreads();
  for (int32_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const int32_t blk = FROM_POS ? part_cell(pos, k, nx).blk : bin[k];
    atomicAdd(&hist[(blk >> shift) & (nd - 1)], 1);
  }
  __syncthreads();
  int32_t *reg = H + (int64_t)tb[b] * nd;  // bucket b: digit-major, tiles minor
  for (int d = threadIdx.x; d < nd; d += 256) reg[(int64_t)d * ntb[b] + j] = hist[d];
}

template <bool FROM_POS>
__global__ void __launch_bounds__(CS_TH)
k_cs_scatter(const float *__restrict__ pos, const int32_t *__restrict__ lab,
             const float4 *__restrict__ sin, const int32_t *__restrict__ bin,
             const int32_t *__restrict__ lin, int32_t nx, int shift, int dbits,
             int32_t nbk, const int32_t *__restrict__ bs,
             const int32_t *__restrict__ ntb, const int32_t *__restrict__ tb,
             const int32_t *__restrict__ tmap, const int32_t *__restrict__ O,
             float4 *__restrict__ sout, int32_t *__restrict__ bout,
             int32_t *__restrict__ lout) {
  __shared__ float4 ex[CS_C];  // the chunk in digit order, one array at a time
  __shared__ int32_t gfill[1 << CS_DB], lcnt[1 << CS_DB], lst[1 << CS_DB];
  int32_t *exi = reinterpret_cast<int32_t *>(ex);
  int32_t b, j, k0, k1;
  if (!cs_tile(blockIdx.x, tb, tmap, bs, nbk, &b, &j, &k0, &k1)) return;
  const int32_t nd = 1 << dbits;
  const int32_t *reg = O + (int64_t)tb[b] * nd;
  for (int d = threadIdx.x; d < nd; d += CS_TH)
    gfill[d] = reg[(int64_t)d * ntb[b] + j];
  for (int32_t c0 = k0; c0 < k1; c0 += CS_C) {
    const int32_t cn = min(CS_C, k1 - c0);
    for (int d = threadIdx.x; d < nd; d += CS_TH) lcnt[d] = 0;
    __syncthreads();
    float4 rr[CS_U];
    int32_t bb[CS_U], ll[CS_U], slot[CS_U], dst[CS_U];
#pragma unroll
    for (int u = 0; u < CS_U; u++) {
      const int32_t i = threadIdx.x + CS_TH * u;
      slot[u] = 0;
      if (i < cn) {
        cs_load<FROM_POS>(c0 + i, pos, lab, sin, bin, lin, nx, &rr[u], &bb[u],
                          &ll[u]);
        slot[u] = atomicAdd(&lcnt[(bb[u] >> shift) & (nd - 1)], 1);
      }
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the chunk's counts (nd <= 128)
      const int d0 = 2 * threadIdx.x;
      const int32_t v0 = d0 < nd ? lcnt[d0] : 0, v1 = d0 + 1 < nd ? lcnt[d0 + 1] : 0;
      const int32_t inc = wave_incl_scan(v0 + v1);
      if (d0 < nd) lst[d0] = inc - v0 - v1;
      if (d0 + 1 < nd) lst[d0 + 1] = inc - v1;
    }
    __syncthreads();
    // block ids to their slots in digit order; each thread then takes the
    // slots i = threadIdx.x + CS_TH * u and finds their destinations
#pragma unroll
    for (int u = 0; u < CS_U; u++)
      if (threadIdx.x + CS_TH * u < cn) {
        slot[u] += lst[(bb[u] >> shift) & (nd - 1)];
        exi[slot[u]] = bb[u];
      }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CS_U; u++) {
      const int32_t i = threadIdx.x + CS_TH * u;
      if (i < cn) {
        const int32_t bk = exi[i];
        const int32_t d = (bk >> shift) & (nd - 1);
        dst[u] = gfill[d] + (i - lst[d]);
        bout[dst[u]] = bk;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CS_U; u++)
      if (threadIdx.x + CS_TH * u < cn) ex[slot[u]] = rr[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CS_U; u++) {
      const int32_t i = threadIdx.x + CS_TH * u;
      if (i < cn) sout[dst[u]] = ex[i];
    }
    if (lout) {
      __syncthreads();
#pragma unroll
      for (int u = 0; u < CS_U; u++)
        if (threadIdx.x + CS_TH * u < cn) exi[slot[u]] = ll[u];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < CS_U; u++) {
        const int32_t i = threadIdx.x + CS_TH * u;
        if (i < cn) lout[dst[u]] = exi[i];
      }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < nd; d += CS_TH) gfill[d] += lcnt[d];
    __syncthreads();
  }
}

__global__ void k_set_pair(int32_t *__restrict__ p, int32_t a, int32_t b) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    p[0] = a;
    p[1] = b;
  }
}

// tiles per bucket (entry nbk is not used by the scan's total)
__global__ void k_cs_ntiles(const int32_t *__restrict__ bs, int32_t nbk,
                            int32_t *__restrict__ ntb) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbk) ntb[b] = (bs[b + 1] - bs[b] + CS_T - 1) / CS_T;
  else if (b == nbk) ntb[b] = 0;
}

__global__ void k_cs_tmap(const int32_t *__restrict__ ntb,
                          const int32_t *__restrict__ tb, int32_t nbk,
                          int32_t *__restrict__ tmap) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbk) return;
  for (int32_t j = 0; j < ntb[b]; j++) tmap[tb[b] + j] = b;
}

// starts of the buckets of the next pass: bucket b splits into nd
__global__ void k_cs_split(const int32_t *__restrict__ O, int dbits,
                           int32_t nbk, const int32_t *__restrict__ bs,
                           const int32_t *__restrict__ ntb,
                           const int32_t *__restrict__ tb, int64_t n,
                           int32_t *__restrict__ bs2) {
  const int64_t B = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t nd = 1 << dbits;
  const int64_t nb2 = (int64_t)nbk * nd;
  if (B > nb2) return;
  if (B == nb2) {
    bs2[B] = (int32_t)n;
    return;
  }
  const int32_t b = (int32_t)(B >> dbits), d = (int32_t)(B & (nd - 1));
  bs2[B] = ntb[b] > 0 ? O[(int64_t)tb[b] * nd + (int64_t)d * ntb[b]] : bs[b];
}

// Dense cells (at least DENSE records) are reordered after the sort; those
// of at most SUB_REG records fit the registers of one workgroup.
constexpr int DENSE = 64;
// cells of at most SUB_REG records are sorted in registers, SUB_V records
// per thread: one wave per cell up to SUB_W records, SUB_T threads above
constexpr int SUB_V = 8, SUB_T = 512, SUB_REG = SUB_T * SUB_V,
              SUB_W = 64 * SUB_V;

// One workgroup per block: counting sort of its records by local cell, the
// final records and the block's 512 cell offsets.  A block of at most BS_CAP
// records is read once into registers, placed in LDS in cell order and
// written out as one contiguous run; a larger one is listed for
// k_block_sort_big, which reads it twice (count, then place) and writes it
// record by record.
constexpr int BS_CAP = 1024;  // 22 KB of LDS: 6 workgroups per CU (by registers)
constexpr int BS_U = BS_CAP / 256;

__device__ inline int32_t local_cell(const float4 &r) {
  return ((int32_t)r.z * 8 + (int32_t)r.y) * 8 + (int32_t)r.x;
}

// The block's cell offsets from its 512 counts, run by the first wave: cnt
// becomes the exclusive scan, cell_start[b * 512 + i] = a + scan, and the
// dense cells go to their lists by size (one atomic per block and list).
__device__ inline void block_cells(int32_t *cnt, int64_t b, int32_t a,
                                   int32_t *__restrict__ cell_start,
                                   int32_t *__restrict__ dlist,
                                   int32_t *__restrict__ dmid,
                                   int32_t *__restrict__ dbig,
                                   int32_t *__restrict__ dcnt, int32_t dcap) {
  // exclusive scan of 512 counters by one wave
  int32_t v[8], s = 0;
  for (int k = 0; k < 8; k++) { v[k] = cnt[threadIdx.x * 8 + k]; s += v[k]; }
  const int32_t inc = wave_incl_scan(s);
  int32_t e = inc - s;
  int nr = 0, nm = 0, ng = 0;  // this lane's dense cells, by list
  for (int k = 0; k < 8; k++) {
    cnt[threadIdx.x * 8 + k] = e;
    cell_start[b * 512 + threadIdx.x * 8 + k] = a + e;
    e += v[k];
    nr += v[k] >= DENSE && v[k] <= SUB_W;
    nm += v[k] > SUB_W && v[k] <= SUB_REG;
    ng += v[k] > SUB_REG;
  }
  // dense cells to their lists: one atomic per block and list
  const int32_t ir = wave_incl_scan(nr), im = wave_incl_scan(nm),
                ig = wave_incl_scan(ng);
  const int32_t tr = __shfl(ir, 63, 64), tm = __shfl(im, 63, 64),
                tg = __shfl(ig, 63, 64);
  int32_t br = 0, bm = 0, bg = 0;
  if (threadIdx.x == 0) {
    if (tr > 0) br = atomicAdd(&dcnt[0], tr);
    if (tg > 0) bg = atomicAdd(&dcnt[2], tg);
    if (tm > 0) bm = atomicAdd(&dcnt[4], tm);
  }
  br = __shfl(br, 0, 64) + ir - nr;
  bm = __shfl(bm, 0, 64) + im - nm;
  bg = __shfl(bg, 0, 64) + ig - ng;
  for (int k = 0; k < 8; k++) {
    const int32_t cell = (int32_t)(b * 512 + threadIdx.x * 8 + k);
    if (v[k] >= DENSE && v[k] <= SUB_W) {
      if (br < dcap) dlist[br] = cell;
      br++;
    } else if (v[k] > SUB_W && v[k] <= SUB_REG) {
      if (bm < dcap) dmid[bm] = cell;
      bm++;
    } else if (v[k] > SUB_REG) {
      if (bg < dcap) dbig[bg] = cell;
      bg++;
    }
  }
}

__global__ void __launch_bounds__(256)
k_block_sort(const int32_t *__restrict__ bstart, int64_t nblk,
             const float4 *__restrict__ stage,
             const int32_t *__restrict__ stage_lab,
             int32_t *__restrict__ cell_start, float4 *__restrict__ rec,
             int32_t *__restrict__ perm, int32_t *__restrict__ lab_sorted,
             int32_t *__restrict__ dlist, int32_t *__restrict__ dmid,
             int32_t *__restrict__ dbig, int32_t *__restrict__ dcnt,
             int32_t dcap, int32_t *__restrict__ bigblk) {
  __shared__ int32_t cnt[512];
  __shared__ float4 sr[BS_CAP];
  __shared__ int32_t sl[BS_CAP];
  // A staged block's records are requested one block ahead (while the
  // block before is placed), and block ranges two ahead, so that the loads
  // of a block are not waited for one after the other.
  auto load = [&](int32_t a, int32_t n, float4 *rr, int32_t *ll) {
    if (n > BS_CAP) return;
#pragma unroll
    for (int k = 0; k < BS_U; k++) {
      const int32_t t = threadIdx.x + 256 * k;
      if (t < n) {
        const float4 v = stage[a + t];
        rr[k].x = v.x;
        rr[k].y = v.y;
        rr[k].z = v.z;
        rr[k].w = v.w;
        ll[k] = stage_lab ? stage_lab[a + t] : 0;
      }
    }
  };
  const int64_t G = gridDim.x;
  int64_t b = blockIdx.x;
  int32_t a = 0, n = 0, an = 0, nn = 0;
  if (b < nblk) {
    a = bstart[b];
    n = bstart[b + 1] - a;
  }
  if (b + G < nblk) {
    an = bstart[b + G];
    nn = bstart[b + G + 1] - an;
  }
  float4 rr[BS_U];
  int32_t ll[BS_U];
  load(a, n, rr, ll);
  for (; b < nblk; b += G) {
    float4 rn[BS_U];
    int32_t ln[BS_U];
    load(an, nn, rn, ln);
    int32_t a2 = 0, n2 = 0;
    if (b + 2 * G < nblk) {
      a2 = bstart[b + 2 * G];
      n2 = bstart[b + 2 * G + 1] - a2;
    }
    if (n > BS_CAP) {
      // placed by k_block_sort_big
      if (threadIdx.x == 0) bigblk[atomicAdd(&dcnt[6], 1)] = (int32_t)b;
    } else {
      for (int i = threadIdx.x; i < 512; i += 256) cnt[i] = 0;
      __syncthreads();
      int32_t dd[BS_U];
#pragma unroll
      for (int k = 0; k < BS_U; k++) {
        const int32_t t = threadIdx.x + 256 * k;
        if (t < n) dd[k] = atomicAdd(&cnt[local_cell(rr[k])], 1);
      }
      __syncthreads();
      if (threadIdx.x < 64)
        block_cells(cnt, b, a, cell_start, dlist, dmid, dbig, dcnt, dcap);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < BS_U; k++) {
        const int32_t t = threadIdx.x + 256 * k;
        if (t < n) {
          const int32_t p = cnt[local_cell(rr[k])] + dd[k];
          sr[p] = rr[k];
          sl[p] = ll[k];
        }
      }
      __syncthreads();
      for (int32_t t = threadIdx.x; t < n; t += 256) {
        const float4 r = sr[t];
        rec[a + t] = r;
        if (perm) perm[a + t] = __float_as_int(r.w);
        if (lab_sorted) lab_sorted[a + t] = sl[t];
      }
      __syncthreads();
    }
    a = an;
    n = nn;
    an = a2;
    nn = n2;
#pragma unroll
    for (int k = 0; k < BS_U; k++) {
      rr[k] = rn[k];
      ll[k] = ln[k];
    }
  }
}

// Blocks of more than BS_CAP records, one BB_T-thread workgroup each: the
// records are counted per cell, then placed from HBM, each pass issuing a
// thread's BB_U records before using them.  Placed by the 256-thread block
// sort with one record in flight per thread, the zoom catalogue's central
// blocks (up to ~10^5 records each) held that launch for 30 ms.
constexpr int BB_T = 1024, BB_U = 4;

__global__ void __launch_bounds__(BB_T)
k_block_sort_big(const int32_t *__restrict__ bstart,
                 const float4 *__restrict__ stage,
                 const int32_t *__restrict__ stage_lab,
                 int32_t *__restrict__ cell_start, float4 *__restrict__ rec,
                 int32_t *__restrict__ perm, int32_t *__restrict__ lab_sorted,
                 int32_t *__restrict__ dlist, int32_t *__restrict__ dmid,
                 int32_t *__restrict__ dbig, int32_t *__restrict__ dcnt,
                 int32_t dcap, const int32_t *__restrict__ bigblk) {
  __shared__ int32_t cnt[512];
  __shared__ int32_t item;
  const int32_t nbig = dcnt[6];
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(&dcnt[7], 1);
    for (int i = threadIdx.x; i < 512; i += BB_T) cnt[i] = 0;
    __syncthreads();
    const int32_t it = item;
    if (it >= nbig) break;
    const int64_t b = bigblk[it];
    const int32_t a = bstart[b], n = bstart[b + 1] - a;
    for (int32_t t0 = threadIdx.x; t0 < n; t0 += BB_T * BB_U) {
      float4 r[BB_U];
#pragma unroll
      for (int k = 0; k < BB_U; k++)
        if (t0 + k * BB_T < n) r[k] = stage[a + t0 + k * BB_T];
#pragma unroll
      for (int k = 0; k < BB_U; k++)
        if (t0 + k * BB_T < n) atomicAdd(&cnt[local_cell(r[k])], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64)
      block_cells(cnt, b, a, cell_start, dlist, dmid, dbig, dcnt, dcap);
    __syncthreads();
    for (int32_t t0 = threadIdx.x; t0 < n; t0 += BB_T * BB_U) {
      float4 r[BB_U];
      int32_t lv[BB_U];
#pragma unroll
      for (int k = 0; k < BB_U; k++) {
        const int32_t t = t0 + k * BB_T;
        if (t < n) {
          r[k] = stage[a + t];
          lv[k] = stage_lab ? stage_lab[a + t] : 0;
        }
      }
#pragma unroll
      for (int k = 0; k < BB_U; k++) {
        if (t0 + k * BB_T < n) {
          const int32_t d = a + atomicAdd(&cnt[local_cell(r[k])], 1);
          rec[d] = r[k];
          if (perm) perm[d] = __float_as_int(r[k].w);
          if (lab_sorted) lab_sorted[d] = lv[k];
        }
      }
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

// Cells of at most T * SUB_V records (BIG false): the records read once into
// registers (SUB_V per thread) and placed in place; one wave per cell up to
// SUB_W records, SUB_T threads up to SUB_REG (8 records per thread keep the
// registers to a quarter of the file: 16 took all of it).  Larger cells (BIG true):
// SUB_TB threads, staged through the workgroup's slice, each pass issuing a
// thread's SUB_U records before using them (that launch: 0.88 -> 0.42 ms at
// 1024^3).  A cell's steps are serial round trips to memory, so the
// launches over the many small cells are bound by the cells in flight: one
// 256-thread workgroup per cell of up to 4096 records took 5.7 ms.
constexpr int SUB_TB = 1024, SUB_U = 4;

template <int T, bool BIG>
__global__ void __launch_bounds__(T)
k_subsort(const int32_t *__restrict__ cs, const int32_t *__restrict__ list,
          const int32_t *__restrict__ nlist, int32_t cap,
          int32_t *__restrict__ queue,
          float4 *__restrict__ rec, int32_t *__restrict__ perm,
          int32_t *__restrict__ lab, float4 *__restrict__ trec,
          int32_t *__restrict__ tperm, int32_t *__restrict__ tlab,
          int64_t tcap) {
  __shared__ int32_t hist[512];
  __shared__ int32_t item;
  const int32_t nl = min(*nlist, cap);
  float4 *tr = BIG ? trec + (int64_t)blockIdx.x * tcap : nullptr;
  int32_t *tp = BIG && tperm ? tperm + (int64_t)blockIdx.x * tcap : nullptr;
  int32_t *tl = BIG && tlab ? tlab + (int64_t)blockIdx.x * tcap : nullptr;
  // cells are taken SUB_B at a time: one counter shared by every workgroup
  // of the launch, hundreds of thousands of small cells
  constexpr int SUB_B = BIG ? 1 : 16;
  int32_t it = 0, end = 0;
  for (;;) {
    if (it == end) {
      if (threadIdx.x == 0) item = atomicAdd(queue, SUB_B);
      __syncthreads();
      it = item;
      end = it + SUB_B;
      __syncthreads();
    }
    for (int i = threadIdx.x; i < 512; i += T) hist[i] = 0;
    __syncthreads();
    if (it >= nl) break;
    const int32_t c = list[it++];
    const int32_t a = cs[c], n = cs[c + 1] - a;
    if (BIG ? n > tcap : n > T * SUB_V) {
      // would not fit: keep the scatter order
      __syncthreads();
      continue;
    }
    // the register path reads its records here, once
    float4 rr[BIG ? 1 : SUB_V];
    int32_t pp[BIG ? 1 : SUB_V], ll[BIG ? 1 : SUB_V];
    if constexpr (BIG) {
      for (int32_t t0 = threadIdx.x; t0 < n; t0 += T * SUB_U) {
        float4 r[SUB_U];
#pragma unroll
        for (int k = 0; k < SUB_U; k++)
          if (t0 + k * T < n) r[k] = rec[a + t0 + k * T];
#pragma unroll
        for (int k = 0; k < SUB_U; k++)
          if (t0 + k * T < n) atomicAdd(&hist[morton9(r[k])], 1);
      }
    } else {
#pragma unroll
      for (int k = 0; k < SUB_V; k++) {
        const int32_t t = threadIdx.x + k * T;
        if (t < n) {
          rr[k] = rec[a + t];
          pp[k] = perm ? perm[a + t] : 0;
          ll[k] = lab ? lab[a + t] : 0;
        }
      }
#pragma unroll
      for (int k = 0; k < SUB_V; k++)
        if (threadIdx.x + k * T < n) atomicAdd(&hist[morton9(rr[k])], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of 512 counters by one wave
      int32_t v[8], s = 0;
      for (int k = 0; k < 8; k++) { v[k] = hist[threadIdx.x * 8 + k]; s += v[k]; }
      const int32_t inc = wave_incl_scan(s);
      int32_t e = inc - s;
      for (int k = 0; k < 8; k++) { hist[threadIdx.x * 8 + k] = e; e += v[k]; }
    }
    __syncthreads();
    if constexpr (!BIG) {
      // in place: everything was read into registers before any write
      int32_t dd[SUB_V];
#pragma unroll
      for (int k = 0; k < SUB_V; k++)
        if (threadIdx.x + k * T < n) dd[k] = atomicAdd(&hist[morton9(rr[k])], 1);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < SUB_V; k++) {
        const int32_t t = threadIdx.x + k * T;
        if (t < n) {
          rec[a + dd[k]] = rr[k];
          if (perm) perm[a + dd[k]] = pp[k];
          if (lab) lab[a + dd[k]] = ll[k];
        }
      }
    } else {
      for (int32_t t0 = threadIdx.x; t0 < n; t0 += T * SUB_U) {
        float4 r[SUB_U];
        int32_t pv[SUB_U], lv[SUB_U];
#pragma unroll
        for (int k = 0; k < SUB_U; k++) {
          const int32_t t = t0 + k * T;
          if (t < n) {
            r[k] = rec[a + t];
            pv[k] = perm ? perm[a + t] : 0;
            lv[k] = lab ? lab[a + t] : 0;
          }
        }
#pragma unroll
        for (int k = 0; k < SUB_U; k++) {
          if (t0 + k * T < n) {
            const int32_t d = atomicAdd(&hist[morton9(r[k])], 1);
            tr[d] = r[k];
            if (perm) tp[d] = pv[k];
            if (lab) tl[d] = lv[k];
          }
        }
      }
      __threadfence_block();
      __syncthreads();
      for (int32_t t0 = threadIdx.x; t0 < n; t0 += T * SUB_U) {
        float4 r[SUB_U];
        int32_t pv[SUB_U], lv[SUB_U];
#pragma unroll
        for (int k = 0; k < SUB_U; k++) {
          const int32_t t = t0 + k * T;
          if (t < n) {
            r[k] = tr[t];
            pv[k] = perm ? tp[t] : 0;
            lv[k] = lab ? tl[t] : 0;
          }
        }
#pragma unroll
        for (int k = 0; k < SUB_U; k++) {
          const int32_t t = t0 + k * T;
          if (t < n) {
            rec[a + t] = r[k];
            if (perm) perm[a + t] = pv[k];
            if (lab) lab[a + t] = lv[k];
          }
        }
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
  const int64_t n = npart;
  int bits = 0;
  while ((int64_t)1 << bits < nblk) bits++;
  // at least one pass (a 0-bit digit when there is a single block)
  const int npass = bits > 0 ? (bits + CS_DB - 1) / CS_DB : 1;
  const int64_t ntile0 = (n + CS_T - 1) / CS_T;
  // bucket tables (two generations), tiles, histograms
  const int64_t maxt = ntile0 + nblk / 2 + 64;  // tiles of any pass, at most
  const int64_t nH = maxt << CS_DB;
  int32_t *hb = (int32_t *)psod_scratch(
      ctx, S_CS_H, (size_t)(2 * (nblk + 64) + 2 * (nblk + 64) + maxt + 64 +
                            2 * nH + 256) * 4);
  float4 *sa = (float4 *)psod_scratch(ctx, S_CS_A, (size_t)(n > 0 ? n : 1) * 24);
  float4 *sc = (float4 *)psod_scratch(ctx, S_CS_C, (size_t)(n > 0 ? n : 1) * 24);
  if (!hb || !sa || !sc) return PSOD_E_NOMEM;
  int32_t *bsA = hb, *bsB = bsA + nblk + 64, *ntb = bsB + nblk + 64,
          *tb = ntb + nblk + 64, *tmap = tb + nblk + 64, *H = tmap + maxt + 64,
          *O = H + nH;
  // record buffers: float4 records, then block ids, then labels
  float4 *recb[2] = {sa, sc};
  int32_t *blkb[2] = {(int32_t *)(sa + n), (int32_t *)(sc + n)};
  int32_t *labb[2] = {d_lab_sorted ? blkb[0] + n : nullptr,
                      d_lab_sorted ? blkb[1] + n : nullptr};
  int32_t *bs = bsA, *bs2 = bsB;
  int32_t nbk = 1;
  int cur = 0;  // buffer holding the last pass's output
  k_set_pair<<<1, 64, 0, st>>>(bs, 0, (int32_t)n);  // one bucket: [0, n)
  PSOD_LAUNCH_CHECK();
  for (int p = 0; p < npass && n > 0; p++) {
    const int hi = bits - CS_DB * p;
    const int dbits = hi < CS_DB ? hi : CS_DB;
    const int shift = hi - dbits;
    const int32_t nd = 1 << dbits;
    const int64_t mt = ntile0 + nbk;  // tiles of this pass, at most
    k_cs_ntiles<<<grid_for(nbk + 1, 256), 256, 0, st>>>(bs, nbk, ntb);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(psod_scan_i32(ctx, ntb, tb, nbk + 1, nullptr));
    k_cs_tmap<<<grid_for(nbk, 256), 256, 0, st>>>(ntb, tb, nbk, tmap);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(psod_fill_i32(ctx, H, mt * nd, 0));
    const int src = cur, dst = p == 0 ? 0 : 1 - cur;
    if (p == 0) {
      k_cs_hist<true><<<(unsigned)mt, 256, 0, st>>>(d_pos, nullptr, nx, shift,
                                                    dbits, nbk, bs, ntb, tb,
                                                    tmap, H);
    } else {
      k_cs_hist<false><<<(unsigned)mt, 256, 0, st>>>(nullptr, blkb[src], nx,
                                                     shift, dbits, nbk, bs,
                                                     ntb, tb, tmap, H);
    }
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(psod_scan_i32(ctx, H, O, mt * nd, nullptr));
    if (p == 0) {
      k_cs_scatter<true><<<(unsigned)mt, CS_TH, 0, st>>>(
          d_pos, d_labels, nullptr, nullptr, nullptr, nx, shift, dbits, nbk,
          bs, ntb, tb, tmap, O, recb[dst], blkb[dst], labb[dst]);
    } else {
      k_cs_scatter<false><<<(unsigned)mt, CS_TH, 0, st>>>(
          nullptr, nullptr, recb[src], blkb[src], labb[src], nx, shift, dbits,
          nbk, bs, ntb, tb, tmap, O, recb[dst], blkb[dst], labb[dst]);
    }
    PSOD_LAUNCH_CHECK();
    k_cs_split<<<grid_for((int64_t)nbk * nd + 1, 256), 256, 0, st>>>(
        O, dbits, nbk, bs, ntb, tb, n, bs2);
    PSOD_LAUNCH_CHECK();
    nbk *= nd;
    int32_t *t = bs; bs = bs2; bs2 = t;
    cur = dst;
  }
  if (n == 0) PSOD_TRY(psod_fill_i32(ctx, bs, nblk + 1, 0));
  const int32_t *bstart = bs;  // nblk + 1 block starts
  const float4 *stage = recb[cur];
  const int32_t *slab = labb[cur];
  // lists of the dense cells by size, filled by the block sort: ctl[0],
  // ctl[4] and ctl[2] count them (up to SUB_W, up to SUB_REG, larger
  // records), ctl[1], ctl[5] and ctl[3] are the subsort's queues (at most
  // npart / DENSE cells are dense)
  const int32_t cap = (int32_t)std::min<int64_t>(ncell, npart / DENSE + 32);
  // blocks of more than BS_CAP records: listed by the block sort (count
  // in ctl[6], queue ctl[7]) and placed by k_block_sort_big
  int32_t *list = (int32_t *)psod_scratch(
      ctx, S_HIST, (size_t)(3 * (int64_t)cap + 96 + nblk + 32) * 4);
  int32_t *ctl = (int32_t *)psod_scratch(ctx, S_MISC, 64);
  if (!list || !ctl) return PSOD_E_NOMEM;
  int32_t *big = list + cap + 32;
  int32_t *mid = big + cap + 32;
  int32_t *bigblk = mid + cap + 32;
  PSOD_TRY(psod_fill_i32(ctx, ctl, 8, 0));
  k_block_sort<<<grid_for(nblk, 1, (int64_t)ctx->num_cus * 32), 256, 0, st>>>(
      bstart, nblk, stage, slab, d_cell_start, (float4 *)d_rec, d_perm,
      d_lab_sorted, list, mid, big, ctl, cap, bigblk);
  PSOD_LAUNCH_CHECK();
  k_block_sort_big<<<(unsigned)ctx->num_cus, BB_T, 0, st>>>(
      bstart, stage, slab, d_cell_start, (float4 *)d_rec, d_perm,
      d_lab_sorted, list, mid, big, ctl, cap, bigblk);
  PSOD_LAUNCH_CHECK();
  k_set_last<<<1, 64, 0, st>>>(d_cell_start, ncell, (int32_t)npart);
  PSOD_LAUNCH_CHECK();
  if (npart > 0) {
    // Morton order of the 1/8-cell sub-cells inside dense cells
    // cells that fit the registers: one wave per cell up to SUB_W records,
    // SUB_T threads up to SUB_REG, no staging slices; larger cells: one
    // workgroup per CU, each with a staging slice
    k_subsort<64, false><<<(unsigned)ctx->num_cus * 12, 64, 0, st>>>(
        d_cell_start, list, ctl, cap, ctl + 1, (float4 *)d_rec, d_perm,
        d_lab_sorted, nullptr, nullptr, nullptr, 0);
    PSOD_LAUNCH_CHECK();
    k_subsort<SUB_T, false><<<(unsigned)ctx->num_cus * 2, SUB_T, 0, st>>>(
        d_cell_start, mid, ctl + 4, cap, ctl + 5, (float4 *)d_rec, d_perm,
        d_lab_sorted, nullptr, nullptr, nullptr, 0);
    PSOD_LAUNCH_CHECK();
    // cells above SUB_REG records: at most npart / SUB_REG of them, each
    // staged in a slice of at most min(SUB_TMP, npart) records
    const int64_t tcap = std::min<int64_t>(SUB_TMP, npart);
    const unsigned nsub = (unsigned)std::min<int64_t>(ctx->num_cus,
                                                      npart / SUB_REG + 1);
    char *tmp = (char *)psod_scratch(ctx, S_KEYS, (size_t)nsub * tcap * 24);
    if (!tmp) return PSOD_E_NOMEM;
    k_subsort<SUB_TB, true><<<nsub, SUB_TB, 0, st>>>(
        d_cell_start, big, ctl + 2, cap, ctl + 3, (float4 *)d_rec, d_perm,
        d_lab_sorted, (float4 *)tmp, (int32_t *)(tmp + (size_t)nsub * tcap * 16),
        (int32_t *)(tmp + (size_t)nsub * tcap * 20), tcap);
    PSOD_LAUNCH_CHECK();
  }
  return psod_finish(ctx);
}
