// cic.hip -- a2 CIC deposit, a3 CIC^-1 (+ FoF seeds), a4 density order.
//
// Deposit (main file:1077-1265).  One workgroup owns the (8*TB*R)^3 nodes of
// a tile of TB^3 8^3-cell blocks (R = ng/nx in {1, 2}).  A node receives
// mass from the cells on both sides of it, so the workgroup reads its own
// blocks' particles (TB^2 contiguous runs of the cell-sorted records) plus
// the particles of the lower-neighbour boundary cells, and accumulates only
// the nodes it owns.  Every node is written exactly once with a plain
// store: no zero-fill and no global atomics.
//
// Node sums are 64-bit fixed point (weights * 2^36) in LDS, exact sums of
// the float32 weights and independent of particle order.  A particle's
// offset in its CIC cell is exact: the record holds u in [0, 8) relative to
// its 8^3 block, and R*u - floor(R*u) is exact for R in {1, 2}.  Each lane
// takes K consecutive particles of the sorted list and keeps the 8 node sums
// in registers while they share the same lower node (dense cells), adding
// them to LDS only when it changes (the upper x plane stays in registers
// when the next particle lies in the next cell along x).  Tiles are dispatched densest first, so
// that the long tiles of halo cores overlap the rest.
#include "psod_internal.h"

#include <algorithm>
#include <mutex>

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);
int psod_fill_u64(psod_ctx *c, uint64_t *p, int64_t n, uint64_t v);

// weights in 2^-36 units: a node holds up to 2^28 unit weights
constexpr float FIX_SCALE = 68719476736.0f;  // 2^36
constexpr double FIX_INV = 1.0 / 68719476736.0;

// Tile of TB^3 blocks (8*TB cells per axis), R nodes per cell per axis.
// Halo ranges: the boundary cells of the lower-neighbour blocks of the tile,
// enumerated per neighbour block type; see tile_seg().
template <int TB>
struct TileSegs {
  static constexpr int NA = TB * TB * 64;  // x-1 face: single cells
  static constexpr int NB = TB * TB * 8;   // y-1 face: rows of 8 cells
  static constexpr int NC = TB * TB;       // z-1 face: 64-cell slabs
  static constexpr int ND = TB * 8;        // x-1,y-1 edge: single cells
  static constexpr int NE = TB * 8;        // x-1,z-1 edge: single cells
  static constexpr int NF = TB;            // y-1,z-1 edge: rows of 8
  static constexpr int N = NA + NB + NC + ND + NE + NF + 1;
};

// s-th halo range of a tile: neighbour block offset (in blocks, relative to
// the tile's first block), local cell key inside it and run length
template <int TB>
__device__ inline void tile_seg(int s, int *bi, int *bj, int *bk, int *loc,
                                int *len) {
  using T = TileSegs<TB>;
  if (s < T::NA) {  // lx=7 of block (-1, j, k); cell (ly, lz)
    const int j = (s >> 6) % TB, k = (s >> 6) / TB, c = s & 63;
    *bi = -1; *bj = j; *bk = k;
    *loc = ((c >> 3) * 8 + (c & 7)) * 8 + 7; *len = 1; return;
  }
  s -= T::NA;
  if (s < T::NB) {  // ly=7 of block (i, -1, k); row lz
    const int i = (s >> 3) % TB, k = (s >> 3) / TB, lz = s & 7;
    *bi = i; *bj = -1; *bk = k;
    *loc = (lz * 8 + 7) * 8; *len = 8; return;
  }
  s -= T::NB;
  if (s < T::NC) {  // lz=7 of block (i, j, -1)
    *bi = s % TB; *bj = s / TB; *bk = -1;
    *loc = 7 * 64; *len = 64; return;
  }
  s -= T::NC;
  if (s < T::ND) {  // lx=7, ly=7 of block (-1, -1, k); cell lz
    *bi = -1; *bj = -1; *bk = s >> 3;
    *loc = ((s & 7) * 8 + 7) * 8 + 7; *len = 1; return;
  }
  s -= T::ND;
  if (s < T::NE) {  // lx=7, lz=7 of block (-1, j, -1); cell ly
    *bi = -1; *bj = s >> 3; *bk = -1;
    *loc = (7 * 8 + (s & 7)) * 8 + 7; *len = 1; return;
  }
  s -= T::NE;
  if (s < T::NF) {  // ly=7, lz=7 of block (i, -1, -1)
    *bi = s; *bj = -1; *bk = -1;
    *loc = (7 * 8 + 7) * 8; *len = 8; return;
  }
  *bi = -1; *bj = -1; *bk = -1;  // corner cell
  *loc = 511; *len = 1;
}

constexpr int DEP_T = 512;  // threads of a deposit workgroup (one tile)
constexpr int DEP_K = 4;  // consecutive particles per lane
constexpr int32_t DEP_CAPD = 8192;  // dense tiles dispatched first, at most

// Dense tiles (at least dense_tile own particles, a tunable) are listed so
// that they are dispatched first: their long runs then overlap the rest
// instead of trailing it.  listed[tile] = 1 for a tile in dlist.
__global__ void k_tile_dense(const int32_t *__restrict__ cell_start,
                             int32_t nb, int32_t TB, int32_t ntile, int32_t capd,
                             int32_t dense_tile,
                             int32_t *__restrict__ listed,
                             int32_t *__restrict__ dcount,
                             int32_t *__restrict__ dlist) {
  const int32_t tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= ntile) return;
  const int32_t nt = nb / TB;
  const int tx = tile % nt, ty = (tile / nt) % nt, tz = tile / (nt * nt);
  int32_t n = 0;
  for (int k = 0; k < TB; k++)
    for (int j = 0; j < TB; j++) {
      const int32_t b0 = ((tz * TB + k) * nb + (ty * TB + j)) * nb + tx * TB;
      n += cell_start[(b0 + TB) * BS3] - cell_start[b0 * BS3];
    }
  int32_t f = 0;
  if (n >= dense_tile) {
    const int32_t k = atomicAdd(dcount, 1);
    if (k < capd) {
      dlist[k] = tile;
      f = 1;
    }
  }
  listed[tile] = f;
}

// Tile of workgroup b: b < capd takes the b-th listed dense tile; the others
// take the remaining tiles in an XCD-aware order (workgroups go round-robin
// over the 8 XCDs, each with its own L2: XCD x sweeps a contiguous slab of
// tiles, so that a tile finds its neighbours' boundary cells in its L2).
__device__ inline int32_t tile_of_block(int32_t b, int32_t ntile, int32_t capd,
                                        const int32_t *dcount,
                                        const int32_t *dlist,
                                        const int32_t *listed) {
  if (b < capd) return b < min(*dcount, capd) ? dlist[b] : -1;
  const int32_t q = b - capd;
  const int32_t tile = ntile % 8 == 0 ? (q % 8) * (ntile / 8) + q / 8 : q;
  return (capd > 0 && listed[tile]) ? -1 : tile;
}

// Deposit of the (8*TB*R)^3 nodes owned by one tile.  Own particles come
// from TB^2 contiguous runs of TB blocks; halo particles from the ranges of
// tile_seg(), found by binary search on their prefix counts.
template <int R, int TB>
__global__ void __launch_bounds__(DEP_T)
k_cic_deposit(const int32_t *__restrict__ cell_start,
              const float4 *__restrict__ rec, int32_t nx, int32_t ng,
              int32_t shift, double norm, int32_t capd,
              const int32_t *__restrict__ dcount,
              const int32_t *__restrict__ dlist,
              const int32_t *__restrict__ listed, float *__restrict__ grid) {
  constexpr int NN = 8 * TB * R;  // owned nodes per axis
  constexpr int NS = TileSegs<TB>::N;
  __shared__ unsigned long long acc[NN * NN * NN];
  __shared__ int32_t hstart[NS];
  __shared__ int32_t hpre[NS + 1];
  __shared__ int8_t hoff[NS][3];
  __shared__ int32_t ostart[TB * TB], opre[TB * TB + 1];
  __shared__ int32_t obnd[TB * TB][TB];  // own-list position of each block
  __shared__ int32_t wtot[DEP_T / 64];
  const int32_t nb = nx >> 3;
  const int32_t nt = nb / TB;  // tiles per axis
  const int32_t tile = tile_of_block(blockIdx.x, nt * nt * nt, capd, dcount,
                                     dlist, listed);
  if (tile < 0) return;  // uniform per workgroup: before any barrier
  const int tx = tile % nt, ty = (tile / nt) % nt, tz = tile / (nt * nt);
  const int bx0 = tx * TB, by0 = ty * TB, bz0 = tz * TB;
  const int l = threadIdx.x & 63;
  for (int i = threadIdx.x; i < NN * NN * NN; i += DEP_T) acc[i] = 0ull;
  // own runs: blocks (bx0..bx0+TB-1, by0+j, bz0+k) are consecutive keys
  if (threadIdx.x < TB * TB) {
    const int j = threadIdx.x % TB, k = threadIdx.x / TB;
    const int32_t b0 = ((bz0 + k) * nb + (by0 + j)) * nb + bx0;
    ostart[threadIdx.x] = cell_start[b0 * BS3];
    for (int q = 0; q < TB; q++)
      obnd[threadIdx.x][q] = cell_start[(b0 + q) * BS3];
    opre[threadIdx.x + 1] = cell_start[(b0 + TB) * BS3];
  }
  // halo ranges and their prefix counts (blocks of DEP_T ranges)
  int32_t carry = 0;
  for (int base = 0; base < NS; base += DEP_T) {
    const int s = base + threadIdx.x;
    int32_t cnt = 0;
    if (s < NS) {
      int bi, bj, bk, loc, len;
      tile_seg<TB>(s, &bi, &bj, &bk, &loc, &len);
      const int32_t nbk = (wrapi(bz0 + bk, nb) * nb + wrapi(by0 + bj, nb)) * nb +
                          wrapi(bx0 + bi, nb);
      const int32_t k0 = nbk * BS3 + loc;
      const int32_t a = cell_start[k0];
      cnt = cell_start[k0 + len] - a;
      hstart[s] = a;
      // offset of the neighbour block relative to the tile origin, blocks
      hoff[s][0] = (int8_t)bi;
      hoff[s][1] = (int8_t)bj;
      hoff[s][2] = (int8_t)bk;
    }
    const int w = threadIdx.x >> 6;
    const int32_t inc = wave_incl_scan(cnt);
    if (l == 63) wtot[w] = inc;
    __syncthreads();
    int32_t off = carry;
    for (int i = 0; i < w; i++) off += wtot[i];
    if (s < NS) hpre[s] = off + inc - cnt;
    for (int i = 0; i < DEP_T / 64; i++) carry += wtot[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    hpre[NS] = carry;
    opre[0] = 0;
    // prefix of the own runs (their ends were stored in opre[1..]) and the
    // own-list position of every block
    int32_t p = 0;
    for (int r = 0; r < TB * TB; r++) {
      const int32_t n = opre[r + 1] - ostart[r];
      for (int q = 0; q < TB; q++) obnd[r][q] = p + (obnd[r][q] - ostart[r]);
      opre[r + 1] = p + n;
      p += n;
    }
  }
  __syncthreads();
  const int32_t nown = opre[TB * TB];
  const int32_t ntot = nown + hpre[NS];
  // register sums of the current lower node and their flush to LDS
  unsigned long long ra[8] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
  int32_t rid = -1, rix = 0, riy = 0, riz = 0;
  // the 4 nodes of x offset h of the current lower node
  auto flush = [&](int h) {
    if (rid < 0) return;
#pragma unroll
    for (int c = h; c < 8; c += 2) {
      const int ax = rix + (c & 1), ay = riy + ((c >> 1) & 1), az = riz + (c >> 2);
      if ((unsigned)ax < (unsigned)NN && (unsigned)ay < (unsigned)NN &&
          (unsigned)az < (unsigned)NN)
        atomicAdd(&acc[(az * NN + ay) * NN + ax], ra[c]);
    }
  };
  // lane l of wave w takes list positions base + w*64*K + l*K + j, j < K.
  // All K records of a lane are requested before the first is used, so that
  // a wave keeps K loads per lane in flight.
  const int wv = threadIdx.x >> 6;
  for (int32_t base = 0; base < ntot; base += DEP_T * DEP_K) {
    float4 rv[DEP_K];
    int32_t ob[DEP_K];  // block offset of the record, packed; 0 = none
#pragma unroll
    for (int j = 0; j < DEP_K; j++) {
      const int32_t t = base + wv * 64 * DEP_K + l * DEP_K + j;
      ob[j] = 0;
      if (t < ntot) {
        int32_t k = 0, bxo = 0, byo = 0, bzo = 0;
        if (t < nown) {
          int r = 0;
          while (r + 1 < TB * TB && opre[r + 1] <= t) r++;
          k = ostart[r] + (t - opre[r]);
          byo = r % TB;
          bzo = r / TB;
          for (int q = 1; q < TB; q++)
            if (t >= obnd[r][q]) bxo = q;
        } else {
          const int32_t th = t - nown;
          int lo = 0, hi = NS;  // last s with hpre[s] <= th
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (hpre[mid] <= th) lo = mid; else hi = mid;
          }
          k = hstart[lo] + (th - hpre[lo]);
          bxo = hoff[lo][0];
          byo = hoff[lo][1];
          bzo = hoff[lo][2];
        }
        ob[j] = ((bzo + 2) << 8) | ((byo + 2) << 4) | (bxo + 2);
        rv[j] = rec[k];
      }
    }
#pragma unroll
    for (int j = 0; j < DEP_K; j++) {
      if (ob[j] == 0) break;
      const int32_t bxo = (ob[j] & 15) - 2, byo = ((ob[j] >> 4) & 15) - 2,
                    bzo = (ob[j] >> 8) - 2;
      // lower node (tile-relative) and exact offsets
      const float4 r = rv[j];
      const float qx = (float)R * r.x, qy = (float)R * r.y, qz = (float)R * r.z;
      const float fx0 = floorf(qx), fy0 = floorf(qy), fz0 = floorf(qz);
      const int32_t ix = 8 * R * bxo + (int32_t)fx0;
      const int32_t iy = 8 * R * byo + (int32_t)fy0;
      const int32_t iz = 8 * R * bzo + (int32_t)fz0;
      const float fx = qx - fx0, fy = qy - fy0, fz = qz - fz0;
      const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy},
                  wz[2] = {1.0f - fz, fz};
      const int32_t id = ((iz + 16) * 64 + (iy + 16)) * 64 + (ix + 16);
      if (id != rid) {
        // the next cell along x: its lower x plane is the current node's
        // upper one, whose sums stay in registers
        const bool carry = rid >= 0 && id == rid + 1;
        flush(0);
        if (!carry) flush(1);
#pragma unroll
        for (int c = 0; c < 8; c += 2) {
          ra[c] = carry ? ra[c + 1] : 0ull;
          ra[c + 1] = 0ull;
        }
        rid = id;
        rix = ix;
        riy = iy;
        riz = iz;
      }
#pragma unroll
      for (int c = 0; c < 8; c++)
        ra[c] += __float2ull_rn(wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2] *
                                FIX_SCALE);
    }
  }
  flush(0);
  flush(1);
  __syncthreads();
  const int32_t gx0 = bx0 * 8 * R, gy0 = by0 * 8 * R, gz0 = bz0 * 8 * R;
  for (int i = threadIdx.x; i < NN * NN * NN; i += DEP_T) {
    const int ax = i % NN, ay = (i / NN) % NN, az = i / (NN * NN);
    const int32_t gx = (gx0 + ax + shift) & (ng - 1);
    const int32_t gy = (gy0 + ay + shift) & (ng - 1);
    const int32_t gz = (gz0 + az + shift) & (ng - 1);
    grid[((int64_t)gz * ng + gy) * ng + gx] =
        (float)((double)acc[i] * FIX_INV * norm);
  }
}

// CIC^-1 of the block's own particles from its (8R+1)^3 nodes.
template <int R>
__global__ void __launch_bounds__(256)
k_cic_interp(const int32_t *__restrict__ cell_start,
             const float4 *__restrict__ rec, int32_t nx, int32_t ng,
             int32_t shift, const float *__restrict__ grid,
             float *__restrict__ rho_sorted, float *__restrict__ rho,
             const int32_t *__restrict__ lab_sorted, int64_t nhalo,
             unsigned long long *__restrict__ seed_key) {
  constexpr int NN = 8 * R + 1;
  __shared__ float g[NN * NN * NN];
  const int32_t nb = nx >> 3;
  const int32_t blk = blockIdx.x;
  const int bx = blk % nb, by = (blk / nb) % nb, bz = blk / (nb * nb);
  const int32_t gx0 = bx * 8 * R, gy0 = by * 8 * R, gz0 = bz * 8 * R;
  const bool seeds = seed_key && lab_sorted;
  // The block's record range first; then the first round's record and label
  // are requested before the nodes are staged, and each round requests the
  // next round's before its own are used, so that the loads of a block are
  // in flight together rather than one after the other.
  const int32_t s0 = cell_start[blk * BS3], s1 = cell_start[blk * BS3 + BS3];
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
  int32_t h = -1;
  if (s0 + (int32_t)threadIdx.x < s1) {
    r = rec[s0 + threadIdx.x];
    if (seeds) h = lab_sorted[s0 + threadIdx.x];
  }
  for (int i = threadIdx.x; i < NN * NN * NN; i += 256) {
    const int ax = i % NN, ay = (i / NN) % NN, az = i / (NN * NN);
    const int32_t gx = (gx0 + ax + shift) & (ng - 1);
    const int32_t gy = (gy0 + ay + shift) & (ng - 1);
    const int32_t gz = (gz0 + az + shift) & (ng - 1);
    g[i] = grid[((int64_t)gz * ng + gy) * ng + gx];
  }
  __syncthreads();
  // uniform trip count so that every lane of a wave takes part in the
  // per-label reduction of the seed keys
  for (int32_t base = s0; base < s1; base += 256) {
    const int32_t k = base + threadIdx.x;
    const bool ok = k < s1;
    float4 rn = make_float4(0.f, 0.f, 0.f, 0.f);
    int32_t hn = -1;
    if (k + 256 < s1) {
      rn = rec[k + 256];
      if (seeds) hn = lab_sorted[k + 256];
    }
    float rf = 0.0f;
    int32_t idx = 0;
    if (ok) {
      const float qx = (float)R * r.x, qy = (float)R * r.y, qz = (float)R * r.z;
      const float fx0 = floorf(qx), fy0 = floorf(qy), fz0 = floorf(qz);
      const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
      const float fx = qx - fx0, fy = qy - fy0, fz = qz - fz0;
      const double wx[2] = {1.0 - (double)fx, (double)fx};
      const double wy[2] = {1.0 - (double)fy, (double)fy};
      const double wz[2] = {1.0 - (double)fz, (double)fz};
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int ax = ix + (c & 1), ay = iy + ((c >> 1) & 1),
                  az = iz + (c >> 2);
        acc += (double)g[(az * NN + ay) * NN + ax] * wx[c & 1] *
               wy[(c >> 1) & 1] * wz[c >> 2];
      }
      rf = (float)acc;
      idx = __float_as_int(r.w);
      if (rho_sorted) rho_sorted[k] = rf;
      if (rho) rho[idx] = rf;
    }
    if (seeds) {
      const bool mem = ok && h >= 0 && h < nhalo;
      const unsigned long long key =
          mem ? (((unsigned long long)__float_as_uint(rf) << 32) |
                 (unsigned long long)(0xFFFFFFFFu - (uint32_t)idx))
              : 0ull;
      // one atomic per distinct label of the wave (members of a halo are
      // contiguous in cell order)
      uint64_t todo = __ballot(mem);
      while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int32_t hl = __shfl(h, leader, 64);
        const bool mine = mem && h == hl;
        unsigned long long v = mine ? key : 0ull;
        for (int o = 32; o > 0; o >>= 1) {
          const unsigned long long w = __shfl_xor(v, o, 64);
          v = w > v ? w : v;
        }
        if ((int)(threadIdx.x & 63) == leader) atomicMax(&seed_key[hl], v);
        todo &= ~__ballot(mine);
      }
    }
    r = rn;
    h = hn;
  }
}

static int check_grid_args(int32_t nx, int32_t ng, int32_t shift) {
  if (nx < 8 || nx > 1024 || (nx & (nx - 1)))
    return psod_set_error(PSOD_E_ARG, "nx=%d must be a power of two in [8,1024]",
                          nx);
  if (ng != nx && ng != 2 * nx)
    return psod_set_error(PSOD_E_ARG, "ng=%d must be nx or 2*nx (nx=%d)", ng,
                          nx);
  if (shift < 0 || shift >= ng)
    return psod_set_error(PSOD_E_ARG, "shift=%d outside [0, ng)", shift);
  return PSOD_OK;
}

// Tiles of a deposit launch and the length of its tile list (ints): the
// listed flags of every tile, the dense-tile count and the dense-tile list.
static int32_t deposit_tiles(int32_t nx, int32_t ng, int32_t *TB,
                             int32_t *ntile, int32_t *capd) {
  const int32_t nb = nx / 8;
  *TB = (ng == nx && nb % 2 == 0) ? 2 : 1;
  *ntile = (nb / *TB) * (nb / *TB) * (nb / *TB);
  *capd = std::min<int32_t>(*ntile, DEP_CAPD);
  return *ntile + *capd + 64;
}

// Queues the deposit on the context's stream with `listed` as its tile list
// (deposit_tiles() ints); returns without waiting.
static int deposit_queue(psod_ctx *ctx, int64_t npart, int32_t nx,
                         const int32_t *d_cell_start, const float *d_rec,
                         int32_t ng, int32_t shift, float *d_grid,
                         int32_t *listed) {
  int32_t TB, ntile, capd;
  (void)deposit_tiles(nx, ng, &TB, &ntile, &capd);
  const int32_t nb = nx / 8;
  const double norm = (double)ng * (double)ng * (double)ng / (double)npart;
  int32_t *dcount = listed + ntile, *dlist = dcount + 32;
  hipStream_t st = ctx->stream;
  const unsigned nblk = (unsigned)(ntile + capd);
  if (capd > 0) {
    PSOD_TRY(psod_fill_i32(ctx, dcount, 1, 0));
    k_tile_dense<<<grid_for(ntile, 256), 256, 0, st>>>(
        d_cell_start, nb, TB, ntile, capd,
        psod_tunables_for(ctx, nx).dense_tile, listed, dcount, dlist);
    PSOD_LAUNCH_CHECK();
  }
  if (ng == nx && TB == 2)
    k_cic_deposit<1, 2><<<nblk, DEP_T, 0, st>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, norm, capd, dcount,
        dlist, listed, d_grid);
  else if (ng == nx)
    k_cic_deposit<1, 1><<<nblk, DEP_T, 0, st>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, norm, capd, dcount,
        dlist, listed, d_grid);
  else
    k_cic_deposit<2, 1><<<nblk, DEP_T, 0, st>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, norm, capd, dcount,
        dlist, listed, d_grid);
  PSOD_LAUNCH_CHECK();
  return PSOD_OK;
}

extern "C" int psod_cic_deposit(psod_ctx *ctx, int64_t npart, int32_t nx,
                                const int32_t *d_cell_start,
                                const float *d_rec, int32_t ng, int32_t shift,
                                float *d_grid) {
  PSOD_TRY(psod_use_device(ctx));
  PSOD_TRY(check_grid_args(nx, ng, shift));
  if (npart <= 0 || !d_cell_start || !d_rec || !d_grid)
    return psod_set_error(PSOD_E_ARG, "psod_cic_deposit: bad arguments");
  int32_t TB, ntile, capd;
  const int32_t nlist = deposit_tiles(nx, ng, &TB, &ntile, &capd);
  int32_t *listed = (int32_t *)psod_scratch(ctx, S_TILE, (size_t)nlist * 4);
  if (!listed) return PSOD_E_NOMEM;
  PSOD_TRY(deposit_queue(ctx, npart, nx, d_cell_start, d_rec, ng, shift,
                         d_grid, listed));
  return psod_finish(ctx);
}

// ---------------------------------------------------------------------------
// The deposit for callers that cannot allocate device memory.  Its tile list
// lives in `scratch`, a fixed array of every device the library runs on,
// large enough for the largest valid grid (nx = 1024, ng = 2048: 2^21 tiles).
// In host code the array's name is not a device address: psod_create
// resolves the array of its device once (hipGetSymbolAddress, with that
// device current) and the entry point passes only that address on.
// ---------------------------------------------------------------------------
constexpr int32_t STATIC_TILE_INTS = (1 << 21) + DEP_CAPD + 64;
constexpr int PSOD_MAX_DEVICES = 64;
static __device__ __align__(256) int32_t scratch[STATIC_TILE_INTS];
static int32_t *g_scratch[PSOD_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Called by psod_create with `device` current.
int psod_static_scratch_init(int device) {
  if (device < 0 || device >= PSOD_MAX_DEVICES)
    return psod_set_error(PSOD_E_ARG, "device %d beyond the %d the library "
                          "supports", device, PSOD_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return PSOD_OK;
  void *p = nullptr;
  PSOD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p)
    return psod_set_error(PSOD_E_DEVICE, "no address for the fixed tile "
                          "list of device %d", device);
  g_scratch[device] = (int32_t *)p;
  return PSOD_OK;
}

extern "C" int psod_cic_deposit_noalloc(psod_ctx *ctx, int64_t npart,
                                        int32_t nx,
                                        const int32_t *d_cell_start,
                                        const float *d_rec, int32_t ng,
                                        int32_t shift, float *d_grid) {
  PSOD_TRY(psod_use_device(ctx));
  PSOD_TRY(check_grid_args(nx, ng, shift));
  if (npart <= 0 || !d_cell_start || !d_rec || !d_grid)
    return psod_set_error(PSOD_E_ARG, "psod_cic_deposit_noalloc: bad "
                          "arguments");
  int32_t TB, ntile, capd;
  const int32_t nlist = deposit_tiles(nx, ng, &TB, &ntile, &capd);
  if (nlist > STATIC_TILE_INTS)
    return psod_set_error(PSOD_E_CAPACITY, "psod_cic_deposit_noalloc: %d "
                          "tile-list entries exceed the fixed %d", nlist,
                          STATIC_TILE_INTS);
  int32_t *listed = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < PSOD_MAX_DEVICES)
      listed = g_scratch[ctx->device];
  }
  if (!listed)
    return psod_set_error(PSOD_E_DEVICE, "psod_cic_deposit_noalloc: the "
                          "fixed tile list of device %d is not set up",
                          ctx->device);
  return deposit_queue(ctx, npart, nx, d_cell_start, d_rec, ng, shift, d_grid,
                       listed);
}

extern "C" int psod_cic_interp(psod_ctx *ctx, int64_t npart, int32_t nx,
                               const int32_t *d_cell_start, const float *d_rec,
                               int32_t ng, int32_t shift, const float *d_grid,
                               float *d_rho_sorted, float *d_rho,
                               const int32_t *d_lab_sorted, int64_t nhalo,
                               uint64_t *d_seed_key) {
  PSOD_TRY(psod_use_device(ctx));
  PSOD_TRY(check_grid_args(nx, ng, shift));
  if (npart <= 0 || !d_cell_start || !d_rec || !d_grid)
    return psod_set_error(PSOD_E_ARG, "psod_cic_interp: bad arguments");
  const int32_t nb = nx / 8;
  const unsigned nblk = (unsigned)nb * nb * nb;
  // the keys are maxima over the members: clear them on this stream, so
  // that no fill queued elsewhere can race the atomics below
  if (d_seed_key && d_lab_sorted && nhalo > 0)
    PSOD_TRY(psod_fill_u64(ctx, d_seed_key, nhalo, 0));
  if (ng == nx)
    k_cic_interp<1><<<nblk, 256, 0, ctx->stream>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, d_grid,
        d_rho_sorted, d_rho, d_lab_sorted, nhalo,
        (unsigned long long *)d_seed_key);
  else
    k_cic_interp<2><<<nblk, 256, 0, ctx->stream>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, d_grid,
        d_rho_sorted, d_rho, d_lab_sorted, nhalo,
        (unsigned long long *)d_seed_key);
  PSOD_LAUNCH_CHECK();
  return psod_finish(ctx);
}

__global__ void k_seed_index(const unsigned long long *__restrict__ key,
                             int64_t n, int32_t *__restrict__ seed) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= n) return;
  const unsigned long long k = key[h];
  seed[h] = (k == 0ull) ? -1
                        : (int32_t)(0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFull));
}

extern "C" int psod_seed_index(psod_ctx *ctx, int64_t nhalo,
                               const uint64_t *d_seed_key, int32_t *d_seed) {
  PSOD_TRY(psod_use_device(ctx));
  if (nhalo <= 0) return PSOD_OK;
  k_seed_index<<<grid_for(nhalo, 256), 256, 0, ctx->stream>>>(
      (const unsigned long long *)d_seed_key, nhalo, d_seed);
  PSOD_LAUNCH_CHECK();
  return psod_finish(ctx);
}

// Members of a label come in runs (a halo's members are neighbours in cell
// order, and catalogues list each group's members together), so each wave
// adds its 64 labels with one atomic per distinct label of the wave; after
// LC_RUNS distinct labels the remaining lanes add their own.  One atomic per
// particle ran at 85 GB/s on 1024^3 particles.  Each wave takes a contiguous
// range of the labels and holds the count of its current label across its
// rounds, adding it when another label's run starts: a halo of 5e7 members
// costs one atomic per wave that meets it, not one per 64 members (all on
// one counter: 9.3 ms for the zoom catalogue's central halo).
constexpr int LC_RUNS = 4;
__global__ void __launch_bounds__(256)
k_label_counts(const int32_t *__restrict__ lab, int64_t n, int64_t nhalo,
               int32_t *__restrict__ cnt) {
  const int l = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t per = ((n + nw - 1) / nw + 63) & ~(int64_t)63;
  const int64_t lo = w * per, hi = min(n, lo + per);
  int32_t hold = -1, hcnt = 0;  // wave-uniform: the label held and its count
  // uniform trip count per wave: every lane takes part in the ballots
  for (int64_t b = lo; b < hi; b += 64) {
    const int64_t i = b + l;
    const int32_t h = i < hi ? lab[i] : -1;
    const bool mem = h >= 0 && h < nhalo;
    uint64_t todo = __ballot(mem);
    for (int r = 0; todo; r++) {
      if (r == LC_RUNS) {
        if ((todo >> l) & 1) atomicAdd(&cnt[h], 1);
        break;
      }
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int32_t hl = __shfl(h, leader, 64);
      const uint64_t same = __ballot(mem && h == hl);
      const int32_t c = (int32_t)__popcll(same);
      if (hl == hold) {
        hcnt += c;
      } else {
        if (l == 0 && hcnt > 0) atomicAdd(&cnt[hold], hcnt);
        hold = hl;
        hcnt = c;
      }
      todo &= ~same;
    }
  }
  if (l == 0 && hcnt > 0) atomicAdd(&cnt[hold], hcnt);
}

extern "C" int psod_label_counts(psod_ctx *ctx, int64_t npart,
                                 const int32_t *d_labels, int64_t nhalo,
                                 int32_t *d_nmem) {
  PSOD_TRY(psod_use_device(ctx));
  PSOD_TRY(psod_fill_i32(ctx, d_nmem, nhalo, 0));
  if (npart > 0) {
    k_label_counts<<<grid_for(npart, 256, (int64_t)ctx->num_cus * 64), 256, 0,
                     ctx->stream>>>(d_labels, npart, nhalo, d_nmem);
    PSOD_LAUNCH_CHECK();
  }
  return psod_finish(ctx);
}

// ---------------------------------------------------------------------------
// a4: density order.  Ascending sort of (0xFFFFFFFF - float bits of rho)
// with the particle index as a stable payload = decreasing rho, ties by
// index; npart_new = #(rho > threshold) + 1 (main file:1029-1032).
// ---------------------------------------------------------------------------
__global__ void k_density_keys(const float *__restrict__ rho, int64_t n,
                               float thr, uint64_t *__restrict__ keys,
                               uint32_t *__restrict__ vals,
                               unsigned long long *__restrict__ nabove) {
  unsigned long long cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float r = rho[i];
    keys[i] = (uint64_t)(0xFFFFFFFFu - __float_as_uint(fmaxf(r, 0.0f)));
    vals[i] = (uint32_t)i;
    cnt += (r > thr) ? 1 : 0;
  }
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(nabove, cnt);
}

__global__ void k_isort_out(const uint32_t *__restrict__ vals, int64_t n,
                            int32_t *__restrict__ isort) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) isort[i] = (int32_t)vals[i] + 1;
}

extern "C" int psod_density_order(psod_ctx *ctx, int64_t npart,
                                  const float *d_rho, double threshold,
                                  int32_t *d_isort, int64_t *npart_new) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || !d_rho || !d_isort || !npart_new)
    return psod_set_error(PSOD_E_ARG, "psod_density_order: bad arguments");
  uint64_t *keys = (uint64_t *)psod_scratch(ctx, S_KEYS, (size_t)npart * 8);
  uint32_t *vals = (uint32_t *)psod_scratch(ctx, S_VALS, (size_t)npart * 4);
  unsigned long long *cnt =
      (unsigned long long *)psod_scratch(ctx, S_MISC, 256);
  if (!keys || !vals || !cnt) return PSOD_E_NOMEM;
  PSOD_TRY(psod_fill_u64(ctx, (uint64_t *)cnt, 1, 0));
  // float threshold: rho > thr in float equals rho > threshold for the
  // thresholds used (exactly representable, e.g. 100)
  k_density_keys<<<grid_for(npart, 256, (int64_t)ctx->num_cus * 32), 256, 0,
                   ctx->stream>>>(d_rho, npart, (float)threshold, keys, vals,
                                  cnt);
  PSOD_LAUNCH_CHECK();
  PSOD_TRY(psod_radix_sort_u64(ctx, npart, keys, vals, 32));
  k_isort_out<<<grid_for(npart, 256), 256, 0, ctx->stream>>>(vals, npart,
                                                             d_isort);
  PSOD_LAUNCH_CHECK();
  unsigned long long h = 0;
  PSOD_HIP(hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  int64_t nn = (int64_t)h + 1;
  if (nn > npart) nn = npart;
  *npart_new = nn;
  return PSOD_OK;
}
