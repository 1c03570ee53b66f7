// cic.hip -- a2 CIC deposit, a3 CIC^-1 (+ FoF seeds), a4 density order.
//
// Deposit (main file:1077-1265).  One workgroup owns the (8R)^3 nodes of
// one 8^3-cell block (R = ng/nx in {1, 2}).  A node receives mass from the
// cells on both sides of it, so the workgroup reads its own block's
// particles (one contiguous range of the cell-sorted records) plus the
// particles of the lower-neighbour boundary cells (92 ranges in total) and
// accumulates only the nodes it owns.  Every node is therefore written
// exactly once with a plain store: no zero-fill and no global atomics.
// Node sums are 64-bit fixed point (weights * 2^36) in LDS, so they are
// exact sums of the float32 weights and independent of particle order.
#include "psod_internal.h"

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);
int psod_fill_u64(psod_ctx *c, uint64_t *p, int64_t n, uint64_t v);

constexpr int NSEG = 92;
// weights in 2^-36 units: a node holds up to 2^28 unit weights
constexpr float FIX_SCALE = 68719476736.0f;  // 2^36
constexpr double FIX_INV = 1.0 / 68719476736.0;

// Segment s of the ranges a block's nodes depend on: its own 512 cells and
// the boundary cells of the 7 lower neighbour blocks.  Returns the local
// cell key inside the (possibly neighbour) block, the run length in cells
// and the block offset (dx, dy, dz) in {-1, 0}.
__device__ inline void seg_decode(int s, int *loc, int *len, int *dx, int *dy,
                                  int *dz) {
  *dx = *dy = *dz = 0;
  if (s == 0) { *loc = 0; *len = 512; return; }
  if (s < 65) {  // x-1 face: lx=7, every (ly, lz)
    const int t = s - 1;
    *loc = ((t >> 3) * 8 + (t & 7)) * 8 + 7; *len = 1; *dx = -1; return;
  }
  if (s < 73) {  // y-1 face: ly=7, rows of 8 cells for each lz
    *loc = ((s - 65) * 8 + 7) * 8; *len = 8; *dy = -1; return;
  }
  if (s == 73) { *loc = 7 * 64; *len = 64; *dz = -1; return; }  // z-1 face
  if (s < 82) {  // x-1, y-1 edge
    *loc = ((s - 74) * 8 + 7) * 8 + 7; *len = 1; *dx = -1; *dy = -1; return;
  }
  if (s < 90) {  // x-1, z-1 edge
    *loc = (7 * 8 + (s - 82)) * 8 + 7; *len = 1; *dx = -1; *dz = -1; return;
  }
  if (s == 90) { *loc = (7 * 8 + 7) * 8; *len = 8; *dy = -1; *dz = -1; return; }
  *loc = 511; *len = 1; *dx = -1; *dy = -1; *dz = -1;  // corner
}

struct SegTable {
  int32_t start[NSEG];   // first sorted slot
  int32_t pre[NSEG + 1]; // prefix of particle counts
  int8_t off[NSEG][3];
};

__device__ inline void build_segments(SegTable &T, const int32_t *cell_start,
                                      int32_t nb, int bx, int by, int bz) {
  const int t = threadIdx.x;
  int32_t cnt = 0;
  if (t < NSEG) {
    int loc, len, dx, dy, dz;
    seg_decode(t, &loc, &len, &dx, &dy, &dz);
    const int32_t nbk = (wrapi(bz + dz, nb) * nb + wrapi(by + dy, nb)) * nb +
                        wrapi(bx + dx, nb);
    const int32_t k0 = nbk * BS3 + loc;
    const int32_t s0 = cell_start[k0];
    cnt = cell_start[k0 + len] - s0;
    T.start[t] = s0;
    T.off[t][0] = (int8_t)dx;
    T.off[t][1] = (int8_t)dy;
    T.off[t][2] = (int8_t)dz;
  }
  // exclusive prefix over the first 128 threads (2 waves)
  __shared__ int32_t carry;
  if (t < 64) {
    int32_t inc = wave_incl_scan(cnt);
    if (t < NSEG) T.pre[t] = inc - cnt;
    if (t == 63) carry = inc;
  }
  __syncthreads();
  if (t >= 64 && t < 128) {
    int32_t inc = wave_incl_scan(cnt);
    if (t < NSEG) T.pre[t] = carry + inc - cnt;
    if (t == NSEG - 1) T.pre[NSEG] = carry + inc;
  }
  __syncthreads();
}

__device__ inline int find_seg(const SegTable &T, int32_t j) {
  int lo = 0, hi = NSEG;  // last s with pre[s] <= j
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (T.pre[mid] <= j) lo = mid; else hi = mid;
  }
  return lo;
}

template <int R>
__global__ void __launch_bounds__(256)
k_cic_deposit(const int32_t *__restrict__ cell_start,
              const float4 *__restrict__ rec, int32_t nx, int32_t ng,
              int32_t shift, double norm, float *__restrict__ grid) {
  constexpr int NN = 8 * R;  // owned nodes per axis
  __shared__ unsigned long long acc[NN * NN * NN];
  __shared__ SegTable T;
  const int32_t nb = nx >> 3;
  const int32_t blk = blockIdx.x;
  const int bx = blk % nb, by = (blk / nb) % nb, bz = blk / (nb * nb);
  for (int i = threadIdx.x; i < NN * NN * NN; i += 256) acc[i] = 0ull;
  build_segments(T, cell_start, nb, bx, by, bz);
  const int32_t total = T.pre[NSEG];
  for (int32_t j = threadIdx.x; j < total; j += 256) {
    const int s = find_seg(T, j);
    const float4 r = rec[T.start[s] + (j - T.pre[s])];
    // position in CIC cells relative to this block's first node
    const float qx = (float)R * (r.x + 8.0f * (float)T.off[s][0]);
    const float qy = (float)R * (r.y + 8.0f * (float)T.off[s][1]);
    const float qz = (float)R * (r.z + 8.0f * (float)T.off[s][2]);
    const float fx0 = floorf(qx), fy0 = floorf(qy), fz0 = floorf(qz);
    const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    const float fx = qx - fx0, fy = qy - fy0, fz = qz - fz0;
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy},
                wz[2] = {1.0f - fz, fz};
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int ax = ix + (c & 1), ay = iy + ((c >> 1) & 1),
                az = iz + (c >> 2);
      if ((unsigned)ax < (unsigned)NN && (unsigned)ay < (unsigned)NN &&
          (unsigned)az < (unsigned)NN) {
        const float w = wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2];
        const unsigned long long q = __float2ull_rn(w * FIX_SCALE);
        atomicAdd(&acc[(az * NN + ay) * NN + ax], q);
      }
    }
  }
  __syncthreads();
  const int32_t gx0 = bx * NN, gy0 = by * NN, gz0 = bz * NN;
  for (int i = threadIdx.x; i < NN * NN * NN; i += 256) {
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
  for (int i = threadIdx.x; i < NN * NN * NN; i += 256) {
    const int ax = i % NN, ay = (i / NN) % NN, az = i / (NN * NN);
    const int32_t gx = (gx0 + ax + shift) & (ng - 1);
    const int32_t gy = (gy0 + ay + shift) & (ng - 1);
    const int32_t gz = (gz0 + az + shift) & (ng - 1);
    g[i] = grid[((int64_t)gz * ng + gy) * ng + gx];
  }
  __syncthreads();
  const int32_t s0 = cell_start[blk * BS3], s1 = cell_start[blk * BS3 + BS3];
  // uniform trip count so that every lane of a wave takes part in the
  // per-label reduction of the seed keys
  for (int32_t base = s0; base < s1; base += 256) {
    const int32_t k = base + threadIdx.x;
    const bool ok = k < s1;
    float rf = 0.0f;
    int32_t idx = 0;
    if (ok) {
      const float4 r = rec[k];
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
    if (seed_key && lab_sorted) {
      const int32_t h = ok ? lab_sorted[k] : -1;
      const bool mem = h >= 0 && h < nhalo;
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

extern "C" int psod_cic_deposit(psod_ctx *ctx, int64_t npart, int32_t nx,
                                const int32_t *d_cell_start,
                                const float *d_rec, int32_t ng, int32_t shift,
                                float *d_grid) {
  PSOD_TRY(psod_use_device(ctx));
  PSOD_TRY(check_grid_args(nx, ng, shift));
  if (npart <= 0 || !d_cell_start || !d_rec || !d_grid)
    return psod_set_error(PSOD_E_ARG, "psod_cic_deposit: bad arguments");
  const int32_t nb = nx / 8;
  const unsigned nblk = (unsigned)nb * nb * nb;
  const double norm = (double)ng * (double)ng * (double)ng / (double)npart;
  if (ng == nx)
    k_cic_deposit<1><<<nblk, 256, 0, ctx->stream>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, norm, d_grid);
  else
    k_cic_deposit<2><<<nblk, 256, 0, ctx->stream>>>(
        d_cell_start, (const float4 *)d_rec, nx, ng, shift, norm, d_grid);
  PSOD_LAUNCH_CHECK();
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
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
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
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
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}

__global__ void k_label_counts(const int32_t *__restrict__ lab, int64_t n,
                               int64_t nhalo, int32_t *__restrict__ cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t h = lab[i];
    if (h >= 0 && h < nhalo) atomicAdd(&cnt[h], 1);
  }
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
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
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
