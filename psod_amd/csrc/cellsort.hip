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

__global__ void __launch_bounds__(256)
k_cell_hist(const float *__restrict__ pos, int64_t n, int32_t nx,
            int32_t *__restrict__ count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t cx, cy, cz;
    scaled_wrap(pos[3 * i + 0], nx, &cx);
    scaled_wrap(pos[3 * i + 1], nx, &cy);
    scaled_wrap(pos[3 * i + 2], nx, &cz);
    atomicAdd(&count[cell_key(nx, cx, cy, cz)], 1);
  }
}

__global__ void __launch_bounds__(256)
k_cell_scatter(const float *__restrict__ pos, const int32_t *__restrict__ lab,
               int64_t n, int32_t nx, const int32_t *__restrict__ cell_start,
               int32_t *__restrict__ fill, int32_t *__restrict__ perm,
               float4 *__restrict__ rec, int32_t *__restrict__ lab_sorted) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t cx, cy, cz;
    const float xs = scaled_wrap(pos[3 * i + 0], nx, &cx);
    const float ys = scaled_wrap(pos[3 * i + 1], nx, &cy);
    const float zs = scaled_wrap(pos[3 * i + 2], nx, &cz);
    const int32_t key = cell_key(nx, cx, cy, cz);
    const int32_t slot = cell_start[key] + atomicAdd(&fill[key], 1);
    if (perm) perm[slot] = (int32_t)i;
    if (rec) {
      float4 r;
      r.x = xs - (float)(cx & ~7);
      r.y = ys - (float)(cy & ~7);
      r.z = zs - (float)(cz & ~7);
      r.w = __int_as_float((int32_t)i);
      rec[slot] = r;
    }
    if (lab_sorted) lab_sorted[slot] = lab ? lab[i] : -1;
  }
}

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);

extern "C" int psod_cell_sort(psod_ctx *ctx, int64_t npart, const float *d_pos,
                              const int32_t *d_labels, int32_t nx,
                              int32_t *d_cell_start, int32_t *d_perm,
                              float *d_rec, int32_t *d_lab_sorted) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart < 0 || npart >= (1ll << 31) || !d_pos || !d_cell_start)
    return psod_set_error(PSOD_E_ARG, "psod_cell_sort: bad arguments");
  if (nx < 8 || nx > 1024 || (nx & (nx - 1)))
    return psod_set_error(PSOD_E_ARG,
                          "psod_cell_sort: nx=%d must be a power of two in "
                          "[8, 1024]", nx);
  const int64_t ncell = (int64_t)nx * nx * nx;
  int32_t *count = (int32_t *)psod_scratch(ctx, S_COUNT,
                                           (size_t)(ncell + 1) * 4);
  if (!count) return PSOD_E_NOMEM;
  PSOD_TRY(psod_fill_i32(ctx, count, ncell + 1, 0));
  const unsigned g = grid_for(npart, 256, (int64_t)ctx->num_cus * 64);
  if (npart > 0) {
    k_cell_hist<<<g, 256, 0, ctx->stream>>>(d_pos, npart, nx, count);
    PSOD_LAUNCH_CHECK();
  }
  PSOD_TRY(psod_scan_i32(ctx, count, d_cell_start, ncell + 1, nullptr));
  PSOD_TRY(psod_fill_i32(ctx, count, ncell, 0));
  if (npart > 0) {
    k_cell_scatter<<<g, 256, 0, ctx->stream>>>(
        d_pos, d_labels, npart, nx, d_cell_start, count, d_perm,
        (float4 *)d_rec, d_lab_sorted);
    PSOD_LAUNCH_CHECK();
  }
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}
