// host_api.hip -- host-pointer entry points that replace the reference
// subroutines cic, cicm1 and tri_cic for one zone covering the box.
//
// The reference's cube and rho are real(8) (main file:22-23), so these
// entry points compute in FP64 with the reference's expressions and
// operation order: the cell index and offset of main file:1154-1164, the
// weight products m*dex*dey*dez of :1194-1213 evaluated left to right
// without fused operations, the division by dx^3 (:1263), and the
// interpolation sum of :1318-1385.  The node sums are FP64 atomic adds, so
// they differ from the reference's particle-order sums only by the
// rounding of the additions (a few units of 2^-53 relative).  The density
// order sorts pdens = 1/rho in FP64 like tri_cic (:1022-1025), ties by
// particle index.
#include <cmath>

#include "psod_internal.h"

// hipcc contracts a*b + c into one fused operation by default.  Every FP64
// expression of this file is evaluated as written, each operation rounded:
// contraction is off here, and the operations are plain operators (HIP's
// __dmul_rn and __dadd_rn are plain operators defined in a header, outside
// this pragma, so they could still be fused once inlined).
#pragma clang fp contract(off)

int psod_fill_u64(psod_ctx *c, uint64_t *p, int64_t n, uint64_t v);

// cell index, its periodic neighbour and the offset of one coordinate
// (main file:1154-1190): u = (x - xxmin)/dx, ix = int(u) floored for
// negative u, f = u - ix, both indices wrapped into [0, ng).  The reference
// wraps once, which covers its zone layouts; the remainder here gives the
// same index for those and keeps any other input inside the grid.
__device__ inline void cic_index64(float x, double xxmin, double dx,
                                   int32_t ng, int32_t *i0, int32_t *i1,
                                   double *f) {
  const double u = ((double)x - xxmin) / dx;
  int32_t ix = (int32_t)u;
  if (u < 0.0) ix = ix - 1;
  *f = u - (double)ix;
  ix %= ng;
  if (ix < 0) ix += ng;
  *i0 = ix;
  *i1 = ix + 1 == ng ? 0 : ix + 1;
}

struct Corners64 {
  int32_t i[2], j[2], k[2];
  double dd[3], de[3];
};

__device__ inline Corners64 corners64(const float *pos, int64_t p,
                                      double xxmin, double dx, int32_t ng) {
  Corners64 c;
  cic_index64(pos[3 * p + 0], xxmin, dx, ng, &c.i[0], &c.i[1], &c.dd[0]);
  cic_index64(pos[3 * p + 1], xxmin, dx, ng, &c.j[0], &c.j[1], &c.dd[1]);
  cic_index64(pos[3 * p + 2], xxmin, dx, ng, &c.k[0], &c.k[1], &c.dd[2]);
  for (int a = 0; a < 3; a++) c.de[a] = 1.0 - c.dd[a];
  return c;
}

// a * b * c * d evaluated left to right, each product rounded
__device__ inline double prod4(double a, double b, double c, double d) {
  return a * b * c * d;
}

// one particle per thread, eight FP64 atomic adds in the reference's corner
// order (main file:1194-1213)
__global__ void k_cic64_deposit(const float *__restrict__ pos, int64_t n,
                                double m, double xxmin, double dx, int32_t ng,
                                double *__restrict__ grid) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const Corners64 c = corners64(pos, p, xxmin, dx, ng);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int a = q & 1, b = (q >> 1) & 1, z = q >> 2;
      const double w = prod4(m, a ? c.dd[0] : c.de[0], b ? c.dd[1] : c.de[1],
                             z ? c.dd[2] : c.de[2]);
      atomicAdd(&grid[((int64_t)c.k[z] * ng + c.j[b]) * ng + c.i[a]], w);
    }
  }
}

// cube = cube/(dx*dy*dz) (main file:1263)
__global__ void k_cic64_scale(double *__restrict__ grid, int64_t n,
                              double vol) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    grid[i] = grid[i] / vol;
}

// rho = sum of the eight nodes times their weights, in the reference's order
// (main file:1318-1385)
__global__ void k_cic64_interp(const float *__restrict__ pos, int64_t n,
                               double xxmin, double dx, int32_t ng,
                               const double *__restrict__ grid,
                               double *__restrict__ rho) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const Corners64 c = corners64(pos, p, xxmin, dx, ng);
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int a = q & 1, b = (q >> 1) & 1, z = q >> 2;
      const double g = grid[((int64_t)c.k[z] * ng + c.j[b]) * ng + c.i[a]];
      r = r + prod4(g, a ? c.dd[0] : c.de[0], b ? c.dd[1] : c.de[1],
                    z ? c.dd[2] : c.de[2]);
    }
    rho[p] = r;
  }
}

static int check_ng(int32_t ng) {
  if (ng < 2 || ng > 4096)
    return psod_set_error(PSOD_E_ARG, "ng=%d outside [2, 4096]", ng);
  return PSOD_OK;
}

extern "C" int psod_host_cic(psod_ctx *ctx, int64_t npart, const float *pos,
                             const float *mass, double xxmin, int32_t ng,
                             double *grid) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || npart >= (1ll << 31) || !pos || !grid)
    return psod_set_error(PSOD_E_ARG, "psod_host_cic: bad arguments");
  PSOD_TRY(check_ng(ng));
  if (mass)
    for (int64_t i = 1; i < npart; i++)
      if (mass[i] != mass[0])
        return psod_set_error(PSOD_E_ARG,
                              "psod_host_cic: particle masses differ (the "
                              "zoom variant's variable masses are not "
                              "supported)");
  // mpf = mpf/nparttotal in real(4) (main file:662) unless given
  const double m = (double)(mass ? mass[0] : (float)(1.0 / (double)npart));
  const double dx = 1.0 / (double)ng;
  const int64_t ng3 = (int64_t)ng * ng * ng;
  float *dpos = (float *)psod_scratch(ctx, S_HOST_A, (size_t)npart * 12);
  double *dgrid = (double *)psod_scratch(ctx, S_HOST_E, (size_t)ng3 * 8);
  if (!dpos || !dgrid) return PSOD_E_NOMEM;
  hipStream_t st = ctx->stream;
  PSOD_HIP(hipMemcpyAsync(dpos, pos, (size_t)npart * 12,
                          hipMemcpyHostToDevice, st));
  PSOD_TRY(psod_fill_u64(ctx, (uint64_t *)dgrid, ng3, 0));
  k_cic64_deposit<<<grid_for(npart, 256, (int64_t)ctx->num_cus * 64), 256, 0,
                    st>>>(dpos, npart, m, xxmin, dx, ng, dgrid);
  PSOD_LAUNCH_CHECK();
  k_cic64_scale<<<grid_for(ng3, 256, (int64_t)ctx->num_cus * 64), 256, 0,
                  st>>>(dgrid, ng3, dx * dx * dx);
  PSOD_LAUNCH_CHECK();
  PSOD_HIP(hipMemcpyAsync(grid, dgrid, (size_t)ng3 * 8, hipMemcpyDeviceToHost,
                          st));
  PSOD_HIP(hipStreamSynchronize(st));
  return PSOD_OK;
}

extern "C" int psod_host_cicm1(psod_ctx *ctx, int64_t npart, const float *pos,
                               double xxmin, int32_t ng, const double *grid,
                               double *rho) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || npart >= (1ll << 31) || !pos || !grid || !rho)
    return psod_set_error(PSOD_E_ARG, "psod_host_cicm1: bad arguments");
  PSOD_TRY(check_ng(ng));
  const double dx = 1.0 / (double)ng;
  const int64_t ng3 = (int64_t)ng * ng * ng;
  float *dpos = (float *)psod_scratch(ctx, S_HOST_A, (size_t)npart * 12);
  double *dgrid = (double *)psod_scratch(ctx, S_HOST_E, (size_t)ng3 * 8);
  double *drho = (double *)psod_scratch(ctx, S_HOST_F, (size_t)npart * 8);
  if (!dpos || !dgrid || !drho) return PSOD_E_NOMEM;
  hipStream_t st = ctx->stream;
  PSOD_HIP(hipMemcpyAsync(dpos, pos, (size_t)npart * 12,
                          hipMemcpyHostToDevice, st));
  PSOD_HIP(hipMemcpyAsync(dgrid, grid, (size_t)ng3 * 8, hipMemcpyHostToDevice,
                          st));
  k_cic64_interp<<<grid_for(npart, 256, (int64_t)ctx->num_cus * 64), 256, 0,
                   st>>>(dpos, npart, xxmin, dx, ng, dgrid, drho);
  PSOD_LAUNCH_CHECK();
  PSOD_HIP(hipMemcpyAsync(rho, drho, (size_t)npart * 8, hipMemcpyDeviceToHost,
                          st));
  PSOD_HIP(hipStreamSynchronize(st));
  return PSOD_OK;
}

// ---------------------------------------------------------------------------
// a4 in FP64: ascending sort of pdens = 1/rho (main file:1022-1025) as the
// bits of a positive double, particle index as the stable payload
// ---------------------------------------------------------------------------
__global__ void k_pdens_keys(const double *__restrict__ rho, int64_t n,
                             double lim, uint64_t *__restrict__ keys,
                             uint32_t *__restrict__ vals,
                             unsigned long long *__restrict__ nbelow) {
  unsigned long long cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double pd = 1.0 / rho[i];
    // rho > 0 for every particle (its own mass is deposited); a
    // non-positive or NaN value sorts last
    keys[i] = pd > 0.0 ? (uint64_t)__double_as_longlong(pd) : ~0ull;
    vals[i] = (uint32_t)i;
    cnt += (pd > 0.0 && pd < lim) ? 1 : 0;
  }
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(nbelow, cnt);
}

__global__ void k_isort1(const uint32_t *__restrict__ vals, int64_t n,
                         int32_t *__restrict__ isort) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) isort[i] = (int32_t)vals[i] + 1;
}

extern "C" int psod_density_order_f64(psod_ctx *ctx, int64_t npart,
                                      const double *d_rho, double threshold,
                                      int32_t *d_isort, int64_t *npart_new) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || npart >= (1ll << 31) || !d_rho || !d_isort ||
      !npart_new || !(threshold > 0.0))
    return psod_set_error(PSOD_E_ARG, "psod_density_order_f64: bad arguments");
  uint64_t *keys = (uint64_t *)psod_scratch(ctx, S_KEYS, (size_t)npart * 8);
  uint32_t *vals = (uint32_t *)psod_scratch(ctx, S_VALS, (size_t)npart * 4);
  unsigned long long *cnt =
      (unsigned long long *)psod_scratch(ctx, S_MISC, 256);
  if (!keys || !vals || !cnt) return PSOD_E_NOMEM;
  hipStream_t st = ctx->stream;
  PSOD_TRY(psod_fill_u64(ctx, (uint64_t *)cnt, 1, 0));
  k_pdens_keys<<<grid_for(npart, 256, (int64_t)ctx->num_cus * 32), 256, 0,
                 st>>>(d_rho, npart, 1.0 / threshold, keys, vals, cnt);
  PSOD_LAUNCH_CHECK();
  PSOD_TRY(psod_radix_sort_u64(ctx, npart, keys, vals, 64));
  k_isort1<<<grid_for(npart, 256), 256, 0, st>>>(vals, npart, d_isort);
  PSOD_LAUNCH_CHECK();
  unsigned long long h = 0;
  PSOD_HIP(hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipStreamSynchronize(st));
  // first 1-based rank whose pdens >= 1/threshold, at most npart (:1029)
  int64_t nn = (int64_t)h + 1;
  if (nn > npart) nn = npart;
  *npart_new = nn;
  return PSOD_OK;
}

extern "C" int psod_host_tri_cic(psod_ctx *ctx, int64_t npart,
                                 const double *rho, double threshold,
                                 int32_t *isort, int64_t *npart_new) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || npart >= (1ll << 31) || !rho || !isort || !npart_new)
    return psod_set_error(PSOD_E_ARG, "psod_host_tri_cic: bad arguments");
  double *drho = (double *)psod_scratch(ctx, S_HOST_F, (size_t)npart * 8);
  int32_t *dis = (int32_t *)psod_scratch(ctx, S_HOST_G, (size_t)npart * 4);
  if (!drho || !dis) return PSOD_E_NOMEM;
  PSOD_HIP(hipMemcpyAsync(drho, rho, (size_t)npart * 8, hipMemcpyHostToDevice,
                          ctx->stream));
  PSOD_TRY(psod_density_order_f64(ctx, npart, drho, threshold, dis,
                                  npart_new));
  PSOD_HIP(hipMemcpyAsync(isort, dis, (size_t)npart * 4,
                          hipMemcpyDeviceToHost, ctx->stream));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}
