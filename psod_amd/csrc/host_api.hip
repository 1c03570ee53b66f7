// host_api.hip -- host-pointer entry points that replace the reference
// subroutines cic, cicm1 and tri_cic for one zone covering the box.
#include <cmath>

#include "psod_internal.h"

static int grid_nx(int32_t ng, int32_t *nx) {
  if (ng < 8 || ng > 2048 || (ng & (ng - 1)))
    return psod_set_error(PSOD_E_ARG, "ng=%d must be a power of two in "
                          "[8, 2048]", ng);
  *nx = ng <= 1024 ? ng : ng / 2;
  return PSOD_OK;
}

static int node_shift(double xxmin, int32_t ng, int32_t *shift) {
  const double s = -xxmin * (double)ng;
  const double r = std::nearbyint(s);
  if (std::fabs(s - r) > 1e-9)
    return psod_set_error(PSOD_E_ARG,
                          "xxmin=%g is not a whole number of cells of 1/%d",
                          xxmin, ng);
  int64_t k = (int64_t)r % ng;
  if (k < 0) k += ng;
  *shift = (int32_t)k;
  return PSOD_OK;
}

struct SortedCat {
  float *pos;
  int32_t *cs;
  int32_t *perm;
  float *rec;
};

static int upload_and_sort(psod_ctx *ctx, int64_t npart, const float *pos,
                           int32_t nx, SortedCat *sc) {
  const int64_t ncell = (int64_t)nx * nx * nx;
  sc->pos = (float *)psod_scratch(ctx, S_HOST_A, (size_t)npart * 12);
  sc->cs = (int32_t *)psod_scratch(ctx, S_HOST_B, (size_t)(ncell + 1) * 4);
  sc->perm = (int32_t *)psod_scratch(ctx, S_HOST_C, (size_t)npart * 4);
  sc->rec = (float *)psod_scratch(ctx, S_HOST_D, (size_t)npart * 16);
  if (!sc->pos || !sc->cs || !sc->perm || !sc->rec) return PSOD_E_NOMEM;
  PSOD_HIP(hipMemcpyAsync(sc->pos, pos, (size_t)npart * 12,
                          hipMemcpyHostToDevice, ctx->stream));
  return psod_cell_sort(ctx, npart, sc->pos, nullptr, nx, sc->cs, sc->perm,
                        sc->rec, nullptr);
}

extern "C" int psod_host_cic(psod_ctx *ctx, int64_t npart, const float *pos,
                             const float *mass, double xxmin, int32_t ng,
                             float *grid) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || !pos || !grid)
    return psod_set_error(PSOD_E_ARG, "psod_host_cic: bad arguments");
  if (mass)
    for (int64_t i = 1; i < npart; i++)
      if (mass[i] != mass[0])
        return psod_set_error(PSOD_E_ARG,
                              "psod_host_cic: particle masses differ (the "
                              "zoom variant's variable masses are not "
                              "supported)");
  int32_t nx, shift;
  PSOD_TRY(grid_nx(ng, &nx));
  PSOD_TRY(node_shift(xxmin, ng, &shift));
  SortedCat sc;
  PSOD_TRY(upload_and_sort(ctx, npart, pos, nx, &sc));
  const int64_t ng3 = (int64_t)ng * ng * ng;
  float *dgrid = (float *)psod_scratch(ctx, S_HOST_E, (size_t)ng3 * 4);
  if (!dgrid) return PSOD_E_NOMEM;
  PSOD_TRY(psod_cic_deposit(ctx, npart, nx, sc.cs, sc.rec, ng, shift, dgrid));
  PSOD_HIP(hipMemcpyAsync(grid, dgrid, (size_t)ng3 * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}

extern "C" int psod_host_cicm1(psod_ctx *ctx, int64_t npart, const float *pos,
                               double xxmin, int32_t ng, const float *grid,
                               float *rho) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || !pos || !grid || !rho)
    return psod_set_error(PSOD_E_ARG, "psod_host_cicm1: bad arguments");
  int32_t nx, shift;
  PSOD_TRY(grid_nx(ng, &nx));
  PSOD_TRY(node_shift(xxmin, ng, &shift));
  SortedCat sc;
  PSOD_TRY(upload_and_sort(ctx, npart, pos, nx, &sc));
  const int64_t ng3 = (int64_t)ng * ng * ng;
  float *dgrid = (float *)psod_scratch(ctx, S_HOST_E, (size_t)ng3 * 4);
  float *drho = (float *)psod_scratch(ctx, S_HOST_F, (size_t)npart * 4);
  if (!dgrid || !drho) return PSOD_E_NOMEM;
  PSOD_HIP(hipMemcpyAsync(dgrid, grid, (size_t)ng3 * 4, hipMemcpyHostToDevice,
                          ctx->stream));
  PSOD_TRY(psod_cic_interp(ctx, npart, nx, sc.cs, sc.rec, ng, shift, dgrid,
                           nullptr, drho, nullptr, 0, nullptr));
  PSOD_HIP(hipMemcpyAsync(rho, drho, (size_t)npart * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}

extern "C" int psod_host_tri_cic(psod_ctx *ctx, int64_t npart,
                                 const float *rho, double threshold,
                                 int32_t *isort, int64_t *npart_new) {
  PSOD_TRY(psod_use_device(ctx));
  if (npart <= 0 || !rho || !isort || !npart_new)
    return psod_set_error(PSOD_E_ARG, "psod_host_tri_cic: bad arguments");
  float *drho = (float *)psod_scratch(ctx, S_HOST_F, (size_t)npart * 4);
  int32_t *dis = (int32_t *)psod_scratch(ctx, S_HOST_G, (size_t)npart * 4);
  if (!drho || !dis) return PSOD_E_NOMEM;
  PSOD_HIP(hipMemcpyAsync(drho, rho, (size_t)npart * 4, hipMemcpyHostToDevice,
                          ctx->stream));
  PSOD_TRY(psod_density_order(ctx, npart, drho, threshold, dis, npart_new));
  PSOD_HIP(hipMemcpyAsync(isort, dis, (size_t)npart * 4,
                          hipMemcpyDeviceToHost, ctx->stream));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}
