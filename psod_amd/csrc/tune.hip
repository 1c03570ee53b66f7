// tune.hip -- tunables of the batched pass and the autotuner that chooses
// them per device and workload (no reference counterpart: the reference's
// knobs are command-line flags, main file:811-852, fixed per run).
//
// The autotuner prepares the FoF-seeded pass once on the caller's
// catalogue, then sweeps one tunable at a time over its candidates with the
// others at their best so far (coordinate descent), timing only the stage
// the tunable affects: the CIC deposit for dense_tile, the batched SOD for
// the others.  Every candidate gives the same results (the tunables only
// move work between code paths that compute the same values), so the
// choice is a matter of time alone.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "psod_internal.h"

int psod_pass_prepare(psod_ctx *ctx, int64_t npart, int32_t nx, int32_t ng,
                      const float *d_pos, const int32_t *d_labels,
                      int64_t nhalo, float *d_grid, PassBuffers *b);

static int32_t clamp_or(int32_t v, int32_t def, int32_t lo, int32_t hi) {
  if (v <= 0) return def;
  return v < lo ? lo : v > hi ? hi : v;
}

psod_tunables psod_tunables_for(const psod_ctx *c, int32_t nx) {
  psod_tunables t = {TUNE_SMALL_MAX, TUNE_BOUND_MIN, TUNE_DENSE_TILE,
                     TUNE_SMALL_WG, TUNE_GIANT_MIN};
  for (int i = 0; i < c->ntune_file; i++)
    if (c->tune_file[i].nx == nx) {
      const psod_tunables &f = c->tune_file[i].t;
      if (f.small_max > 0) t.small_max = f.small_max;
      if (f.bound_min > 0) t.bound_min = f.bound_min;
      if (f.dense_tile > 0) t.dense_tile = f.dense_tile;
      if (f.small_wg_per_cu > 0) t.small_wg_per_cu = f.small_wg_per_cu;
      if (f.giant_min > 0) t.giant_min = f.giant_min;
    }
  const psod_tunables &s = c->tune_set;
  if (s.small_max > 0) t.small_max = s.small_max;
  if (s.bound_min > 0) t.bound_min = s.bound_min;
  if (s.dense_tile > 0) t.dense_tile = s.dense_tile;
  if (s.small_wg_per_cu > 0) t.small_wg_per_cu = s.small_wg_per_cu;
  if (s.giant_min > 0) t.giant_min = s.giant_min;
  t.small_max = clamp_or(t.small_max, TUNE_SMALL_MAX, 64, TUNE_SMALL_MAX);
  t.bound_min = clamp_or(t.bound_min, TUNE_BOUND_MIN, 1, 1 << 30);
  t.dense_tile = clamp_or(t.dense_tile, TUNE_DENSE_TILE, 1, 1 << 30);
  t.small_wg_per_cu = clamp_or(t.small_wg_per_cu, TUNE_SMALL_WG, 1, 32);
  t.giant_min = clamp_or(t.giant_min, TUNE_GIANT_MIN, 1, 1 << 30);
  return t;
}

extern "C" int psod_set_tunables(psod_ctx *ctx, const psod_tunables *t) {
  if (!ctx || !t) return psod_set_error(PSOD_E_ARG, "psod_set_tunables: null");
  if (t->small_max < 0 || t->small_max > TUNE_SMALL_MAX || t->bound_min < 0 ||
      t->dense_tile < 0 || t->small_wg_per_cu < 0 ||
      t->small_wg_per_cu > 32 || t->giant_min < 0)
    return psod_set_error(PSOD_E_ARG, "psod_set_tunables: out of range "
                          "(small_max %d <= %d, small_wg_per_cu %d <= 32)",
                          t->small_max, TUNE_SMALL_MAX, t->small_wg_per_cu);
  ctx->tune_set = *t;
  return PSOD_OK;
}

extern "C" int psod_get_tunables(psod_ctx *ctx, int32_t nx, psod_tunables *t) {
  if (!ctx || !t) return psod_set_error(PSOD_E_ARG, "psod_get_tunables: null");
  *t = psod_tunables_for(ctx, nx);
  return PSOD_OK;
}

extern "C" int psod_load_tunables(psod_ctx *ctx, const char *path,
                                  int32_t *nloaded) {
  if (!ctx || !path)
    return psod_set_error(PSOD_E_ARG, "psod_load_tunables: null");
  if (nloaded) *nloaded = 0;
  FILE *f = fopen(path, "r");
  if (!f) return PSOD_OK;  // no file yet: nothing tuned
  char line[512];
  int n = 0;
  while (fgets(line, sizeof(line), f)) {
    char tag[32], dev[256];
    int cus, nx, a, b, cc, d, e;
    long long np, nh;
    double ms;
    if (sscanf(line, "%31s %255s %d %d %lld %lld %d %d %d %d %d %lf", tag,
               dev, &cus, &nx, &np, &nh, &a, &b, &cc, &d, &e, &ms) != 12)
      continue;
    if (strcmp(tag, "psod-tune") != 0 || strcmp(dev, ctx->device_name) != 0 ||
        cus != ctx->num_cus)
      continue;
    int slot = -1;
    for (int i = 0; i < ctx->ntune_file; i++)
      if (ctx->tune_file[i].nx == nx) slot = i;
    if (slot < 0) {
      if (ctx->ntune_file == PSOD_TUNE_ENTRIES) continue;
      slot = ctx->ntune_file++;
    }
    ctx->tune_file[slot].nx = nx;
    ctx->tune_file[slot].t = psod_tunables{a, b, cc, d, e};
    n++;
  }
  fclose(f);
  if (nloaded) *nloaded = n;
  return PSOD_OK;
}

namespace {
using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// the context's tunables for nx with `t` in force (for this call only)
struct TuneScope {
  psod_ctx *c;
  psod_tunables saved;
  TuneScope(psod_ctx *ctx, const psod_tunables &t) : c(ctx), saved(ctx->tune_set) {
    c->tune_set = t;
  }
  ~TuneScope() { c->tune_set = saved; }
};
}  // namespace

extern "C" int psod_autotune(psod_ctx *ctx, const psod_sod_params *p,
                             int64_t npart, int32_t nx, const float *d_pos,
                             const float *d_vel, const int64_t *d_ids,
                             const int32_t *d_labels, int64_t nhalo,
                             int32_t reps, const char *cache_file,
                             psod_tunables *chosen, double *stage_ms) {
  PSOD_TRY(psod_use_device(ctx));
  if (!p || !d_pos || !d_labels || npart <= 0 || nhalo <= 0 || p->nx != nx)
    return psod_set_error(PSOD_E_ARG, "psod_autotune: bad arguments");
  if (reps < 1) reps = 1;
  PassBuffers b;
  PSOD_TRY(psod_pass_prepare(ctx, npart, nx, nx, d_pos, d_labels, nhalo,
                             nullptr, &b));
  // per-halo outputs: one block of the tuning scratch
  const int64_t nh = nhalo;
  const size_t nb = (size_t)p->nbins;
  const size_t per = 4 + 8 + 4 + 4 + 16 + 32 * 4 + 4 + 8 + 4 * nb + 48 + 8;
  char *o = (char *)psod_scratch(ctx, S_TUNE, (size_t)nh * per + 4096);
  if (!o) return PSOD_E_NOMEM;
  psod_halo_out out;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = o + off;
    off += (bytes + 255) & ~(size_t)255;
    return (void *)q;
  };
  out.centre = (int32_t *)take(nh * 4);
  out.centre_id = (int64_t *)take(nh * 8);
  out.status = (int32_t *)take(nh * 4);
  out.w = (int32_t *)take(nh * 4);
  out.count = (int32_t *)take(nh * 16);
  out.radius = (double *)take(nh * 32);
  out.overdens = (double *)take(nh * 32);
  out.rint = (double *)take(nh * 32);
  out.mint = (double *)take(nh * 32);
  out.count2 = (int32_t *)take(nh * 4);
  out.overdens2 = (double *)take(nh * 8);
  out.profile = (int32_t *)take(nh * 4 * nb);
  out.com_pos = (double *)take(nh * 24);
  out.com_vel = (double *)take(nh * 24);
  out.ngather = (int32_t *)take(nh * 4);
  out.nboundary = (int32_t *)take(nh * 4);
  if (off > (size_t)nh * per + 4096)
    return psod_set_error(PSOD_E_ARG, "psod_autotune: output layout");
  psod_catalogue cat;
  cat.npart = npart;
  cat.nx = nx;
  cat.pos = d_pos;
  cat.vel = (p->flags & PSOD_F_VELOCITY) ? d_vel : nullptr;
  cat.ids = d_ids;
  cat.cell_start = b.cell_start;
  cat.rec = b.rec;
  float *grid = b.grid;

  psod_tunables best = psod_tunables_for(ctx, nx);
  auto time_sod = [&](const psod_tunables &t, double *ms) -> int {
    TuneScope scope(ctx, t);
    double m = 1e30;
    for (int r = 0; r < reps; r++) {
      PSOD_HIP(hipStreamSynchronize(ctx->stream));
      const auto t0 = Clock::now();
      PSOD_TRY(psod_sod_fof(ctx, p, &cat, nhalo, b.seed, b.nmem, &out));
      const double e = ms_since(t0);
      m = e < m ? e : m;
    }
    *ms = m;
    return PSOD_OK;
  };
  auto time_dep = [&](const psod_tunables &t, double *ms) -> int {
    TuneScope scope(ctx, t);
    double m = 1e30;
    for (int r = 0; r < reps; r++) {
      PSOD_HIP(hipStreamSynchronize(ctx->stream));
      const auto t0 = Clock::now();
      PSOD_TRY(psod_cic_deposit(ctx, npart, nx, b.cell_start, b.rec, nx, 0,
                                grid));
      const double e = ms_since(t0);
      m = e < m ? e : m;
    }
    *ms = m;
    return PSOD_OK;
  };
  // deposit: dense-tile threshold
  double best_dep;
  PSOD_TRY(time_dep(best, &best_dep));
  for (int32_t v : {4096, 8192, 16384, 32768, 65536}) {
    psod_tunables t = best;
    t.dense_tile = v;
    double ms;
    PSOD_TRY(time_dep(t, &ms));
    if (ms < best_dep) {
      best_dep = ms;
      best = t;
    }
  }
  // SOD: one-wave cap, bound threshold, one-wave workgroups per CU
  double best_sod;
  PSOD_TRY(time_sod(best, &best_sod));
  for (int k = 0; k < 4; k++) {
    static const int32_t cand[4][5] = {{256, 512, 768, 1024, 0},
                                       {512, 1024, 2048, 4096, 8192},
                                       {4, 8, 16, 0, 0},
                                       {1 << 19, 1 << 20, 1 << 21, 1 << 22,
                                        1 << 30}};
    for (int j = 0; j < 5 && cand[k][j] > 0; j++) {
      psod_tunables t = best;
      (k == 0   ? t.small_max
       : k == 1 ? t.bound_min
       : k == 2 ? t.small_wg_per_cu
                : t.giant_min) = cand[k][j];
      double ms;
      PSOD_TRY(time_sod(t, &ms));
      if (ms < best_sod) {
        best_sod = ms;
        best = t;
      }
    }
  }
  // in force for this nx from now on
  int slot = -1;
  for (int i = 0; i < ctx->ntune_file; i++)
    if (ctx->tune_file[i].nx == nx) slot = i;
  if (slot < 0 && ctx->ntune_file < PSOD_TUNE_ENTRIES) slot = ctx->ntune_file++;
  if (slot >= 0) {
    ctx->tune_file[slot].nx = nx;
    ctx->tune_file[slot].t = best;
  }
  if (cache_file) {
    FILE *f = fopen(cache_file, "a");
    if (!f)
      return psod_set_error(PSOD_E_ARG, "psod_autotune: cannot append to %s",
                            cache_file);
    fprintf(f, "psod-tune %s %d %d %lld %lld %d %d %d %d %d %.3f\n",
            ctx->device_name, ctx->num_cus, nx, (long long)npart,
            (long long)nhalo, best.small_max, best.bound_min, best.dense_tile,
            best.small_wg_per_cu, best.giant_min, best_dep + best_sod);
    fclose(f);
  }
  if (chosen) *chosen = best;
  if (stage_ms) {
    stage_ms[0] = best_dep;
    stage_ms[1] = best_sod;
  }
  return PSOD_OK;
}
