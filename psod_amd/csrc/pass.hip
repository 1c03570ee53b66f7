// pass.hip -- the whole FoF-seeded SOD pass behind one call.
//
// The reference's per-zone sequence (main file:659-744) is cic -> cicm1 ->
// tri_cic -> spherover over module globals.  For the FoF-seeded pass the
// same chain is: cell sort, CIC deposit, CIC^-1 with each label's densest
// member, member counts, batched SOD.  The intermediates (cell offsets,
// records, sorted labels, seed keys, seeds, counts, and the grid unless the
// caller wants it) live in the context's scratch and are reused by the next
// call.
#include "psod_internal.h"

extern "C" int psod_pass(psod_ctx *ctx, const psod_sod_params *p,
                         int64_t npart, int32_t nx, int32_t ng,
                         const float *d_pos, const float *d_vel,
                         const int64_t *d_ids, const int32_t *d_labels,
                         int64_t nhalo, float *d_grid,
                         const psod_halo_out *out) {
  PSOD_TRY(psod_use_device(ctx));
  if (!p || !out || !d_pos || !d_labels || npart <= 0 || nhalo < 0)
    return psod_set_error(PSOD_E_ARG, "psod_pass: bad arguments");
  if (nx < 8 || nx > 1024 || (nx & (nx - 1)))
    return psod_set_error(PSOD_E_ARG, "psod_pass: nx=%d must be a power of "
                          "two in [8, 1024]", nx);
  if (p->nx != nx)
    return psod_set_error(PSOD_E_ARG, "psod_pass: params nx=%d, nx=%d",
                          p->nx, nx);
  const int64_t ncell = (int64_t)nx * nx * nx;
  int32_t *cs = (int32_t *)psod_scratch(ctx, S_PASS_CS, (size_t)(ncell + 1) * 4);
  float *rec = (float *)psod_scratch(ctx, S_PASS_REC, (size_t)npart * 16);
  int32_t *labs = (int32_t *)psod_scratch(ctx, S_PASS_LAB, (size_t)npart * 4);
  // seed keys, seeds and member counts
  char *hs = (char *)psod_scratch(ctx, S_PASS_HALO,
                                  (size_t)(nhalo > 0 ? nhalo : 1) * 16 + 512);
  float *grid = d_grid;
  if (!grid)
    grid = (float *)psod_scratch(ctx, S_PASS_GRID,
                                 (size_t)ng * (size_t)ng * (size_t)ng * 4);
  if (!cs || !rec || !labs || !hs || !grid) return PSOD_E_NOMEM;
  uint64_t *keys = (uint64_t *)hs;
  int32_t *seed = (int32_t *)(hs + (size_t)(nhalo > 0 ? nhalo : 1) * 8 + 256);
  int32_t *nmem = seed + (nhalo > 0 ? nhalo : 1);
  PSOD_TRY(psod_cell_sort(ctx, npart, d_pos, d_labels, nx, cs, nullptr, rec,
                          labs));
  PSOD_TRY(psod_cic_deposit(ctx, npart, nx, cs, rec, ng, 0, grid));
  if (nhalo == 0) return PSOD_OK;
  PSOD_TRY(psod_cic_interp(ctx, npart, nx, cs, rec, ng, 0, grid, nullptr,
                           nullptr, labs, nhalo, keys));
  PSOD_TRY(psod_seed_index(ctx, nhalo, keys, seed));
  PSOD_TRY(psod_label_counts(ctx, npart, d_labels, nhalo, nmem));
  psod_catalogue cat;
  cat.npart = npart;
  cat.nx = nx;
  cat.pos = d_pos;
  cat.vel = d_vel;
  cat.ids = d_ids;
  cat.cell_start = cs;
  cat.rec = rec;
  return psod_sod_fof(ctx, p, &cat, nhalo, seed, nmem, out);
}
