// pass.hip -- the whole FoF-seeded SOD pass behind one call.
//
// The reference's per-zone sequence (main file:659-744) is cic -> cicm1 ->
// tri_cic -> spherover over module globals.  For the FoF-seeded pass the
// same chain is: cell sort, CIC deposit, CIC^-1 with each label's densest
// member, member counts, batched SOD.  The intermediates (cell offsets,
// records, sorted labels, seed keys, seeds, counts, and the grid unless the
// caller wants it) live in the context's scratch and are reused by the next
// call.
#include <algorithm>
#include <cstring>

#include "psod_internal.h"

// Everything before the SOD: cell sort, deposit, seeds and member counts,
// into context scratch (the grid into d_grid when it is not NULL).
int psod_pass_prepare(psod_ctx *ctx, int64_t npart, int32_t nx, int32_t ng,
                      const float *d_pos, const int32_t *d_labels,
                      int64_t nhalo, float *d_grid, PassBuffers *b) {
  if (nx < 8 || nx > 1024 || (nx & (nx - 1)))
    return psod_set_error(PSOD_E_ARG, "psod_pass: nx=%d must be a power of "
                          "two in [8, 1024]", nx);
  const int64_t ncell = (int64_t)nx * nx * nx;
  const size_t nh = (size_t)(nhalo > 0 ? nhalo : 1);
  b->cell_start = (int32_t *)psod_scratch(ctx, S_PASS_CS, (size_t)(ncell + 1) * 4);
  b->rec = (float *)psod_scratch(ctx, S_PASS_REC, (size_t)npart * 16);
  b->lab_sorted = (int32_t *)psod_scratch(ctx, S_PASS_LAB, (size_t)npart * 4);
  // seed keys, seeds and member counts
  char *hs = (char *)psod_scratch(ctx, S_PASS_HALO, nh * 16 + 512);
  b->grid = d_grid;
  if (!b->grid)
    b->grid = (float *)psod_scratch(ctx, S_PASS_GRID,
                                    (size_t)ng * (size_t)ng * (size_t)ng * 4);
  if (!b->cell_start || !b->rec || !b->lab_sorted || !hs || !b->grid)
    return PSOD_E_NOMEM;
  b->seed_key = (uint64_t *)hs;
  b->seed = (int32_t *)(hs + nh * 8 + 256);
  b->nmem = b->seed + nh;
  PSOD_TRY(psod_cell_sort(ctx, npart, d_pos, d_labels, nx, b->cell_start,
                          nullptr, b->rec, b->lab_sorted));
  PSOD_TRY(psod_cic_deposit(ctx, npart, nx, b->cell_start, b->rec, ng, 0,
                            b->grid));
  if (nhalo == 0) return PSOD_OK;
  PSOD_TRY(psod_cic_interp(ctx, npart, nx, b->cell_start, b->rec, ng, 0,
                           b->grid, nullptr, nullptr, b->lab_sorted, nhalo,
                           b->seed_key));
  PSOD_TRY(psod_seed_index(ctx, nhalo, b->seed_key, b->seed));
  PSOD_TRY(psod_label_counts(ctx, npart, b->lab_sorted, nhalo, b->nmem));
  return PSOD_OK;
}

extern "C" int psod_pass(psod_ctx *ctx, const psod_sod_params *p,
                         int64_t npart, int32_t nx, int32_t ng,
                         const float *d_pos, const float *d_vel,
                         const int64_t *d_ids, const int32_t *d_labels,
                         int64_t nhalo, float *d_grid,
                         const psod_halo_out *out) {
  PSOD_TRY(psod_use_device(ctx));
  if (!p || !out || !d_pos || !d_labels || npart <= 0 || nhalo < 0)
    return psod_set_error(PSOD_E_ARG, "psod_pass: bad arguments");
  if (p->nx != nx)
    return psod_set_error(PSOD_E_ARG, "psod_pass: params nx=%d, nx=%d",
                          p->nx, nx);
  PassBuffers b;
  PSOD_TRY(psod_pass_prepare(ctx, npart, nx, ng, d_pos, d_labels, nhalo,
                             d_grid, &b));
  if (nhalo == 0) return PSOD_OK;
  psod_catalogue cat;
  cat.npart = npart;
  cat.nx = nx;
  cat.pos = d_pos;
  cat.vel = d_vel;
  cat.ids = d_ids;
  cat.cell_start = b.cell_start;
  cat.rec = b.rec;
  return psod_sod_fof(ctx, p, &cat, nhalo, b.seed, b.nmem, out);
}

// ---------------------------------------------------------------------------
// The pass as one HIP graph.  A graph holds fixed launch shapes and fixed
// pointers, so the pass it replays runs in its device-count form: every
// launch is sized by a bound, the kernels read the work counts on the
// device, and nothing is read back until the graph ends.  The centre
// search's task buffers are sized from an eager pass (twice its task count);
// a replay whose tasks exceed that sets a device flag, and the call then
// runs the pass eagerly and captures it again on the next call.
// ---------------------------------------------------------------------------
int psod_sod_fof_dev(psod_ctx *ctx, const psod_sod_params *p,
                     const psod_catalogue *cat, int64_t nhalo,
                     const int32_t *d_seed, const int32_t *d_nmem,
                     const psod_halo_out *out, bool finish,
                     int32_t *overflow);

namespace {
// everything a captured pass depends on: its arguments and the tunables
struct GraphKey {
  psod_sod_params p;
  int64_t npart, nhalo;
  int32_t nx, ng;
  const void *pos, *vel, *ids, *labels;
  float *grid;
  psod_halo_out out;
  psod_tunables tun;
  int64_t task_cap;
};
}  // namespace

struct GraphState {
  GraphKey key;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

void psod_graph_release(psod_ctx *ctx) {
  if (!ctx || !ctx->graph) return;
  (void)hipGraphExecDestroy(ctx->graph->exec);
  (void)hipGraphDestroy(ctx->graph->graph);
  delete ctx->graph;
  ctx->graph = nullptr;
}

// Queues the whole pass in its device-count form (nothing read back).
static int pass_queue_dev(psod_ctx *ctx, const GraphKey &k) {
  PassBuffers b;
  PSOD_TRY(psod_pass_prepare(ctx, k.npart, k.nx, k.ng, (const float *)k.pos,
                             (const int32_t *)k.labels, k.nhalo, k.grid, &b));
  if (k.nhalo == 0) return PSOD_OK;
  psod_catalogue cat;
  cat.npart = k.npart;
  cat.nx = k.nx;
  cat.pos = (const float *)k.pos;
  cat.vel = (const float *)k.vel;
  cat.ids = (const int64_t *)k.ids;
  cat.cell_start = b.cell_start;
  cat.rec = b.rec;
  return psod_sod_fof_dev(ctx, &k.p, &cat, k.nhalo, b.seed, b.nmem, &k.out,
                          false, nullptr);
}

// After the device-count pass: failures, re-runs, the overflow flag.
static int pass_finish_dev(psod_ctx *ctx, const GraphKey &k,
                           int32_t *overflow) {
  *overflow = 0;
  if (k.nhalo == 0) return PSOD_OK;
  PassBuffers b;
  // the same scratch the queued pass used (no kernel runs here)
  const int64_t ncell = (int64_t)k.nx * k.nx * k.nx;
  b.cell_start = (int32_t *)psod_scratch(ctx, S_PASS_CS, (size_t)(ncell + 1) * 4);
  b.rec = (float *)psod_scratch(ctx, S_PASS_REC, (size_t)k.npart * 16);
  char *hs = (char *)psod_scratch(ctx, S_PASS_HALO, (size_t)k.nhalo * 16 + 512);
  if (!b.cell_start || !b.rec || !hs) return PSOD_E_NOMEM;
  int32_t *seed = (int32_t *)(hs + (size_t)k.nhalo * 8 + 256);
  int32_t *nmem = seed + k.nhalo;
  psod_catalogue cat;
  cat.npart = k.npart;
  cat.nx = k.nx;
  cat.pos = (const float *)k.pos;
  cat.vel = (const float *)k.vel;
  cat.ids = (const int64_t *)k.ids;
  cat.cell_start = b.cell_start;
  cat.rec = b.rec;
  return psod_sod_fof_dev(ctx, &k.p, &cat, k.nhalo, seed, nmem, &k.out, true,
                          overflow);
}

static int both_streams(psod_ctx *ctx) {
  PSOD_HIP(hipStreamSynchronize(ctx->side));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}

extern "C" int psod_pass_graph(psod_ctx *ctx, const psod_sod_params *p,
                               int64_t npart, int32_t nx, int32_t ng,
                               const float *d_pos, const float *d_vel,
                               const int64_t *d_ids, const int32_t *d_labels,
                               int64_t nhalo, float *d_grid,
                               const psod_halo_out *out, int32_t *replayed) {
  PSOD_TRY(psod_use_device(ctx));
  if (replayed) *replayed = 0;
  if (!p || !out || !d_pos || !d_labels || npart <= 0 || nhalo < 0)
    return psod_set_error(PSOD_E_ARG, "psod_pass_graph: bad arguments");
  if (p->nx != nx)
    return psod_set_error(PSOD_E_ARG, "psod_pass_graph: params nx=%d, nx=%d",
                          p->nx, nx);
  GraphKey k;
  memset(&k, 0, sizeof(k));  // padding included: keys compare with memcmp
  k.p.nx = p->nx;            // fields one by one (the caller's padding is
  k.p.ndelta = p->ndelta;    // not copied)
  for (int i = 0; i < PSOD_MAXDELTA; i++) k.p.delta[i] = p->delta[i];
  k.p.delta2 = p->delta2;
  k.p.nbins = p->nbins;
  k.p.rmin_frac = p->rmin_frac;
  k.p.rmax_frac = p->rmax_frac;
  k.p.mmin = p->mmin;
  k.p.mpart = p->mpart;
  k.p.nmin_count = p->nmin_count;
  k.p.rsearch = p->rsearch;
  k.p.eps = p->eps;
  k.p.tol_abs = p->tol_abs;
  k.p.small_max = p->small_max;
  k.p.flags = p->flags;
  k.npart = npart;
  k.nhalo = nhalo;
  k.nx = nx;
  k.ng = ng;
  k.pos = d_pos;
  k.vel = d_vel;
  k.ids = d_ids;
  k.labels = d_labels;
  k.grid = d_grid;
  k.out = *out;
  k.tun = psod_tunables_for(ctx, nx);
  GraphState *g = ctx->graph;
  if (g) {
    k.task_cap = g->key.task_cap;
    if (memcmp(&k, &g->key, sizeof(k)) == 0) {
      PSOD_HIP(hipGraphLaunch(g->exec, ctx->stream));
      int32_t ov = 0;
      const int rc = pass_finish_dev(ctx, k, &ov);
      (void)both_streams(ctx);
      if (rc != PSOD_OK || !ov) {
        if (rc == PSOD_OK && replayed) *replayed = 1;
        return rc;
      }
      // more centre-search tasks than the graph holds: eager this time,
      // captured again with a larger capacity on the next call
      psod_graph_release(ctx);
      return psod_pass(ctx, p, npart, nx, ng, d_pos, d_vel, d_ids, d_labels,
                       nhalo, d_grid, out);
    }
    psod_graph_release(ctx);
  }
  // 1. the eager pass: this call's outputs and its centre-search task count
  PSOD_TRY(psod_pass(ctx, p, npart, nx, ng, d_pos, d_vel, d_ids, d_labels,
                     nhalo, d_grid, out));
  if (nhalo == 0) return PSOD_OK;
  ctx->gcap.tasks = std::max<int64_t>(2 * ctx->sod_stat[7], 1 << 16);
  k.task_cap = ctx->gcap.tasks;
  // 2. the device-count pass once, eagerly: it sizes every buffer of that
  // form, which may not grow while the pass is captured
  PSOD_TRY(pass_queue_dev(ctx, k));
  int32_t ov = 0;
  PSOD_TRY(pass_finish_dev(ctx, k, &ov));
  PSOD_TRY(both_streams(ctx));
  if (ov)
    return psod_set_error(PSOD_E_CAPACITY, "psod_pass_graph: the task "
                          "capacity %lld of an eager count did not hold",
                          (long long)k.task_cap);
  // 3. capture
  hipGraph_t graph = nullptr;
  PSOD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = 1;
  const int rq = pass_queue_dev(ctx, k);
  ctx->capturing = 0;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  if (rq != PSOD_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rq;
  }
  if (ec != hipSuccess || !graph)
    return psod_set_error(PSOD_E_DEVICE, "psod_pass_graph: capture failed "
                          "(%s)", hipGetErrorString(ec));
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiateWithFlags(&exec, graph, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return psod_set_error(PSOD_E_DEVICE, "psod_pass_graph: instantiate "
                          "failed (%s)", hipGetErrorString(ei));
  }
  g = new GraphState;
  g->key = k;
  g->graph = graph;
  g->exec = exec;
  ctx->graph = g;
  return PSOD_OK;
}
