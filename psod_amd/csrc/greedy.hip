// greedy.hip -- `spherover` (main file:1396-2201) with the reference's
// greedy semantics: seeds in decreasing density, a seed is skipped when an
// earlier sphere captured it, and every grown sphere removes its particles
// (accepted or not) before later seeds gather.
//
// Exact sequential semantics in parallel: seed s only reads and writes the
// (2 ncut + 1)^3 cells around its cell, so two seeds whose cubes do not
// intersect commute.  Each round marks, per 8^3 block, the earliest
// unprocessed live seed; a seed whose neighbourhood of blocks holds no
// earlier unprocessed live seed runs in this round.  Rounds repeat until
// every seed up to npart_new is processed or dead.  Per-seed arithmetic is
// FP64 on the float32 positions scaled by nx, as in the reference.
#include <cmath>
#include <cstring>
#include <vector>

#include "sod_common.h"

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);

namespace {

constexpr int GT = 256;
constexpr int64_t GCAP = 1 << 18;  // gathered particles per workgroup slice

struct GreedyCfg {
  int32_t nx, ncut, nbins, nb;
  double seuil, seuil2, mmin, lr0, lr1, rmaxf;
  int64_t nsel;
};

struct GEntry {   // one gathered particle
  double x, y, z;  // scaled position relative to the seed cell origin
  int32_t slot, idx;
  int32_t off;     // Chebyshev offset of its cell from the seed cell
  int32_t rank;    // candidate order: walk rank of its cell (3^3) or -1
};

struct GRec {
  int64_t ipart, seed, centre, iamas, masse, masse2, npart_sel;
  double amass, xc, yc, zc, rayon, overdens, overdens2;
  int32_t accepted, status;
};

__global__ void k_inverse_perm(const int32_t *__restrict__ perm, int64_t n,
                               int32_t *__restrict__ slot_of) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) slot_of[perm[k]] = (int32_t)k;
}

__global__ void k_set_u8(uint8_t *__restrict__ p, int64_t n, uint8_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// per block: smallest order of an unprocessed live seed in [lo, hi)
__global__ void k_mark_blocks(GreedyCfg c, const int32_t *__restrict__ seeds,
                              const int32_t *__restrict__ slot_of,
                              const float *__restrict__ pos,
                              const uint8_t *__restrict__ alive,
                              const uint8_t *__restrict__ done, int64_t lo,
                              int64_t hi, int32_t *__restrict__ bmin) {
  const int64_t o = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= hi || done[o]) return;
  const int32_t s = seeds[o];
  if (!alive[slot_of[s]]) return;
  int32_t b[3];
  for (int d = 0; d < 3; d++) {
    int32_t cc;
    float f;
    split_coord(pos[3 * (int64_t)s + d], c.nx, &cc, &f);
    b[d] = cc >> 3;
  }
  atomicMin(&bmin[(b[2] * c.nb + b[1]) * c.nb + b[0]], (int32_t)(o - lo));
}

// ready seeds: live, unprocessed, and no earlier unprocessed live seed in
// the blocks a (2ncut+1)^3 cube around either seed can reach
__global__ void k_ready(GreedyCfg c, const int32_t *__restrict__ seeds,
                        const int32_t *__restrict__ slot_of,
                        const float *__restrict__ pos,
                        const uint8_t *__restrict__ alive,
                        uint8_t *__restrict__ done, int64_t lo, int64_t hi,
                        const int32_t *__restrict__ bmin,
                        int32_t *__restrict__ ready,
                        int32_t *__restrict__ nready) {
  const int64_t o = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= hi || done[o]) return;
  const int32_t s = seeds[o];
  if (!alive[slot_of[s]]) {  // captured earlier: nothing to do (:1628)
    done[o] = 1;
    return;
  }
  int32_t b[3];
  for (int d = 0; d < 3; d++) {
    int32_t cc;
    float f;
    split_coord(pos[3 * (int64_t)s + d], c.nx, &cc, &f);
    b[d] = cc >> 3;
  }
  // cubes intersect only if the seed cells are <= 2 ncut apart
  const int reach = (2 * c.ncut + 7) / 8 + 1;
  const int span = min(2 * reach + 1, c.nb);
  const int32_t me = (int32_t)(o - lo);
  for (int i = 0; i < span; i++)
    for (int j = 0; j < span; j++)
      for (int k = 0; k < span; k++) {
        const int32_t bx = wrapi(b[0] - reach + i, c.nb);
        const int32_t by = wrapi(b[1] - reach + j, c.nb);
        const int32_t bz = wrapi(b[2] - reach + k, c.nb);
        if (bmin[(bz * c.nb + by) * c.nb + bx] < me) return;
      }
  ready[atomicAdd(nready, 1)] = me;  // relative to the window start lo
}

struct GShared {
  int32_t cnt, ncand;
  int64_t red_l[GT / 64];
  double red_d[GT / 64];
  int64_t bk[GT / 64];
  double bs[GT / 64];
  int32_t item;
};

__device__ inline int64_t gmin(int64_t v, GShared &S) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, (int64_t)__shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) S.red_l[threadIdx.x >> 6] = v;
  __syncthreads();
  int64_t r = S.red_l[0];
  for (int i = 1; i < GT / 64; i++) r = min(r, S.red_l[i]);
  __syncthreads();
  return r;
}

// one workgroup per ready seed, persistent over the ready list
__global__ void __launch_bounds__(GT)
k_greedy_seed(GreedyCfg c, const int32_t *__restrict__ seeds,
              const int32_t *__restrict__ cell_start,
              const float4 *__restrict__ rec, const float *__restrict__ pos,
              const int64_t *__restrict__ ids, uint8_t *__restrict__ alive,
              uint8_t *__restrict__ done, int64_t lo, int64_t hi,
              const int32_t *__restrict__ ready,
              const int32_t *__restrict__ nready,
              int32_t *__restrict__ queue, GEntry *__restrict__ ent_all,
              double *__restrict__ d2_all, int32_t *__restrict__ ord_all,
              int64_t *__restrict__ structure, GRec *__restrict__ recs,
              double *__restrict__ radb, int32_t *__restrict__ massb,
              int32_t *__restrict__ fail) {
  __shared__ GShared S;
  GEntry *ent = ent_all + (int64_t)blockIdx.x * GCAP;
  double *d2 = d2_all + (int64_t)blockIdx.x * GCAP;
  int32_t *ord = ord_all + (int64_t)blockIdx.x * GCAP;
  const int32_t nr = *nready;
  const double taille = (double)c.nx;
  for (;;) {
    if (threadIdx.x == 0) S.item = atomicAdd(queue, 1);
    __syncthreads();
    const int32_t it = S.item;
    __syncthreads();
    if (it >= nr) break;
    const int64_t o = lo + ready[it];
    if (o < lo || o >= hi) continue;  // never: ready holds window offsets
    const int32_t s = seeds[o];
    GRec R;
    memset(&R, 0, sizeof(R));
    R.ipart = o + 1;
    R.seed = s;
    int32_t cs[3];
    for (int d = 0; d < 3; d++) {
      float f;
      split_coord(pos[3 * (int64_t)s + d], c.nx, &cs[d], &f);
    }
    // S1: live particles of the (2ncut+1)^3 cells, walk x outermost
    // (main file:1658-1716)
    if (threadIdx.x == 0) S.cnt = 0;
    __syncthreads();
    const int side = 2 * c.ncut + 1;
    const int32_t ncell = side * side * side;
    // cells are visited in walk order by one thread each, so that the
    // gathered list keeps the reference's order (needed for exact ties)
    for (int32_t base = 0; base < ncell; base += GT) {
      const int32_t q = base + threadIdx.x;
      int32_t a = 0, b = 0, lc[3] = {0, 0, 0};
      int di = 0, dj = 0, dk = 0;
      int32_t nlive = 0;
      if (q < ncell) {
        walk_off(q, c.ncut, &di, &dj, &dk);
        const int32_t key = key_at(c.nx, cs, di, dj, dk, lc);
        a = cell_start[key];
        b = cell_start[key + 1];
        for (int32_t k = a; k < b; k++) nlive += alive[k];
      }
      // exclusive prefix over this batch of cells
      int32_t inc = wave_incl_scan(nlive);
      const int w = threadIdx.x >> 6;
      if ((threadIdx.x & 63) == 63) S.red_l[w] = inc;
      __syncthreads();
      int32_t off = S.cnt, tot = 0;
      for (int i = 0; i < GT / 64; i++) {
        if (i < w) off += (int32_t)S.red_l[i];
        tot += (int32_t)S.red_l[i];
      }
      off += inc - nlive;
      if (q < ncell) {
        int32_t e = off;
        for (int32_t k = a; k < b; k++) {
          if (!alive[k]) continue;
          if (e < GCAP && e < c.nsel) {
            const float4 r = rec[k];
            GEntry g;
            g.x = (double)cs[0] + (double)di + ((double)r.x - (double)lc[0]);
            g.y = (double)cs[1] + (double)dj + ((double)r.y - (double)lc[1]);
            g.z = (double)cs[2] + (double)dk + ((double)r.z - (double)lc[2]);
            g.slot = k;
            g.idx = __float_as_int(r.w);
            g.off = max(abs(di), max(abs(dj), abs(dk)));
            g.rank = g.off <= 1 ? (di + 1) * 9 + (dj + 1) * 3 + (dk + 1) : -1;
            ent[e] = g;
          }
          e++;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) S.cnt += tot;
      __syncthreads();
    }
    const int32_t n1 = S.cnt;
    if (n1 > c.nsel || n1 > GCAP) {  // "Increase Nsel" (main file:1681)
      if (threadIdx.x == 0) {
        R.status = n1 > c.nsel ? PSOD_E_CAPACITY : -PSOD_E_CAPACITY;
        recs[o] = R;
        atomicMin(&fail[0], (int32_t)o);
      }
      __syncthreads();
      continue;
    }
    __threadfence_block();
    __syncthreads();
    // Mtotsel >= Mmin with unit masses (main file:1759)
    if (!((double)n1 >= c.mmin)) {
      if (threadIdx.x == 0) done[o] = 1;
      continue;
    }
    // centre search over the 3^3 candidates against S2 (offset <= 2),
    // first strict maximum in walk order (main file:1799-1875)
    double bs = 0.0;
    int64_t bk = INT64_MAX;
    for (int32_t t = threadIdx.x; t < n1; t += GT) {
      const GEntry p = ent[t];
      if (p.rank < 0) continue;
      double acc = 0.0;
      for (int32_t u = 0; u < n1; u++) {
        const GEntry q = ent[u];
        if (q.off > 2) continue;
        double dx = fabs(q.x - p.x);
        dx = fmin(dx, taille - dx);
        double dy = fabs(q.y - p.y);
        dy = fmin(dy, taille - dy);
        double dz = fabs(q.z - p.z);
        dz = fmin(dz, taille - dz);
        const double dd = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx),
                                              __dmul_rn(dy, dy)),
                                    __dmul_rn(dz, dz));
        if (dd <= 1.0) acc += 1.0 / sqrt(dd + (1.0 / 64.0) * (1.0 / 64.0));
      }
      const int64_t key = ((int64_t)p.rank << 32) | (uint32_t)p.idx;
      if (acc > bs || (acc == bs && key < bk)) { bs = acc; bk = key; }
    }
    for (int sh = 32; sh > 0; sh >>= 1) {
      const double s2 = __shfl_xor(bs, sh, 64);
      const int64_t k2 = __shfl_xor(bk, sh, 64);
      if (s2 > bs || (s2 == bs && k2 < bk)) { bs = s2; bk = k2; }
    }
    if ((threadIdx.x & 63) == 0) {
      S.bs[threadIdx.x >> 6] = bs;
      S.bk[threadIdx.x >> 6] = bk;
    }
    __syncthreads();
    bs = S.bs[0];
    bk = S.bk[0];
    for (int i = 1; i < GT / 64; i++)
      if (S.bs[i] > bs || (S.bs[i] == bs && S.bk[i] < bk)) {
        bs = S.bs[i];
        bk = S.bk[i];
      }
    __syncthreads();
    const int32_t centre = (int32_t)(bk & 0xFFFFFFFF);
    double xc[3];
    centre_scaled(pos, centre, c.nx, xc);
    // distances (main file:2002-2020) and sort by (d2, gather order)
    for (int32_t t = threadIdx.x; t < n1; t += GT) {
      const GEntry p = ent[t];
      double dx = fabs((double)(pos[3 * (int64_t)p.idx] * (float)c.nx) - xc[0]);
      dx = fmin(dx, taille - dx);
      double dy = fabs((double)(pos[3 * (int64_t)p.idx + 1] * (float)c.nx) - xc[1]);
      dy = fmin(dy, taille - dy);
      double dz = fabs((double)(pos[3 * (int64_t)p.idx + 2] * (float)c.nx) - xc[2]);
      dz = fmin(dz, taille - dz);
      d2[t] = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)),
                        __dmul_rn(dz, dz));
      ord[t] = t;
    }
    int32_t np = 1;
    while (np < n1) np <<= 1;
    for (int32_t t = n1 + threadIdx.x; t < np; t += GT) {
      d2[t] = INFINITY;
      ord[t] = INT32_MAX;
    }
    __threadfence_block();
    __syncthreads();
    for (int32_t k = 2; k <= np; k <<= 1)
      for (int32_t j = k >> 1; j > 0; j >>= 1) {
        for (int32_t i = threadIdx.x; i < np; i += GT) {
          const int32_t p = i ^ j;
          if (p > i) {
            const double a = d2[i], b = d2[p];
            const int32_t ia = ord[i], ib = ord[p];
            const bool gt = a > b || (a == b && ia > ib);
            if (gt == ((i & k) == 0)) {
              d2[i] = b; d2[p] = a; ord[i] = ib; ord[p] = ia;
            }
          }
        }
        __threadfence_block();
        __syncthreads();
      }
    // growth (main file:2040-2051): stop at the first i >= 10 with
    // O(i) <= seuil, i == npart_sel or r_i >= ncut
    int64_t stop = INT64_MAX;
    for (int64_t i = 10 + threadIdx.x; i <= n1; i += GT) {
      const double r = sqrt(d2[i - 1]);
      if (od_k(i, 1.0, r) <= c.seuil || i >= n1 || r >= (double)c.ncut)
        stop = min(stop, i);
    }
    stop = gmin(stop, S);
    if (stop == INT64_MAX || od_k(stop, 1.0, sqrt(d2[stop - 1])) > c.seuil) {
      // "Increase selection radius" (main file:2072-2089)
      if (threadIdx.x == 0) {
        R.status = PSOD_E_UNCLOSED;
        R.npart_sel = n1;
        recs[o] = R;
        atomicMin(&fail[0], (int32_t)o);
      }
      __syncthreads();
      continue;
    }
    // second threshold: last i in [2, stop] above seuil2 (:2062-2069)
    int64_t m2 = 0;
    for (int64_t i = 2 + threadIdx.x; i <= stop; i += GT)
      if (od_k(i, 1.0, sqrt(d2[i - 1])) > c.seuil2) m2 = max(m2, i);
    m2 = -gmin(-m2, S);
    const int64_t masse = stop - 1;
    const double rayon = sqrt(d2[masse - 1]);
    const double volume = 2.0 / 3.0 * TWOPI_REF * (rayon * rayon * rayon);
    const double amass = (double)masse;
    // profile (main file:2103-2115)
    for (int j = threadIdx.x; j < c.nbins; j += GT) {
      // the last radius is rayon * rmax_frac exactly
      const double rj =
          j == c.nbins - 1 ? rayon * c.rmaxf
                           : rayon * exp(c.lr0 + (double)j * (c.lr1 - c.lr0) /
                                                     ((double)c.nbins - 1.0));
      int32_t lo2 = 0, hi2 = n1;
      while (lo2 < hi2) {
        const int32_t mid = (lo2 + hi2) >> 1;
        if (sqrt(d2[mid]) < rj) lo2 = mid + 1; else hi2 = mid;
      }
      radb[o * c.nbins + j] = rj;
      massb[o * c.nbins + j] = lo2;
    }
    const bool acc = amass >= c.mmin && xc[0] <= taille && xc[0] > 0.0 &&
                     xc[1] <= taille && xc[1] > 0.0 && xc[2] <= taille &&
                     xc[2] > 0.0;
    const int64_t iamas = ids ? ids[centre] : (int64_t)centre;
    // exclusion of the first masse particles, accepted or not
    for (int64_t t = threadIdx.x; t < masse; t += GT) {
      const GEntry p = ent[ord[t]];
      alive[p.slot] = 0;
      structure[p.idx] = acc ? iamas : 0;
    }
    if (threadIdx.x == 0) {
      R.centre = centre;
      R.iamas = iamas;
      R.masse = masse;
      R.masse2 = m2;
      R.npart_sel = n1;
      R.amass = amass;
      R.xc = xc[0];
      R.yc = xc[1];
      R.zc = xc[2];
      R.rayon = rayon;
      R.overdens = volume > 0.0 ? amass / volume : 0.0;
      if (m2 > 0) {
        const double r2 = sqrt(d2[m2 - 1]);
        R.overdens2 = r2 > 1e-30 ? od_k(m2, 1.0, r2) : 0.0;
      }
      R.accepted = acc ? 1 : 0;
      R.status = 1;  // a processed seed
      recs[o] = R;
      done[o] = 1;
    }
    __syncthreads();
  }
}

__global__ void k_alive_orig(const uint8_t *__restrict__ alive_s,
                             const int32_t *__restrict__ perm, int64_t n,
                             uint8_t *__restrict__ alive) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) alive[perm[k]] = alive_s[k];
}

}  // namespace

extern "C" int psod_host_spherover(psod_ctx *ctx, int64_t npart,
                                   const float *pos, const int64_t *ids,
                                   const int32_t *isort, int64_t npart_new,
                                   const psod_greedy_params *p, int64_t maxrec,
                                   int64_t *nrec, int64_t *rec_i64,
                                   double *rec_f64, double *rad_bins,
                                   int32_t *mass_bins, uint8_t *alive,
                                   int64_t *structure) {
  PSOD_TRY(psod_use_device(ctx));
  if (!pos || !isort || !p || !nrec || npart <= 0 || npart_new < 0)
    return psod_set_error(PSOD_E_ARG, "spherover: bad arguments");
  if (p->nx < 8 || p->nx > 1024 || (p->nx & (p->nx - 1)) || p->ncut < 1 ||
      2 * p->ncut + 1 > p->nx || p->nbins < 2 || p->nsel < 1 ||
      !(p->rmin_frac > 0.0) || !(p->rmax_frac > 0.0))
    return psod_set_error(PSOD_E_ARG, "spherover: bad parameters");
  *nrec = 0;
  if (npart_new > npart) npart_new = npart;
  // the seeds index the device's particle arrays: reject any outside them
  // before anything is queued
  for (int64_t o = 0; o < npart_new; o++)
    if (isort[o] < 1 || isort[o] > npart)
      return psod_set_error(PSOD_E_ARG,
                            "spherover: isort(%lld) = %d outside [1, %lld]",
                            (long long)(o + 1), isort[o], (long long)npart);
  GreedyCfg c;
  c.nx = p->nx;
  c.ncut = p->ncut;
  c.nbins = p->nbins;
  c.nb = p->nx / 8;
  c.seuil = p->seuil;
  c.seuil2 = p->seuil2;
  c.mmin = p->mmin;
  c.lr0 = log(p->rmin_frac);
  c.lr1 = log(p->rmax_frac);
  c.rmaxf = p->rmax_frac;
  c.nsel = p->nsel;
  hipStream_t st = ctx->stream;
  const int64_t ncellt = (int64_t)p->nx * p->nx * p->nx;
  const int64_t nn = npart_new > 0 ? npart_new : 1;
  float *dpos = (float *)psod_scratch(ctx, S_HOST_A, (size_t)npart * 12);
  int32_t *cs = (int32_t *)psod_scratch(ctx, S_HOST_B, (size_t)(ncellt + 1) * 4);
  int32_t *perm = (int32_t *)psod_scratch(ctx, S_HOST_C, (size_t)npart * 4);
  float *drec = (float *)psod_scratch(ctx, S_HOST_D, (size_t)npart * 16);
  char *A = (char *)psod_scratch(ctx, S_GREEDY,
                                 (size_t)npart * (4 + 1 + 8 + 8 + 1) +
                                     (size_t)nn * (4 + 1 + 4) + 4096);
  if (!dpos || !cs || !perm || !drec || !A) return PSOD_E_NOMEM;
  int32_t *slot_of = (int32_t *)A;
  int64_t *dstruct = (int64_t *)(A + (((size_t)npart * 4 + 255) & ~(size_t)255));
  int64_t *dids = dstruct + npart;
  uint8_t *alive_s = (uint8_t *)(dids + npart);
  uint8_t *alive_o = alive_s + npart;
  int32_t *dseeds = (int32_t *)(((uintptr_t)(alive_o + npart) + 255) & ~(uintptr_t)255);
  uint8_t *done = (uint8_t *)(dseeds + nn);
  int32_t *ready = (int32_t *)(((uintptr_t)(done + nn) + 255) & ~(uintptr_t)255);
  const int32_t NG = ctx->num_cus * 2;
  const int32_t nbk = c.nb * c.nb * c.nb;
  char *B = (char *)psod_scratch(
      ctx, S_GREEDY2,
      (size_t)NG * GCAP * (sizeof(GEntry) + 8 + 4) + (size_t)nbk * 4 + 1024);
  GRec *recs = (GRec *)psod_scratch(ctx, S_GREEDY3,
                                    (size_t)nn * (sizeof(GRec) +
                                                  (size_t)p->nbins * 12) + 256);
  if (!B || !recs) return PSOD_E_NOMEM;
  GEntry *ent = (GEntry *)B;
  double *d2 = (double *)(ent + (size_t)NG * GCAP);
  int32_t *ord = (int32_t *)(d2 + (size_t)NG * GCAP);
  int32_t *bmin = ord + (size_t)NG * GCAP;
  int32_t *ctr = bmin + nbk;  // [0] nready, [1] queue, [2] fail
  double *radb = (double *)(recs + nn);
  int32_t *massb = (int32_t *)(radb + (size_t)nn * p->nbins);

  PSOD_HIP(hipMemcpyAsync(dpos, pos, (size_t)npart * 12, hipMemcpyHostToDevice, st));
  PSOD_TRY(psod_cell_sort(ctx, npart, dpos, nullptr, p->nx, cs, perm, drec,
                          nullptr));
  if (ids)
    PSOD_HIP(hipMemcpyAsync(dids, ids, (size_t)npart * 8, hipMemcpyHostToDevice, st));
  k_inverse_perm<<<grid_for(npart, 256), 256, 0, st>>>(perm, npart, slot_of);
  k_set_u8<<<grid_for(npart, 256, 4096), 256, 0, st>>>(alive_s, npart, 1);
  PSOD_HIP(hipMemsetAsync(dstruct, 0, (size_t)npart * 8, st));
  PSOD_HIP(hipMemsetAsync(done, 0, (size_t)nn, st));
  PSOD_HIP(hipMemsetAsync(recs, 0, (size_t)nn * sizeof(GRec), st));
  // seeds: isort is 1-based particle indices
  std::vector<int32_t> hs(nn);
  for (int64_t o = 0; o < npart_new; o++) hs[o] = isort[o] - 1;
  PSOD_HIP(hipMemcpyAsync(dseeds, hs.data(), (size_t)npart_new * 4,
                          hipMemcpyHostToDevice, st));
  PSOD_TRY(psod_fill_i32(ctx, ctr + 2, 1, INT32_MAX));
  PSOD_HIP(hipStreamSynchronize(st));
  // rounds over a sliding window of seeds
  const int64_t WIN = 1 << 16;
  std::vector<uint8_t> hdone;
  int64_t lo = 0;
  int rounds = 0;
  while (lo < npart_new) {
    const int64_t hi = std::min<int64_t>(npart_new, lo + WIN);
    PSOD_TRY(psod_fill_i32(ctx, bmin, nbk, INT32_MAX));
    PSOD_TRY(psod_fill_i32(ctx, ctr, 2, 0));
    const unsigned gw = grid_for(hi - lo, 256);
    k_mark_blocks<<<gw, 256, 0, st>>>(c, dseeds, slot_of, dpos, alive_s, done,
                                      lo, hi, bmin);
    k_ready<<<gw, 256, 0, st>>>(c, dseeds, slot_of, dpos, alive_s, done, lo,
                                hi, bmin, ready, ctr);
    int32_t nready = 0;
    PSOD_HIP(hipMemcpyAsync(&nready, ctr, 4, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    if (nready > 0) {
      k_greedy_seed<<<(unsigned)std::min(nready, NG), GT, 0, st>>>(
          c, dseeds, cs, (const float4 *)drec, dpos, ids ? dids : nullptr,
          alive_s, done, lo, hi, ready, ctr, ctr + 1, ent, d2, ord, dstruct,
          recs, radb, massb, ctr + 2);
      PSOD_LAUNCH_CHECK();
    }
    int32_t fail = INT32_MAX;
    PSOD_HIP(hipMemcpyAsync(&fail, ctr + 2, 4, hipMemcpyDeviceToHost, st));
    hdone.resize((size_t)(hi - lo));
    PSOD_HIP(hipMemcpyAsync(hdone.data(), done + lo, (size_t)(hi - lo),
                            hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    rounds++;
    if (fail != INT32_MAX) break;
    int64_t adv = 0;
    while (adv < hi - lo && hdone[(size_t)adv]) adv++;
    lo += adv;
    if ((adv == 0 && nready == 0) || rounds > npart_new + 2)
      return psod_set_error(PSOD_E_DEVICE, "spherover: no progress");
  }
  int32_t fail = INT32_MAX;
  PSOD_HIP(hipMemcpyAsync(&fail, ctr + 2, 4, hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipStreamSynchronize(st));
  // records in seed order, up to the first failing seed
  std::vector<GRec> hr((size_t)nn);
  std::vector<double> hrb((size_t)nn * p->nbins);
  std::vector<int32_t> hmb((size_t)nn * p->nbins);
  PSOD_HIP(hipMemcpyAsync(hr.data(), recs, (size_t)nn * sizeof(GRec),
                          hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipMemcpyAsync(hrb.data(), radb, hrb.size() * 8,
                          hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipMemcpyAsync(hmb.data(), massb, hmb.size() * 4,
                          hipMemcpyDeviceToHost, st));
  if (alive) {
    k_alive_orig<<<grid_for(npart, 256), 256, 0, st>>>(alive_s, perm, npart,
                                                       alive_o);
    PSOD_HIP(hipMemcpyAsync(alive, alive_o, (size_t)npart,
                            hipMemcpyDeviceToHost, st));
  }
  if (structure)
    PSOD_HIP(hipMemcpyAsync(structure, dstruct, (size_t)npart * 8,
                            hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipStreamSynchronize(st));
  const int64_t upto = fail == INT32_MAX ? npart_new : (int64_t)fail;
  int64_t k = 0;
  for (int64_t o = 0; o < upto; o++) {
    const GRec &g = hr[(size_t)o];
    if (g.status != 1) continue;
    if (k >= maxrec) return psod_set_error(PSOD_E_CAPACITY,
                                           "spherover: more than %lld records",
                                           (long long)maxrec);
    if (rec_i64) {
      int64_t *ri = rec_i64 + 6 * k;
      ri[0] = g.ipart; ri[1] = g.seed; ri[2] = g.centre; ri[3] = g.iamas;
      ri[4] = g.masse; ri[5] = g.masse2;
    }
    if (rec_f64) {
      double *rf = rec_f64 + 8 * k;
      rf[0] = g.amass; rf[1] = g.xc; rf[2] = g.yc; rf[3] = g.zc;
      rf[4] = g.rayon; rf[5] = g.overdens; rf[6] = g.overdens2;
      rf[7] = g.accepted;
    }
    if (rad_bins)
      memcpy(rad_bins + k * p->nbins, &hrb[(size_t)o * p->nbins],
             sizeof(double) * p->nbins);
    if (mass_bins)
      memcpy(mass_bins + k * p->nbins, &hmb[(size_t)o * p->nbins],
             sizeof(int32_t) * p->nbins);
    k++;
  }
  *nrec = k;
  if (fail != INT32_MAX) {
    const GRec &g = hr[(size_t)fail];
    if (g.status == -PSOD_E_CAPACITY)
      return psod_set_error(PSOD_E_CAPACITY,
                            "spherover: seed %lld gathers more than %lld "
                            "particles (device slice)",
                            (long long)(fail + 1), (long long)GCAP);
    return psod_set_error(g.status,
                          g.status == PSOD_E_CAPACITY
                              ? "spherover: Nsel exceeded at seed %lld"
                              : "spherover: sphere of seed %lld not closed "
                                "within ncut",
                          (long long)(fail + 1));
  }
  (void)rounds;
  return PSOD_OK;
}
