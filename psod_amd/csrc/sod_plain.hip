// sod_plain.hip -- the plain method: the reference's SOD algorithm in FP64
// with one launch sequence per halo (main file:1795-2115).
//
// Per halo:
//   k_plain_centre  one thread per candidate of the 3^3 cells around the
//                   seed cell; FP64 S(c) over the particles of the 3^3 cells
//                   around the candidate's cell (every neighbour within one
//                   cell lies there), plain FP64 sum;
//   k_plain_grow    one workgroup: first strict maximum of S in the cell
//                   walk order, FP64 squared distances of every particle
//                   within the reach, bitonic sort in device memory, the
//                   growth walk, profile, second threshold, centre of mass;
//                   doubles the reach until every radius fits.
// It is the comparator of the batched split-precision pass and produces
// the same outputs.
#include "sod_common.h"

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);

constexpr int PT = 1024;

struct PlainHalo {
  int32_t seed, nmem, cs[3], ncand, w0, pad;
};

__global__ void k_plain_plan(SodCfg c, int64_t nhalo,
                             const int32_t *__restrict__ seed,
                             const int32_t *__restrict__ nmem,
                             const float *__restrict__ pos,
                             const int32_t *__restrict__ cell_start,
                             PlainHalo *__restrict__ ph) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nhalo) return;
  PlainHalo p;
  p.seed = seed[h];
  p.nmem = nmem ? nmem[h] : 0;
  p.ncand = 0;
  p.pad = 0;
  p.cs[0] = p.cs[1] = p.cs[2] = 0;
  p.w0 = reach_estimate(c, p.nmem);
  if (p.seed >= 0) {
    const CellRef sr = ref_of_particle(pos, p.seed, c.nx);
    for (int d = 0; d < 3; d++) p.cs[d] = sr.c[d];
    for (int q = 0; q < 27; q++) {
      int di, dj, dk;
      walk_off(q, 1, &di, &dj, &dk);
      int32_t l[3];
      const int32_t key = key_at(c.nx, p.cs, di, dj, dk, l);
      p.ncand += cell_start[key + 1] - cell_start[key];
    }
  }
  ph[h] = p;
}

// candidate t of the 27-cell walk: cell rank q and sorted slot
__device__ inline void plain_cand(const SodCfg &c, const int32_t *cs,
                                  const int32_t *cell_start, int32_t t,
                                  int *q, int32_t *slot) {
  int32_t acc = 0;
  for (int r = 0; r < 27; r++) {
    int di, dj, dk;
    walk_off(r, 1, &di, &dj, &dk);
    int32_t l[3];
    const int32_t key = key_at(c.nx, cs, di, dj, dk, l);
    const int32_t a = cell_start[key], n = cell_start[key + 1] - a;
    if (t < acc + n) {
      *q = r;
      *slot = a + (t - acc);
      return;
    }
    acc += n;
  }
  *q = -1;
  *slot = -1;
}

__global__ void __launch_bounds__(256)
k_plain_centre(SodCfg c, const PlainHalo *__restrict__ ph, int64_t h,
               const int32_t *__restrict__ cell_start,
               const float4 *__restrict__ rec, double *__restrict__ s_out) {
  const PlainHalo p = ph[h];
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.ncand) return;
  int q;
  int32_t slot;
  plain_cand(c, p.cs, cell_start, t, &q, &slot);
  int ci, cj, ck;
  walk_off(q, 1, &ci, &cj, &ck);
  int32_t lc[3];
  key_at(c.nx, p.cs, ci, cj, ck, lc);
  const float4 rc = rec[slot];
  const double px = rel1d(rc.x, lc[0], ci, 0.0);
  const double py = rel1d(rc.y, lc[1], cj, 0.0);
  const double pz = rel1d(rc.z, lc[2], ck, 0.0);
  double s = 0.0;
  for (int r = 0; r < 27; r++) {
    int ei, ej, ek;
    walk_off(r, 1, &ei, &ej, &ek);
    const int di = ci + ei, dj = cj + ej, dk = ck + ek;
    int32_t ln[3];
    const int32_t key = key_at(c.nx, p.cs, di, dj, dk, ln);
    for (int32_t k = cell_start[key]; k < cell_start[key + 1]; k++) {
      const float4 r4 = rec[k];
      const double dx = rel1d(r4.x, ln[0], di, 0.0) - px;
      const double dy = rel1d(r4.y, ln[1], dj, 0.0) - py;
      const double dz = rel1d(r4.z, ln[2], dk, 0.0) - pz;
      const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx),
                                            __dmul_rn(dy, dy)),
                                  __dmul_rn(dz, dz));
      if (d2 <= c.rs2) s += 1.0 / sqrt(d2 + c.eps2);
    }
  }
  s_out[t] = s;
}

// block reductions (PT threads)
struct PlainShared {
  int64_t red_l[PT / 64];
  double red_d[PT / 64];
  int32_t cnt;
  int64_t bk[PT / 64];
  double bs[PT / 64];
};

__device__ inline int64_t pblock_min(int64_t v, PlainShared &S) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, (int64_t)__shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) S.red_l[threadIdx.x >> 6] = v;
  __syncthreads();
  int64_t r = S.red_l[0];
  for (int i = 1; i < PT / 64; i++) r = min(r, S.red_l[i]);
  __syncthreads();
  return r;
}
__device__ inline double pblock_sum(double v, PlainShared &S) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) S.red_d[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < PT / 64; i++) r += S.red_d[i];
  __syncthreads();
  return r;
}

// ascending (d2, idx) bitonic sort of n entries padded to np in memory
__device__ void plain_sort(double *d2, int32_t *ix, int32_t np) {
  for (int32_t k = 2; k <= np; k <<= 1) {
    for (int32_t j = k >> 1; j > 0; j >>= 1) {
      for (int32_t i = threadIdx.x; i < np; i += PT) {
        const int32_t p = i ^ j;
        if (p > i) {
          const double a = d2[i], b = d2[p];
          const int32_t ia = ix[i], ib = ix[p];
          const bool gt = a > b || (a == b && ia > ib);
          const bool up = (i & k) == 0;
          if (gt == up) { d2[i] = b; d2[p] = a; ix[i] = ib; ix[p] = ia; }
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  }
}

__global__ void __launch_bounds__(PT)
k_plain_grow(SodCfg c, const PlainHalo *__restrict__ ph, int64_t h,
             const int32_t *__restrict__ cell_start,
             const float4 *__restrict__ rec, const float *__restrict__ pos,
             const float *__restrict__ vel, const int64_t *__restrict__ ids,
             const double *__restrict__ s_in, double *__restrict__ d2buf,
             int32_t *__restrict__ ixbuf, int64_t cap,
             int32_t *__restrict__ err, psod_halo_out out) {
  __shared__ PlainShared S;
  const PlainHalo p = ph[h];
  HaloResult res;
  zero_result(res);
  if (p.seed < 0 || p.ncand == 0) {
    if (threadIdx.x == 0) store_result(out, h, c.nbins, res);
    for (int j = threadIdx.x; j < c.nbins; j += PT)
      out.profile[h * c.nbins + j] = 0;
    return;
  }
  // first strict maximum in the walk order (cell rank, then index)
  double bs = -1.0;
  int64_t bk = INT64_MAX;
  for (int32_t t = threadIdx.x; t < p.ncand; t += PT) {
    int q;
    int32_t slot;
    plain_cand(c, p.cs, cell_start, t, &q, &slot);
    const int64_t key = ((int64_t)q << 32) | (uint32_t)__float_as_int(rec[slot].w);
    const double s = s_in[t];
    if (s > bs || (s == bs && key < bk)) { bs = s; bk = key; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double s2 = __shfl_xor(bs, o, 64);
    const int64_t k2 = __shfl_xor(bk, o, 64);
    if (s2 > bs || (s2 == bs && k2 < bk)) { bs = s2; bk = k2; }
  }
  if ((threadIdx.x & 63) == 0) {
    S.bs[threadIdx.x >> 6] = bs;
    S.bk[threadIdx.x >> 6] = bk;
  }
  __syncthreads();
  bs = S.bs[0];
  bk = S.bk[0];
  for (int i = 1; i < PT / 64; i++)
    if (S.bs[i] > bs || (S.bs[i] == bs && S.bk[i] < bk)) {
      bs = S.bs[i];
      bk = S.bk[i];
    }
  __syncthreads();
  const int32_t centre = (int32_t)(bk & 0xFFFFFFFF);
  res.centre = centre;
  res.centre_id = ids ? ids[centre] : (int64_t)centre;
  if (c.flags & PSOD_F_CENTRE_ONLY) {
    if (threadIdx.x == 0) store_result(out, h, c.nbins, res);
    for (int j = threadIdx.x; j < c.nbins; j += PT)
      out.profile[h * c.nbins + j] = 0;
    return;
  }
  const CellRef cr = ref_of_particle(pos, centre, c.nx);
  double xcs[3];
  centre_scaled(pos, centre, c.nx, xcs);
  int32_t W = p.w0;
  for (;;) {
    if (2 * W + 1 > c.nx) {
      res.status = PSOD_H_UNCLOSED;
      if (threadIdx.x == 0) atomicAdd(&err[1], 1);
      break;
    }
    const int side = 2 * W + 1;
    const int32_t ncell = side * side * side;
    if (threadIdx.x == 0) S.cnt = 0;
    __syncthreads();
    // FP64 squared distances of every particle of the cube
    for (int32_t q = threadIdx.x; q < ncell; q += PT) {
      int di, dj, dk;
      walk_off(q, W, &di, &dj, &dk);
      int32_t lc[3];
      const int32_t key = key_at(c.nx, cr.c, di, dj, dk, lc);
      const int32_t a = cell_start[key], b = cell_start[key + 1];
      if (b == a) continue;
      const int32_t o = atomicAdd(&S.cnt, b - a);
      for (int32_t k = a; k < b; k++) {
        if (o + (k - a) >= cap) break;
        const float4 r4 = rec[k];
        const double dx = rel1d(r4.x, lc[0], di, (double)cr.f[0]);
        const double dy = rel1d(r4.y, lc[1], dj, (double)cr.f[1]);
        const double dz = rel1d(r4.z, lc[2], dk, (double)cr.f[2]);
        d2buf[o + (k - a)] = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx),
                                                 __dmul_rn(dy, dy)),
                                       __dmul_rn(dz, dz));
        ixbuf[o + (k - a)] = __float_as_int(r4.w);
      }
    }
    __syncthreads();
    const int32_t n = S.cnt;
    int32_t np = 1;
    while (np < n) np <<= 1;
    if (np > cap) {
      if (threadIdx.x == 0) atomicAdd(&err[0], 1);
      res.status = 0;
      break;
    }
    for (int32_t i = n + threadIdx.x; i < np; i += PT) {
      d2buf[i] = INFINITY;
      ixbuf[i] = INT32_MAX;
    }
    __threadfence_block();
    __syncthreads();
    plain_sort(d2buf, ixbuf, np);
    // first crossings; r_k must stay below the reach
    bool complete = true;
    int64_t kst[PSOD_MAXDELTA];
    for (int di = 0; di < c.ndelta; di++) {
      int64_t first = INT64_MAX;
      for (int64_t k = c.nmin_count + threadIdx.x; k <= n; k += PT) {
        const double r = sqrt(d2buf[k - 1]);
        if (r >= (double)W || od_k(k, c.mpart, r) <= c.delta[di])
          first = min(first, k * 2 + (r >= (double)W ? 1 : 0));
      }
      first = pblock_min(first, S);
      if (first == INT64_MAX || (first & 1)) {
        complete = false;
        kst[di] = (first == INT64_MAX) ? -1 : -2;
      } else {
        kst[di] = first >> 1;
      }
    }
    if (complete && c.rmax_frac * sqrt(d2buf[kst[0] - 2]) > (double)W)
      complete = false;
    if (!complete) {
      int32_t wn = 2 * W;
      if (kst[0] > 0) {
        const int32_t need = (int32_t)ceil(c.rmax_frac * sqrt(d2buf[kst[0] - 2])) + 1;
        if (need > wn) wn = need;
      }
      W = wn;
      __syncthreads();
      continue;
    }
    res.w = W;
    res.ngather = n;
    for (int di = 0; di < c.ndelta; di++)
      finish_delta(c, di, kst[di], sqrt(d2buf[kst[di] - 2]),
                   sqrt(d2buf[kst[di] - 1]), res);
    const double R0 = res.radius[0];
    // last k in [2, k*0] above delta2
    if (c.delta2 > 0.0) {
      int64_t best = -1;
      for (int64_t k = 2 + threadIdx.x; k <= kst[0]; k += PT)
        if (od_k(k, c.mpart, sqrt(d2buf[k - 1])) > c.delta2)
          best = max(best, k);
      best = -pblock_min(-best, S);
      if (best > 0) {
        res.count2 = (int32_t)best;
        const double r = sqrt(d2buf[best - 1]);
        res.overdens2 = r > 1e-30 ? od_k(best, c.mpart, r) : 0.0;
      }
    }
    for (int j = threadIdx.x; j < c.nbins; j += PT) {
      const double rj = prof_radius(c, R0, j);
      int32_t lo = 0, hi = n;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (sqrt(d2buf[mid]) < rj) lo = mid + 1; else hi = mid;
      }
      out.profile[h * c.nbins + j] = lo;
    }
    // boundary particles, centre of mass of the count[0] closest
    int64_t nbd = 0;
    double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    const int32_t nin = res.count[0];
    const bool flagv = (c.flags & PSOD_F_VELOCITY) && vel;
    for (int32_t k = threadIdx.x; k < n; k += PT) {
      if (fabs(sqrt(d2buf[k]) - R0) <= c.tol_abs) nbd++;
      if (k < nin) {
        const int32_t pj = ixbuf[k];
        for (int d = 0; d < 3; d++) {
          double dd = (double)(pos[3 * (int64_t)pj + d] * (float)c.nx) - xcs[d];
          if (dd > 0.5 * c.nx) dd -= c.nx;
          if (dd < -0.5 * c.nx) dd += c.nx;
          sp[d] += dd;
          if (flagv) sv[d] += (double)vel[3 * (int64_t)pj + d];
        }
      }
    }
    res.nboundary = (int32_t)pblock_sum((double)nbd, S);
    for (int d = 0; d < 3; d++) {
      const double s = pblock_sum(sp[d], S), v = pblock_sum(sv[d], S);
      double cp = (xcs[d] + s / (double)nin) / (double)c.nx;
      cp -= floor(cp);
      res.com_pos[d] = cp;
      res.com_vel[d] = v / (double)nin;
    }
    bool acc = (double)res.count[0] * c.mpart >= c.mmin;
    for (int d = 0; d < 3; d++) acc = acc && xcs[d] > 0.0 && xcs[d] <= c.nx;
    res.status = PSOD_H_CLOSED | (acc ? PSOD_H_ACCEPTED : 0);
    break;
  }
  if (threadIdx.x == 0) store_result(out, h, c.nbins, res);
}

extern "C" int psod_sod_fof_plain(psod_ctx *ctx, const psod_sod_params *p,
                                  const psod_catalogue *cat, int64_t nhalo,
                                  const int32_t *d_seed, const int32_t *d_nmem,
                                  const psod_halo_out *out) {
  PSOD_TRY(psod_use_device(ctx));
  PSOD_TRY(check_sod_args(p, cat, nhalo, d_seed, out));
  if (nhalo <= 0) return PSOD_OK;
  if (!out->profile)
    return psod_set_error(PSOD_E_ARG, "sod: profile output is required");
  const SodCfg c = make_cfg(p, cat->npart);
  hipStream_t st = ctx->stream;
  PlainHalo *ph = (PlainHalo *)psod_scratch(ctx, S_PLAIN,
                                            (size_t)nhalo * sizeof(PlainHalo));
  if (!ph) return PSOD_E_NOMEM;
  k_plain_plan<<<grid_for(nhalo, 256), 256, 0, st>>>(
      c, nhalo, d_seed, d_nmem, cat->pos, cat->cell_start, ph);
  PSOD_LAUNCH_CHECK();
  PlainHalo *hph = (PlainHalo *)malloc((size_t)nhalo * sizeof(PlainHalo));
  hipError_t e = hipMemcpyAsync(hph, ph, (size_t)nhalo * sizeof(PlainHalo),
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    free(hph);
    PSOD_HIP(e);
  }
  int32_t maxc = 1;
  for (int64_t h = 0; h < nhalo; h++) maxc = max(maxc, hph[h].ncand);
  const int64_t cap = (int64_t)1 << 26;
  char *buf = (char *)psod_scratch(ctx, S_PLAIN2,
                                   (size_t)cap * 12 + (size_t)maxc * 8 + 256);
  if (!buf) {
    free(hph);
    return PSOD_E_NOMEM;
  }
  double *d2buf = (double *)buf;
  int32_t *ixbuf = (int32_t *)(buf + (size_t)cap * 8);
  double *sbuf = (double *)(buf + (size_t)cap * 12);
  int32_t *err = (int32_t *)(sbuf + maxc + 8);
  int rc = psod_fill_i32(ctx, err, 2, 0);
  for (int64_t h = 0; h < nhalo && rc == PSOD_OK; h++) {
    const int32_t nc = hph[h].seed >= 0 ? hph[h].ncand : 0;
    if (nc > 0) {
      k_plain_centre<<<grid_for(nc, 256), 256, 0, st>>>(
          c, ph, h, cat->cell_start, (const float4 *)cat->rec, sbuf);
    }
    k_plain_grow<<<1, PT, 0, st>>>(c, ph, h, cat->cell_start,
                                   (const float4 *)cat->rec, cat->pos,
                                   cat->vel, cat->ids, sbuf, d2buf, ixbuf,
                                   cap, err, *out);
    if (hipGetLastError() != hipSuccess)
      rc = psod_set_error(PSOD_E_DEVICE, "plain SOD launch failed");
  }
  free(hph);
  PSOD_TRY(rc);
  int32_t he[2];
  PSOD_HIP(hipMemcpyAsync(he, err, 8, hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipStreamSynchronize(st));
  if (he[0] > 0)
    return psod_set_error(PSOD_E_CAPACITY,
                          "plain sod: %d halos gather more than %lld particles",
                          he[0], (long long)cap);
  if (he[1] > 0)
    return psod_set_error(PSOD_E_UNCLOSED,
                          "plain sod: %d halos do not close", he[1]);
  return PSOD_OK;
}
