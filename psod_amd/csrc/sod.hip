// sod.hip -- a6-a12: batched split-precision SOD of FoF-seeded halos.
//
// Each halo (FoF label) is one work item:
//   seed   = its member with the largest CIC density (cic.hip);
//   centre = argmax over the particles of the 3^3 cells around the seed cell
//            of S(c) = sum over particles within 1 cell of 1/sqrt(d2+eps^2)
//            (main file:1795-1875);
//   growth = first k >= 10 with O(k) <= Delta for each threshold, R = r_{k-1}
//            (main file:2040-2100), last k above Delta2 (:2062-2069), the
//            log profile (:2103-2115), centre of mass inside R.
// Unlike the reference, which gathers a fixed (2*ncut+1)^3 cube around the
// seed, the growth here gathers every particle within a reach W of the
// centre and doubles W until every requested radius fits inside it.
//
// Positions are split: records hold float32 offsets inside 8^3-cell blocks,
// so relative coordinates are (whole cells) + (offset difference) and stay
// exact to float32 resolution of the in-cell offsets.  S is accumulated in
// 2^-24 fixed point, so it does not depend on the order of the neighbours.
//
// Small halos (<= CAP5 particles around the seed, reach <= WMAX_S) run one
// per wavefront from a persistent queue; everything they touch stays in LDS
// and their distances are sorted there.  Large halos split the centre search
// into (halo, candidate cell, 64 candidates) tasks and grow with radial
// histograms (bins uniform in r^3) whose ambiguous bins are resolved exactly
// from the collected particles, one workgroup per halo.
#include "sod_common.h"

#include <cstring>
#include <type_traits>
#include <vector>

int psod_fill_i32(psod_ctx *c, int32_t *p, int64_t n, int32_t v);
int psod_fill_u64(psod_ctx *c, uint64_t *p, int64_t n, uint64_t v);

constexpr int CAP5 = 1024;   // small path: particles in the 5^3 cells
constexpr int CAPG = 2048;   // small path: particles in the growth cube
constexpr int WMAX_S = 4;    // small path: largest reach
constexpr int NCUBE_S = (2 * WMAX_S + 1) * (2 * WMAX_S + 1) * (2 * WMAX_S + 1);
// large path: growth cubes of at most this many particles are sorted whole
// in LDS (PSOD_LIGHT_CAP lowers it; 0 sends every halo through the bins)
constexpr int32_t LIGHT_CAP = 8192;

__device__ inline int32_t ldcg(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
SodCfg make_cfg(const psod_sod_params *p, int64_t npart) {
  SodCfg c;
  memset(&c, 0, sizeof(c));
  c.nx = p->nx;
  c.ndelta = p->ndelta;
  c.nbins = p->nbins;
  c.nmin_count = p->nmin_count > 1 ? p->nmin_count : 2;
  c.flags = p->flags;
  for (int i = 0; i < PSOD_MAXDELTA; i++)
    c.delta[i] = i < p->ndelta ? p->delta[i] : 0.0;
  c.delta2 = p->delta2;
  c.lr0 = log(p->rmin_frac);
  c.lr1 = log(p->rmax_frac);
  c.rmin_frac = p->rmin_frac;
  c.rmax_frac = p->rmax_frac;
  c.mmin = p->mmin;
  c.mpart = p->mpart;
  c.rs2 = p->rsearch * p->rsearch;
  c.eps2 = p->eps * p->eps;
  c.rs2f = (float)c.rs2;
  c.eps2f = (float)c.eps2;
  c.tol_abs = p->tol_abs;
  c.nbar = (double)npart / ((double)p->nx * p->nx * p->nx);
  c.giant_min = INT64_MAX;  // no staged growth unless the pass sets it
  c.light_cap = LIGHT_CAP;
  if (const char *e = getenv("PSOD_LIGHT_CAP")) {
    const long v = atol(e);
    if (v >= 0 && v < LIGHT_CAP) c.light_cap = (int32_t)v;
  }
  return c;
}

int check_sod_args(const psod_sod_params *p, const psod_catalogue *cat,
                   int64_t nhalo, const int32_t *seed,
                   const psod_halo_out *out) {
  if (!p || !cat || !out || (nhalo > 0 && !seed))
    return psod_set_error(PSOD_E_ARG, "sod: null argument");
  if (p->nx != cat->nx || p->nx < 8 || p->nx > 1024 || (p->nx & (p->nx - 1)))
    return psod_set_error(PSOD_E_ARG, "sod: nx=%d (catalogue %d)", p->nx,
                          cat->nx);
  if (p->ndelta < 1 || p->ndelta > PSOD_MAXDELTA)
    return psod_set_error(PSOD_E_ARG, "sod: ndelta=%d", p->ndelta);
  for (int i = 0; i < p->ndelta; i++)
    if (!(p->delta[i] > 0.0))
      return psod_set_error(PSOD_E_ARG, "sod: delta[%d]=%g", i, p->delta[i]);
  if (p->ndelta > 1)
    for (int i = 1; i < p->ndelta; i++)
      if (p->delta[i] < p->delta[0])
        return psod_set_error(PSOD_E_ARG,
                              "sod: delta[0] must be the smallest threshold");
  if (p->nbins < 2 || p->nbins > 256 || !(p->rmin_frac > 0.0) ||
      !(p->rmax_frac >= p->rmin_frac) || !(p->rmax_frac >= 1.0))
    return psod_set_error(PSOD_E_ARG, "sod: profile nbins=%d [%g, %g]",
                          p->nbins, p->rmin_frac, p->rmax_frac);
  if (!(p->mpart > 0.0) || p->nmin_count < 2)
    return psod_set_error(PSOD_E_ARG, "sod: mpart=%g nmin_count=%d", p->mpart,
                          p->nmin_count);
  if (!cat->pos || !cat->cell_start || !cat->rec || cat->npart <= 0)
    return psod_set_error(PSOD_E_ARG, "sod: catalogue is incomplete");
  if ((p->flags & PSOD_F_VELOCITY) && !cat->vel)
    return psod_set_error(PSOD_E_ARG, "sod: velocity requested without vel");
  return PSOD_OK;
}

// ---------------------------------------------------------------------------
// plan: seed cell, particles around it, reach estimate, small/large
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_plan(SodCfg c, int64_t nhalo, const int32_t *__restrict__ seed,
       const int32_t *__restrict__ nmem, const float *__restrict__ pos,
       const int32_t *__restrict__ cell_start, HaloPlan *__restrict__ plan,
       int32_t *__restrict__ small, int32_t *__restrict__ large,
       int32_t *__restrict__ counters, psod_halo_out out) {
  // one wave per halo: the lanes count the 125 cells around the seed
  const int l = threadIdx.x & 63;
  const int64_t h = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (h >= nhalo) return;  // uniform per wave
  HaloPlan pl;
  pl.seed = seed[h];
  pl.nmem = nmem ? nmem[h] : 0;
  pl.centre = -1;
  pl.ntask = 0;
  pl.task0 = 0;
  pl.pad = 0;
  if (pl.seed < 0) {
    pl.klass = -1;
    pl.w0 = 0;
    pl.n5 = 0;
    pl.cs[0] = pl.cs[1] = pl.cs[2] = 0;
    for (int j = l; j < c.nbins; j += 64) out.profile[h * c.nbins + j] = 0;
    if (l == 0) {
      plan[h] = pl;
      HaloResult r;
      zero_result(r);
      store_result(out, h, c.nbins, r);
    }
    return;
  }
  const CellRef sr = ref_of_particle(pos, pl.seed, c.nx);
  int32_t n5 = 0;
  for (int q = l; q < 125; q += 64) {
    int di, dj, dk;
    walk_off(q, 2, &di, &dj, &dk);
    int32_t lc[3];
    const int32_t key = key_at(c.nx, sr.c, di, dj, dk, lc);
    n5 += cell_start[key + 1] - cell_start[key];
  }
  n5 = wave_sum(n5);
  if (l != 0) return;
  for (int d = 0; d < 3; d++) pl.cs[d] = sr.c[d];
  pl.n5 = n5;
  pl.w0 = reach_estimate(c, pl.nmem);
  pl.klass = (n5 <= c.small_max && pl.w0 <= WMAX_S) ? 0 : 1;
  plan[h] = pl;
}

// the small (class 0) and large (class 1) lists, one atomic per wave and
// list
__global__ void __launch_bounds__(256)
k_plan_lists(int64_t nhalo, const HaloPlan *__restrict__ plan,
             int32_t *__restrict__ small, int32_t *__restrict__ large,
             int32_t *__restrict__ counters) {
  const int l = threadIdx.x & 63;
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = h < nhalo ? plan[h].klass : -1;
  const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
  const uint64_t m0 = __ballot(k == 0), m1 = __ballot(k == 1);
  int32_t b0 = 0, b1 = 0;
  if (l == 0) {
    if (m0) b0 = atomicAdd(&counters[0], __popcll(m0));
    if (m1) b1 = atomicAdd(&counters[1], __popcll(m1));
  }
  b0 = __shfl(b0, 0, 64);
  b1 = __shfl(b1, 0, 64);
  if (k == 0) small[b0 + __popcll(m0 & below)] = (int32_t)h;
  if (k == 1) large[b1 + __popcll(m1 & below)] = (int32_t)h;
}

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
// exclusive prefix of v over the 64 lanes plus the wave total
__device__ inline int32_t wave_excl(int32_t v, int32_t *tot) {
  const int32_t inc = wave_incl_scan(v);
  *tot = __shfl(inc, 63, 64);
  return inc - v;
}

__device__ inline bool better(uint64_t a, uint64_t ka, uint64_t b,
                              uint64_t kb) {
  return a > b || (a == b && ka < kb);
}

__device__ inline void wave_argmax(uint64_t &acc, uint64_t &key) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a2 = __shfl_xor(acc, o, 64);
    const uint64_t k2 = __shfl_xor(key, o, 64);
    if (better(a2, k2, acc, key)) { acc = a2; key = k2; }
  }
}

// Bitonic sort of np (a power of two) u64 keys in LDS by the NT threads of
// the workgroup (NT a multiple of 64): thread t orders the pairs t, t + NT,
// ...; pair t of step j is (i, i + j), i being t with a zero bit inserted
// at j.  For j <= 32 a wave's pairs lie in its own blocks of 128 keys
// (128 w + 2 NT m), so those steps need a wave barrier only; a workgroup
// barrier comes before and after every step of larger j, and at the end.
// (The form that ran every key index and left the upper half idle took
// twice the iterations.)
template <int NT>
__device__ inline void lds_bitonic(uint64_t *K, int32_t np) {
  const int32_t half = np >> 1;
  for (int32_t k = 2; k <= np; k <<= 1) {
    for (int32_t j = k >> 1; j > 0; j >>= 1) {
      for (int32_t t = threadIdx.x; t < half; t += NT) {
        const int32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1)), p = i | j;
        const uint64_t x = K[i], y = K[p];
        if ((x > y) == ((i & k) == 0)) {
          K[i] = y;
          K[p] = x;
        }
      }
      const int32_t jn = j > 1 ? j >> 1 : k;  // the next step's j
      if (NT > 64 && (j > 32 || jn > 32)) {
        __syncthreads();
      } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  }
  __syncthreads();
}

__device__ inline int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

__device__ inline uint64_t d2key(float d2, int32_t idx) {
  return ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)idx;
}
__device__ inline float key_d2(uint64_t k) {
  return __uint_as_float((uint32_t)(k >> 32));
}
__device__ inline int32_t key_idx(uint64_t k) {
  return (int32_t)(uint32_t)(k & 0xFFFFFFFFull);
}

// ---------------------------------------------------------------------------
// small halos: one wavefront (= one workgroup of 64) per halo
// ---------------------------------------------------------------------------
struct SmallA {
  float4 nb[CAP5];        // relative to the seed cell origin, w = index bits
  int8_t rank[CAP5];      // candidate cell rank 0..26, -1 otherwise
  int16_t cidx[CAP5];     // compacted candidate list
  int32_t pre[128];       // prefix over the 125 cells
  int32_t st[128];        // first slot of each of the 125 cells
};
struct SmallB {
  uint64_t key[CAPG];
  int32_t pre[NCUBE_S + 1];
  int32_t st[NCUBE_S];
};
union SmallLds {
  SmallA a;
  SmallB b;
};

__device__ inline int bsearch_last_le(const int32_t *pre, int n, int32_t j) {
  int lo = 0, hi = n;  // last c in [0, n) with pre[c] <= j
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= j) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ void small_growth_and_output(const SodCfg &c, SmallLds &S,
                                        const int32_t *cell_start,
                                        const float4 *rec, const float *pos,
                                        const float *vel, const int64_t *ids,
                                        int32_t centre, int32_t w0,
                                        psod_halo_out out, int64_t h,
                                        HaloPlan *plan, int32_t *large,
                                        int32_t *counters);

// At most 176 vector registers as allocated (blocks of 8): two of these
// waves then leave room for two waves of the centre-search bound on the
// same SIMD.  Its key partition pushed the allocation to 184 (bound 63.4 ->
// 67.4 ms beside it at 1024^3); held to 176, the bound took 62.0 ms.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_num_vgpr(168)))
k_sod_small(SodCfg c, const int32_t *__restrict__ cell_start,
            const float4 *__restrict__ rec, const float *__restrict__ pos,
            const float *__restrict__ vel, const int64_t *__restrict__ ids,
            HaloPlan *__restrict__ plan, const int32_t *__restrict__ small,
            int32_t *__restrict__ large, int32_t *__restrict__ counters,
            psod_halo_out out) {
  __shared__ SmallLds S;
  __shared__ int32_t s_item;
  const int l = threadIdx.x;
  const int32_t nsmall = counters[0];
  for (;;) {
    if (l == 0) s_item = atomicAdd(&counters[2], 1);
    __syncthreads();
    const int32_t it = s_item;
    __syncthreads();
    if (it >= nsmall) break;
    const int64_t h = small[it];
    HaloPlan pl = plan[h];
    // ---- the 125 cells around the seed cell ---------------------------
    int32_t cnt0 = 0, cnt1 = 0, st0 = 0, st1 = 0;
    {
      int di, dj, dk;
      int32_t lc[3];
      walk_off(l, 2, &di, &dj, &dk);
      int32_t key = key_at(c.nx, pl.cs, di, dj, dk, lc);
      st0 = cell_start[key];
      cnt0 = cell_start[key + 1] - st0;
      if (l + 64 < 125) {
        walk_off(l + 64, 2, &di, &dj, &dk);
        key = key_at(c.nx, pl.cs, di, dj, dk, lc);
        st1 = cell_start[key];
        cnt1 = cell_start[key + 1] - st1;
      }
    }
    int32_t t0, t1;
    const int32_t e0 = wave_excl(cnt0, &t0);
    const int32_t e1 = wave_excl(cnt1, &t1);
    S.a.pre[l] = e0;
    S.a.st[l] = st0;
    if (l + 64 < 125) {
      S.a.pre[l + 64] = t0 + e1;
      S.a.st[l + 64] = st1;
    }
    const int32_t n5 = t0 + t1;
    if (l == 0) S.a.pre[125] = n5;
    __syncthreads();
    // n5 <= CAP5 was checked by the plan
    for (int32_t j = l; j < n5; j += 64) {
      const int q = bsearch_last_le(S.a.pre, 125, j);
      const float4 r = rec[S.a.st[q] + (j - S.a.pre[q])];
      int di, dj, dk;
      walk_off(q, 2, &di, &dj, &dk);
      int32_t lc[3];
      key_at(c.nx, pl.cs, di, dj, dk, lc);
      float4 v;
      v.x = cell_rel(r.x, lc[0], di);
      v.y = cell_rel(r.y, lc[1], dj);
      v.z = cell_rel(r.z, lc[2], dk);
      v.w = r.w;
      S.a.nb[j] = v;
      const bool inner = di >= -1 && di <= 1 && dj >= -1 && dj <= 1 &&
                         dk >= -1 && dk <= 1;
      S.a.rank[j] = inner ? (int8_t)((di + 1) * 9 + (dj + 1) * 3 + (dk + 1))
                          : (int8_t)-1;
    }
    __syncthreads();
    // compact the candidates
    int32_t ncand = 0;
    for (int32_t base = 0; base < n5; base += 64) {
      const int32_t j = base + l;
      const bool isc = j < n5 && S.a.rank[j] >= 0;
      const uint64_t m = ballot64(isc);
      const uint64_t lower = (l == 0) ? 0ull : (~0ull >> (64 - l));
      if (isc) S.a.cidx[ncand + __popcll(m & lower)] = (int16_t)j;
      ncand += __popcll(m);
    }
    __syncthreads();
    // ---- centre search -------------------------------------------------
    uint64_t bacc = 0, bkey = ~0ull;
    // every lane takes part in each round (the neighbour loop is uniform:
    // scalar bounds, broadcast LDS reads, four in flight)
    const int32_t ncu = __builtin_amdgcn_readfirstlane(ncand);
    const int32_t n5u = __builtin_amdgcn_readfirstlane(n5);
    for (int32_t cb = 0; cb < ncu; cb += 64) {
      const bool have = cb + l < ncu;
      const int j = have ? S.a.cidx[cb + l] : 0;
      const float4 p = S.a.nb[j];
      uint64_t acc = 0;
      // u32 partial sums of at most 255 terms of at most 2^24 (256 terms
      // at d = 0, duplicated particles, would reach 2^32)
      for (int32_t t0 = 0; t0 < n5u; t0 += 255) {
        const int32_t t1 = min(n5u, t0 + 255);
        uint32_t part = 0;
#pragma unroll 4
        for (int32_t t = t0; t < t1; t++) {
          const float4 q = S.a.nb[t];
          const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
          const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const uint32_t term = sterm_fix(d2, c.eps2f);
          part += d2 <= c.rs2f ? term : 0u;
        }
        acc += part;
      }
      if (have) {
        const uint64_t key = ((uint64_t)(uint32_t)S.a.rank[j] << 32) |
                             (uint32_t)__float_as_int(p.w);
        if (better(acc, key, bacc, bkey)) { bacc = acc; bkey = key; }
      }
    }
    wave_argmax(bacc, bkey);
    const int32_t centre = key_idx(bkey);
    __syncthreads();
    small_growth_and_output(c, S, cell_start, rec, pos, vel, ids, centre,
                            pl.w0, out, h, plan, large, counters);
    __syncthreads();
  }
}

// growth of a small halo around `centre`; defers to the large path when the
// cube does not fit
__device__ void small_growth_and_output(const SodCfg &c, SmallLds &S,
                                        const int32_t *cell_start,
                                        const float4 *rec, const float *pos,
                                        const float *vel, const int64_t *ids,
                                        int32_t centre, int32_t w0,
                                        psod_halo_out out, int64_t h,
                                        HaloPlan *plan, int32_t *large,
                                        int32_t *counters) {
  const int l = threadIdx.x;
  const CellRef cr = ref_of_particle(pos, centre, c.nx);
  int32_t W = w0;
  HaloResult res;
  zero_result(res);
  res.centre = centre;
  res.centre_id = ids ? ids[centre] : (int64_t)centre;
  for (;;) {
    if (W > WMAX_S || 2 * W + 1 > c.nx) {
      if (l == 0) {
        plan[h].centre = centre;
        plan[h].w0 = W;
        large[atomicAdd(&counters[1], 1)] = (int32_t)h;
      }
      return;
    }
    const int side = 2 * W + 1, ncell = side * side * side;
    // ranges of the cube cells
    int32_t carry = 0;
    for (int base = 0; base < ncell; base += 64) {
      const int q = base + l;
      int32_t cnt = 0, st = 0;
      if (q < ncell) {
        int di, dj, dk;
        walk_off(q, W, &di, &dj, &dk);
        int32_t lc[3];
        const int32_t key = key_at(c.nx, cr.c, di, dj, dk, lc);
        st = cell_start[key];
        cnt = cell_start[key + 1] - st;
      }
      int32_t tot;
      const int32_t ex = wave_excl(cnt, &tot);
      if (q < ncell) {
        S.b.pre[q] = carry + ex;
        S.b.st[q] = st;
      }
      carry += tot;
    }
    const int32_t ncube = carry;
    if (l == 0) S.b.pre[ncell] = ncube;
    __syncthreads();
    if (ncube > CAPG) {
      __syncthreads();
      W = WMAX_S + 1;  // defer
      continue;
    }
    // Keys of r < W (the only ones that can be ranked before a crossing)
    // go to the front, the others to the back of the key array; only the
    // front is sorted when the back leaves room for its padding (the cube's
    // corners hold half of its particles or more).  Otherwise the back
    // moves down behind the front and every key is sorted.
    const double Wd = (double)W;
    int32_t ninr = 0, nfar = 0;
    const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
    for (int32_t base = 0; base < ncube; base += 64) {
      const int32_t j = base + l;
      uint64_t key = 0;
      bool inner = false;
      if (j < ncube) {
        const int q = bsearch_last_le(S.b.pre, ncell, j);
        const float4 r = rec[S.b.st[q] + (j - S.b.pre[q])];
        int di, dj, dk;
        walk_off(q, W, &di, &dj, &dk);
        int32_t lc[3];
        key_at(c.nx, cr.c, di, dj, dk, lc);
        const float dx = rel1(r.x, lc[0], di, cr.f[0]);
        const float dy = rel1(r.y, lc[1], dj, cr.f[1]);
        const float dz = rel1(r.z, lc[2], dk, cr.f[2]);
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        key = d2key(d2, __float_as_int(r.w));
        inner = (double)sqrtf(d2) < Wd;
      }
      const uint64_t mi = ballot64(j < ncube && inner);
      const uint64_t mf = ballot64(j < ncube && !inner);
      if (j < ncube) {
        if (inner) S.b.key[ninr + __popcll(mi & below)] = key;
        else S.b.key[CAPG - 1 - (nfar + __popcll(mf & below))] = key;
      }
      ninr += __popcll(mi);
      nfar += __popcll(mf);
    }
    __syncthreads();
    int32_t npad = next_pow2(ninr < 64 ? 64 : ninr);
    const bool part = npad + nfar <= CAPG;  // the far keys stay at the back
    if (!part) {
      // the far keys down to [ninr, ncube), in ascending chunks (each read
      // before it is written; destinations lie below their sources)
      for (int32_t b0 = 0; b0 < nfar; b0 += 64) {
        const int32_t t = b0 + l;
        uint64_t v = 0;
        if (t < nfar) v = S.b.key[CAPG - nfar + t];
        __syncthreads();
        if (t < nfar) S.b.key[ninr + t] = v;
        __syncthreads();
      }
      npad = next_pow2(ncube < 64 ? 64 : ncube);
    }
    const int32_t nsort = part ? ninr : ncube;  // keys in sorted order
    for (int32_t j = nsort + l; j < npad; j += 64) S.b.key[j] = ~0ull;
    __syncthreads();
    lds_bitonic<64>(S.b.key, npad);
    // ---- first crossing of every threshold -----------------------------
    bool complete = true;
    int64_t kst[PSOD_MAXDELTA];
    for (int di = 0; di < c.ndelta; di++) {
      kst[di] = -1;
      for (int32_t base = c.nmin_count; base <= ncube && kst[di] < 0;
           base += 64) {
        const int64_t k = base + l;
        bool hit = false, over = false;
        if (k <= nsort) {
          const double r = (double)sqrtf(key_d2(S.b.key[k - 1]));
          over = r >= Wd;
          hit = over || od_k(k, c.mpart, r) <= c.delta[di];
        } else if (k <= ncube) {
          hit = over = true;  // a far key: r >= W
        }
        const uint64_t m = ballot64(hit);
        if (m) {
          const int first = __ffsll((unsigned long long)m) - 1;
          const bool fo = __shfl(over ? 1 : 0, first, 64);
          kst[di] = fo ? -2 : base + first;
        }
      }
      if (kst[di] < 0) complete = false;
    }
    double R0 = 0.0;
    if (complete) {
      R0 = (double)sqrtf(key_d2(S.b.key[kst[0] - 2]));
      if (c.rmax_frac * R0 > (double)W) complete = false;
    }
    if (!complete) {
      int32_t wn = 2 * W;
      if (kst[0] > 0) {
        const int32_t need =
            (int32_t)ceil(c.rmax_frac *
                          (double)sqrtf(key_d2(S.b.key[kst[0] - 2]))) + 1;
        if (need > wn) wn = need;
      }
      W = wn;
      __syncthreads();
      continue;
    }
    res.w = W;
    res.ngather = ncube;
    double xcs[3];
    centre_scaled(pos, centre, c.nx, xcs);
    for (int di = 0; di < c.ndelta; di++) {
      const int64_t ks = kst[di];
      finish_delta(c, di, ks,
                   ref_dist(pos, key_idx(S.b.key[ks - 2]), xcs, c.nx),
                   ref_dist(pos, key_idx(S.b.key[ks - 1]), xcs, c.nx), res);
    }
    R0 = res.radius[0];
    // ---- second threshold: last k in [2, k*0] above delta2 ----------------
    if (c.delta2 > 0.0) {
      int64_t best = 0;
      for (int32_t base = 2; base <= kst[0]; base += 64) {
        const int64_t k = base + l;
        bool hit = false;
        if (k <= kst[0]) {
          const double r = (double)sqrtf(key_d2(S.b.key[k - 1]));
          hit = od_k(k, c.mpart, r) > c.delta2;
        }
        const uint64_t m = ballot64(hit);
        if (m) best = base + 63 - __clzll((unsigned long long)m);
      }
      res.count2 = (int32_t)best;
      if (best > 0) {
        const double r = ref_dist(pos, key_idx(S.b.key[best - 1]), xcs, c.nx);
        res.overdens2 = r > 1e-30 ? od_k(best, c.mpart, r) : 0.0;
      }
    }
    // ---- profile --------------------------------------------------------
    for (int j = l; j < c.nbins; j += 64) {
      const double rj = prof_radius(c, R0, j);
      int lo = 0, hi = nsort;  // first index with r >= rj
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)sqrtf(key_d2(S.b.key[mid])) < rj) lo = mid + 1;
        else hi = mid;
      }
      if (part && rj > Wd)  // far keys (r >= W) below a radius beyond W
        for (int32_t t = CAPG - nfar; t < CAPG; t++)
          if ((double)sqrtf(key_d2(S.b.key[t])) < rj) lo++;
      out.profile[h * c.nbins + j] = lo;
    }
    // ---- boundary particles, centre of mass of the count[0] closest -------
    int32_t nbd = 0;
    double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    const int32_t nin = res.count[0];
    const bool flagv = (c.flags & PSOD_F_VELOCITY) && vel;
    const double xc[3] = {(double)cr.c[0] + (double)cr.f[0],
                          (double)cr.c[1] + (double)cr.f[1],
                          (double)cr.c[2] + (double)cr.f[2]};
    for (int32_t k = l; k < ncube; k += 64) {
      // rank order for the sorted keys, then the far keys at the back
      const uint64_t key =
          S.b.key[k < nsort ? k : CAPG - nfar + (k - nsort)];
      const double r = (double)sqrtf(key_d2(key));
      if (fabs(r - R0) <= c.tol_abs) nbd++;
      if (k < nin) {
        const int32_t pj = key_idx(key);
        for (int d = 0; d < 3; d++) {
          int32_t cj;
          float fj;
          split_coord(pos[3 * (int64_t)pj + d], c.nx, &cj, &fj);
          double dd = ((double)cj + (double)fj) - xc[d];
          if (dd > 0.5 * c.nx) dd -= c.nx;
          if (dd < -0.5 * c.nx) dd += c.nx;
          sp[d] += dd;
          if (flagv) sv[d] += (double)vel[3 * (int64_t)pj + d];
        }
      }
    }
    nbd = wave_sum(nbd);
    for (int d = 0; d < 3; d++) {
      sp[d] = wave_sum(sp[d]);
      sv[d] = wave_sum(sv[d]);
      double cp = (xc[d] + sp[d] / (double)nin) / (double)c.nx;
      cp -= floor(cp);
      res.com_pos[d] = cp;
      res.com_vel[d] = sv[d] / (double)nin;
    }
    res.nboundary = nbd;
    bool acc = (double)res.count[0] * c.mpart >= c.mmin;
    for (int d = 0; d < 3; d++) acc = acc && xc[d] > 0.0 && xc[d] <= c.nx;
    res.status = PSOD_H_CLOSED | (acc ? PSOD_H_ACCEPTED : 0);
    if (l == 0) store_result(out, h, c.nbins, res);
    return;
  }
}

// ---------------------------------------------------------------------------
// large halos: centre search tasks
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_task_count(SodCfg c, const int32_t *__restrict__ cell_start,
             HaloPlan *__restrict__ plan, const int32_t *__restrict__ large,
             int32_t nlarge, const int32_t *__restrict__ dn,
             int32_t *__restrict__ ntask) {
  // nlarge bounds the launch; with dn the halos are the first *dn (entries
  // up to nlarge beyond them count no tasks, for the scan)
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nlarge) return;
  if (dn && e >= *dn) {
    ntask[e] = 0;
    return;
  }
  const HaloPlan pl = plan[large[e]];
  int32_t nt = 0;
  if (pl.centre < 0) {
    for (int q = 0; q < 27; q++) {
      int di, dj, dk;
      walk_off(q, 1, &di, &dj, &dk);
      int32_t lc[3];
      const int32_t key = key_at(c.nx, pl.cs, di, dj, dk, lc);
      const int32_t n = cell_start[key + 1] - cell_start[key];
      nt += (n + 63) / 64;
    }
  }
  ntask[e] = nt;
}

__global__ void __launch_bounds__(256)
k_task_emit(SodCfg c, const int32_t *__restrict__ cell_start,
            HaloPlan *__restrict__ plan, const int32_t *__restrict__ large,
            int32_t nlarge, const int32_t *__restrict__ dn,
            const int32_t *__restrict__ ntask,
            const int32_t *__restrict__ task0,
            const int32_t *__restrict__ overflow, int4 *__restrict__ tasks) {
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nlarge || (dn && e >= *dn)) return;
  const int32_t h = large[e];
  // tasks beyond the capacity of a captured pass: none are written
  const bool none = overflow && *overflow;
  plan[h].ntask = none ? 0 : ntask[e];
  plan[h].task0 = none ? 0 : task0[e];
  if (none || ntask[e] == 0) return;
  const HaloPlan pl = plan[h];
  int32_t t = task0[e];
  for (int q = 0; q < 27; q++) {
    int di, dj, dk;
    walk_off(q, 1, &di, &dj, &dk);
    int32_t lc[3];
    const int32_t key = key_at(c.nx, pl.cs, di, dj, dk, lc);
    const int32_t a = cell_start[key], b = cell_start[key + 1];
    for (int32_t s = a; s < b; s += 64)
      tasks[t++] = make_int4(e, q, s, min(b, s + 64) - s);
  }
}

// Minimum / maximum over the 64 lanes of a wave, returned wave-uniform.  All
// lanes must be active.  The floats are mapped to integers of the same
// order (negative values: magnitude bits flipped), so each step is one
// integer min/max with a DPP operand (the float forms also quiet NaNs,
// an extra instruction per step).  Row shifts by 1, 2, 4, 8 leave each
// 16-lane row's result in its lane 15; broadcasts of lanes 15 and 31 fold
// the rows into lane 63 (register moves between lanes, no LDS round trips).
__device__ inline int32_t f2ord(float f) {
  const int32_t i = __float_as_int(f);
  return i ^ ((i >> 31) & 0x7fffffff);
}
__device__ inline float ord2f(int32_t i) {
  return __int_as_float(i ^ ((i >> 31) & 0x7fffffff));
}
template <bool MAX>
__device__ inline float wave_extf(float x) {
  const int id = MAX ? INT32_MIN : INT32_MAX;
  auto op = [](int32_t a, int32_t b) { return MAX ? max(a, b) : min(a, b); };
  int32_t v = f2ord(x);
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x111, 0xf, 0xf, false));
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x112, 0xf, 0xf, false));
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x114, 0xf, 0xf, false));
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x118, 0xf, 0xf, false));
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));
  v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));
  return ord2f(__builtin_amdgcn_readlane(v, 63));
}
__device__ inline float wave_minf(float v) { return wave_extf<false>(v); }
__device__ inline float wave_maxf(float v) { return wave_extf<true>(v); }

// squared distance between two boxes (0 when they overlap)
__device__ inline float box_gap2(const float *a0, const float *a1,
                                 const float *b0, const float *b1) {
  float s = 0.0f;
  for (int d = 0; d < 3; d++) {
    const float g = fmaxf(0.0f, fmaxf(b0[d] - a1[d], a0[d] - b1[d]));
    s += g * g;
  }
  return s;
}

// neighbour cells of a candidate cell, nearest first (walk_off indices:
// the cell itself, 6 faces, 12 edges, 8 corners)
__constant__ int8_t kCellOrder[27] = {13, 4,  10, 12, 14, 16, 22, 1,  3,
                                      5,  7,  9,  11, 15, 17, 19, 21, 23,
                                      25, 0,  2,  6,  8,  18, 20, 24, 26};

// Halos with fewer than SodCfg::bound_min particles in the 5^3 cells around
// the seed (a tunable, 1024 by default) skip the bound: bounding costs about
// as much as their exact sums.  With the items abandoned cell by cell,
// 1024 gave 140.6 ms of SOD at c3 against 144.2 (2048), 141.6 (512) and
// 149.8 ms (4096).

// Upper bound of the term of any pair at squared distance >= g2 within the
// cutoff.  The float d2 of a pair is at least g2 up to a few ulps, hence the
// 2e-5 shrink; rounding with +8.5 instead of +0.5 covers non-monotone
// rounding of the reciprocal square root.
__device__ inline uint32_t sterm_bound(float g2, float eps2) {
  return (uint32_t)(__builtin_amdgcn_rsqf(fmaf(g2, 1.0f - 2e-5f, eps2)) * SFIX +
                    8.5f);
}

// Next block of work of a wave from a shared counter (the first active
// lane takes it; every lane gets it): waves that finish early take more.
__device__ inline int32_t wave_next(int32_t *ctr) {
  int32_t v = 0;
  if ((threadIdx.x & 63) == 0) v = atomicAdd(ctr, 1);
  return __builtin_amdgcn_readfirstlane(v);
}

// Sums over the wave of u[g] for 8 groups, transposed: lane l ends with the
// sum of group l >> 3.  Each exchange step hands half of the values still
// held to the partner lane (10 64-bit exchanges instead of 8 x 6).
__device__ inline uint64_t group_reduce8(const uint64_t (&u)[8], int l) {
  const bool h5 = (l & 32) != 0, h4 = (l & 16) != 0, h3 = (l & 8) != 0;
  uint64_t a[4], b[2];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint64_t keep = h5 ? u[i + 4] : u[i], send = h5 ? u[i] : u[i + 4];
    a[i] = keep + __shfl_xor(send, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const uint64_t keep = h4 ? a[i + 2] : a[i], send = h4 ? a[i] : a[i + 2];
    b[i] = keep + __shfl_xor(send, 16, 64);
  }
  uint64_t v = (h3 ? b[1] : b[0]) + __shfl_xor(h3 ? b[0] : b[1], 8, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 1, 64);
  return v;
}

// Phase 0 of the centre search, one wave per task: each candidate's S is at
// most the sum, over the neighbours of the 27 cells, of the term at the
// neighbour's distance to the box of the candidate's 16 (candidates
// 16b..16b+15, Morton-compact).  Boxes of 8 bound tighter but cost twice as
// many terms; with the items taken most promising first, 16 is faster
// (c3: 8 -> 236 ms, 16 -> 218, 32 -> 235, 64 -> 263 ms of SOD).  The bound
// of box b serves both groups of 8 inside it: stores the bound per (cell,
// group) in cub[t*27*8 + q*8 + g], the per-group totals in
// cub[(nt*27 + t)*8 + g], the task's bound in ubt[t] and each halo's
// largest task bound in hpil.
__global__ void __launch_bounds__(256)
k_centre_bound(SodCfg c, const int32_t *__restrict__ cell_start,
               const float4 *__restrict__ rec, const HaloPlan *__restrict__ plan,
               const int32_t *__restrict__ large, const int4 *__restrict__ tasks,
               const int32_t *__restrict__ ntask_total, int32_t tcap,
               unsigned long long *__restrict__ hpil,
               unsigned long long *__restrict__ ubt,
               unsigned long long *__restrict__ cub, int32_t *__restrict__ wq) {
  // the wave index is uniform: task fields and everything derived from
  // them go to scalar registers
  const int l = threadIdx.x & 63,
            w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int32_t nt = *ntask_total;
  const float cut = c.rs2f * 1.0001f + 1e-6f;  // conservative chunk culling
  // held in a vector register: the term's fused multiply-add takes it
  // without a move per term (71.8 -> 67.9 ms at 1024^3 with the maxima below)
  const float eps2v = __shfl(c.eps2f, 0, 64);
  for (;;) {
    const int32_t t0 = wave_next(wq) * 2;
    if (t0 >= nt) break;
    const int32_t t1 = min(nt, t0 + 2);
    for (int32_t t = t0; t < t1; t++) {
      const int4 tk = tasks[t];
      const HaloPlan pl = plan[large[tk.x]];
      if (pl.n5 < c.bound_min) {
        // few particles around the seed: every task of the halo is evaluated
        // exactly by the pilots (an all-ones bound makes each one a pilot)
        if (l == 0) {
          ubt[t] = ~0ull;
          atomicMax(&hpil[tk.x], ~0ull);
        }
        continue;
      }
      int ci, cj, ck;
      walk_off(tk.y, 1, &ci, &cj, &ck);
      int32_t lc[3];
      key_at(c.nx, pl.cs, ci, cj, ck, lc);
      float px = 0.f, py = 0.f, pz = 0.f;
      const bool have = l < tk.w;
      if (have) {
        const float4 r = rec[tk.z + l];
        px = cell_rel(r.x, lc[0], ci);
        py = cell_rel(r.y, lc[1], cj);
        pz = cell_rel(r.z, lc[2], ck);
      }
      float cb0[3], cb1[3];
      cb0[0] = wave_minf(have ? px : INFINITY);
      cb0[1] = wave_minf(have ? py : INFINITY);
      cb0[2] = wave_minf(have ? pz : INFINITY);
      cb1[0] = wave_maxf(have ? px : -INFINITY);
      cb1[1] = wave_maxf(have ? py : -INFINITY);
      cb1[2] = wave_maxf(have ? pz : -INFINITY);
      // boxes of the 4 sets of 16 candidates (uniform)
      float g0x = have ? px : INFINITY, g0y = have ? py : INFINITY,
            g0z = have ? pz : INFINITY;
      float g1x = have ? px : -INFINITY, g1y = have ? py : -INFINITY,
            g1z = have ? pz : -INFINITY;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        g0x = fminf(g0x, __shfl_xor(g0x, o, 64));
        g0y = fminf(g0y, __shfl_xor(g0y, o, 64));
        g0z = fminf(g0z, __shfl_xor(g0z, o, 64));
        g1x = fmaxf(g1x, __shfl_xor(g1x, o, 64));
        g1y = fmaxf(g1y, __shfl_xor(g1y, o, 64));
        g1z = fmaxf(g1z, __shfl_xor(g1z, o, 64));
      }
      // minima in scalar registers, maxima in vector registers: the clamp
      // (v_med3) takes one scalar operand, so neither needs a move per term
      const float hb[3] = {__shfl(cb1[0], 0, 64), __shfl(cb1[1], 0, 64),
                           __shfl(cb1[2], 0, 64)};
      float B[4][6];
#pragma unroll
      for (int g = 0; g < 4; g++) {
        B[g][0] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g0x), 16 * g));
        B[g][1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g0y), 16 * g));
        B[g][2] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g0z), 16 * g));
        B[g][3] = __shfl(g1x, 16 * g, 64);
        B[g][4] = __shfl(g1y, 16 * g, 64);
        B[g][5] = __shfl(g1z, 16 * g, 64);
      }
      const int ngr = (tk.w + 7) >> 3;  // groups holding candidates
      const int nbx = (tk.w + 15) >> 4;  // boxes holding candidates
      // record ranges of the 27 cells, one per lane, all loaded at once; a
      // cell beyond reach of the candidates' box gets an empty range
      int32_t qa = 0, qb = 0;
      if (l < 27) {
        int ei, ej, ek;
        walk_off(l, 1, &ei, &ej, &ek);
        const int di = ci + ei, dj = cj + ej, dk = ck + ek;
        const float nb0[3] = {(float)di, (float)dj, (float)dk};
        const float nb1[3] = {(float)di + 1.f, (float)dj + 1.f, (float)dk + 1.f};
        if (box_gap2(cb0, cb1, nb0, nb1) <= cut) {
          int32_t ln[3];
          const int32_t key = key_at(c.nx, pl.cs, di, dj, dk, ln);
          qa = cell_start[key];
          qb = cell_start[key + 1];
        }
      }
      unsigned long long *cb = cub + (int64_t)t * 27 * 8;
      uint64_t rest = __ballot(l < 27 && qa < qb);  // cells with records
      for (int i = l; i < 27 * 8; i += 64)
        if (!((rest >> (i >> 3)) & 1)) cb[i] = 0;
      uint64_t tot = 0;    // bound of group l >> 3 over the cells done
      uint64_t cellv = 0;  // the same for the current cell, flushed part
      // this lane's neighbours of the current cell, per box: at most 255
      // terms of at most 2^24 + 8 each before a flush
      uint32_t ubg[4];
#pragma unroll
      for (int g = 0; g < 4; g++) ubg[g] = 0;
      int nch = 0;  // chunks summed into ubg
      // the position after chunk (q, s): the next chunk of cell q, else the
      // first of the next cell with records (27: none left)
      auto next_pos = [&](int &q, int32_t &s) {
        if (q >= 27) return;
        s += 64;
        if (s >= __builtin_amdgcn_readlane(qb, q)) {
          rest &= ~(1ull << q);
          q = rest ? __ffsll((unsigned long long)rest) - 1 : 27;
          s = q < 27 ? __builtin_amdgcn_readlane(qa, q) : 0;
        }
      };
      // unconditional loads, clamped to the chunk's cell (rec[0] when there
      // is no chunk): no branch around them, so each wait is for the oldest
      auto load = [&](int q, int32_t s) {
        const int32_t i =
            q < 27 ? min(s + l, __builtin_amdgcn_readlane(qb, q) - 1) : 0;
        const float4 v = rec[i];
        return make_float3(v.x, v.y, v.z);
      };
      // the terms of chunk (q, s) of records r; the cell's sums are reduced
      // when it is the cell's last chunk
      auto sum_chunk = [&](int q, int32_t s, const float3 &r, bool cell_end) {
        int ei, ej, ek;
        walk_off(q, 1, &ei, &ej, &ek);
        const int di = ci + ei, dj = cj + ej, dk = ck + ek;
        const bool valid = s + l < __builtin_amdgcn_readlane(qb, q);
        // block-local coordinates of the walked cell (8 divides nx)
        const int32_t lx = (pl.cs[0] + di) & 7, ly = (pl.cs[1] + dj) & 7,
                      lz = (pl.cs[2] + dk) & 7;
        float vx = 0.f, vy = 0.f, vz = 0.f;
        bool near = false;
        if (valid) {
          vx = cell_rel(r.x, lx, di);
          vy = cell_rel(r.y, ly, dj);
          vz = cell_rel(r.z, lz, dk);
          // within reach of the task's box (its groups lie inside it): the
          // coordinate minus its clamp to the box
          const float ex = vx - __builtin_amdgcn_fmed3f(vx, cb0[0], hb[0]);
          const float ey = vy - __builtin_amdgcn_fmed3f(vy, cb0[1], hb[1]);
          const float ez = vz - __builtin_amdgcn_fmed3f(vz, cb0[2], hb[2]);
          near = fmaf(ex, ex, fmaf(ey, ey, ez * ez)) <= cut;
        }
        if (near) {
#pragma unroll
          for (int g = 0; g < 4; g++) {
            if (g < nbx) {
              // distance to the box per axis: the coordinate minus its clamp
              const float gx = vx - __builtin_amdgcn_fmed3f(vx, B[g][0], B[g][3]);
              const float gy = vy - __builtin_amdgcn_fmed3f(vy, B[g][1], B[g][4]);
              const float gz = vz - __builtin_amdgcn_fmed3f(vz, B[g][2], B[g][5]);
              const float g2 = fmaf(gx, gx, fmaf(gy, gy, gz * gz));
              ubg[g] += g2 <= cut ? sterm_bound(g2, eps2v) : 0u;
            }
          }
        }
        if (cell_end || ++nch == 255) {
          uint64_t u[8];
#pragma unroll
          for (int g = 0; g < 8; g++) {
            u[g] = g < ngr ? ubg[g >> 1] : 0;  // both groups of the box
          }
#pragma unroll
          for (int g = 0; g < 4; g++) ubg[g] = 0;
          cellv += group_reduce8(u, l);
          nch = 0;
          if (cell_end) {  // the cell is done
            tot += cellv;
            if ((l & 7) == 0) cb[q * 8 + (l >> 3)] = cellv;
            cellv = 0;
          }
        }
      };
      // chunks of 64 records of the cells in turn, three in flight: slots A,
      // B, C hold consecutive chunks; once a slot's chunk is summed, the slot
      // requests the chunk after the last one requested (tq, ts).  With one
      // chunk in flight the kernel took 66.7 ms at 1024^3, with three 65.3.
      int qA = rest ? __ffsll((unsigned long long)rest) - 1 : 27;
      int32_t sA = qA < 27 ? __builtin_amdgcn_readlane(qa, qA) : 0;
      int qB = qA, qC;
      int32_t sB = sA, sC;
      next_pos(qB, sB);
      qC = qB;
      sC = sB;
      next_pos(qC, sC);
      int tq = qC;
      int32_t ts = sC;
      float3 rA = load(qA, sA), rB = load(qB, sB), rC = load(qC, sC);
      for (;;) {
        if (qA >= 27) break;
        sum_chunk(qA, sA, rA, qB != qA);
        next_pos(tq, ts);
        qA = tq;
        sA = ts;
        rA = load(qA, sA);
        if (qB >= 27) break;
        sum_chunk(qB, sB, rB, qC != qB);
        next_pos(tq, ts);
        qB = tq;
        sB = ts;
        rB = load(qB, sB);
        if (qC >= 27) break;
        sum_chunk(qC, sC, rC, qA != qC);
        next_pos(tq, ts);
        qC = tq;
        sC = ts;
        rC = load(qC, sC);
      }
      if ((l & 7) == 0) cub[((int64_t)tcap * 27 + t) * 8 + (l >> 3)] = tot;
      unsigned long long ubmax = (l >> 3) < ngr ? tot : 0ull;
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_xor(ubmax, o, 64);
        ubmax = v > ubmax ? v : ubmax;
      }
      if (l == 0) {
        ubt[t] = ubmax;
        atomicMax(&hpil[tk.x], ubmax);
      }
    }
  }
}

// A pilot task: its candidates (one per lane), their box, and the record
// ranges of the 27 cells around the candidates' cell (lanes 0..26; a cell
// beyond reach of the box is empty).
struct PilotSetup {
  float px, py, pz;
  int32_t pidx;
  bool have;
  int ci, cj, ck;
  float cb0[3], cb1[3];
  int32_t qa, qb;
};

__device__ __forceinline__ void pilot_setup(const SodCfg &c,
                                            const int32_t *cell_start,
                                            const float4 *rec,
                                            const HaloPlan &pl, const int4 &tk,
                                            int l, float cut, PilotSetup &P) {
  walk_off(tk.y, 1, &P.ci, &P.cj, &P.ck);
  int32_t lc[3];
  key_at(c.nx, pl.cs, P.ci, P.cj, P.ck, lc);
  P.px = P.py = P.pz = 0.f;
  P.pidx = 0;
  P.have = l < tk.w;
  if (P.have) {
    const float4 r = rec[tk.z + l];
    P.px = cell_rel(r.x, lc[0], P.ci);
    P.py = cell_rel(r.y, lc[1], P.cj);
    P.pz = cell_rel(r.z, lc[2], P.ck);
    P.pidx = __float_as_int(r.w);
  }
  P.cb0[0] = wave_minf(P.have ? P.px : INFINITY);
  P.cb0[1] = wave_minf(P.have ? P.py : INFINITY);
  P.cb0[2] = wave_minf(P.have ? P.pz : INFINITY);
  P.cb1[0] = wave_maxf(P.have ? P.px : -INFINITY);
  P.cb1[1] = wave_maxf(P.have ? P.py : -INFINITY);
  P.cb1[2] = wave_maxf(P.have ? P.pz : -INFINITY);
  P.qa = P.qb = 0;
  if (l < 27) {
    int ei, ej, ek;
    walk_off(l, 1, &ei, &ej, &ek);
    const int di = P.ci + ei, dj = P.cj + ej, dk = P.ck + ek;
    const float nb0[3] = {(float)di, (float)dj, (float)dk};
    const float nb1[3] = {(float)di + 1.f, (float)dj + 1.f, (float)dk + 1.f};
    if (box_gap2(P.cb0, P.cb1, nb0, nb1) <= cut) {
      int32_t ln[3];
      const int32_t key = key_at(c.nx, pl.cs, di, dj, dk, ln);
      P.qa = cell_start[key];
      P.qb = cell_start[key + 1];
    }
  }
}

// Exact sums of a pilot task's candidates (this lane's) over at most
// `budget` chunks of 64 records of the cells in reach, starting with record
// s of cell q and going on through the cells of `rest` (cells in walk_off
// order).  The next chunk is requested before the current one is used; a
// chunk whose box lies beyond reach of the candidates' box is skipped.
// Adds the neighbours tested per candidate to npair.  nb is the wave's LDS
// slot of 64 entries.
__device__ __forceinline__ uint64_t pilot_walk(const SodCfg &c,
                                               const float4 *rec,
                                               const HaloPlan &pl,
                                               const PilotSetup &P, float4 *nb,
                                               int l, float cut, uint64_t rest,
                                               int q, int32_t s, int32_t budget,
                                               uint64_t &npair) {
  uint64_t acc = 0;
  float4 rn = make_float4(0.f, 0.f, 0.f, 0.f);
  if (q < 27 && s + l < __builtin_amdgcn_readlane(P.qb, q)) rn = rec[s + l];
  while (q < 27 && budget > 0) {
    budget--;
    const int32_t b = __builtin_amdgcn_readlane(P.qb, q);
    const int32_t m = min(64, b - s);
    const float4 r = rn;
    const int qc = q;
    int32_t sn = s + 64;
    if (sn >= b) {
      rest &= ~(1ull << q);
      q = rest ? __ffsll((unsigned long long)rest) - 1 : 27;
      sn = q < 27 ? __builtin_amdgcn_readlane(P.qa, q) : 0;
    }
    if (q < 27 && budget > 0 && sn + l < __builtin_amdgcn_readlane(P.qb, q))
      rn = rec[sn + l];
    s = sn;
    int ei, ej, ek;
    walk_off(qc, 1, &ei, &ej, &ek);
    const int di = P.ci + ei, dj = P.cj + ej, dk = P.ck + ek;
    // block-local coordinates of the walked cell (8 divides nx)
    const int32_t lx = (pl.cs[0] + di) & 7, ly = (pl.cs[1] + dj) & 7,
                  lz = (pl.cs[2] + dk) & 7;
    float vx = 0.f, vy = 0.f, vz = 0.f;
    if (l < m) {
      vx = cell_rel(r.x, lx, di);
      vy = cell_rel(r.y, ly, dj);
      vz = cell_rel(r.z, lz, dk);
    }
    if (m >= 16) {
      float nb0[3], nb1[3];
      nb0[0] = wave_minf(l < m ? vx : INFINITY);
      nb0[1] = wave_minf(l < m ? vy : INFINITY);
      nb0[2] = wave_minf(l < m ? vz : INFINITY);
      nb1[0] = wave_maxf(l < m ? vx : -INFINITY);
      nb1[1] = wave_maxf(l < m ? vy : -INFINITY);
      nb1[2] = wave_maxf(l < m ? vz : -INFINITY);
      if (box_gap2(P.cb0, P.cb1, nb0, nb1) > cut) continue;
    }
    uint32_t part = 0;  // <= 64 terms of <= 2^24
    // the chunk goes through the wave's LDS slot and is read back with
    // wave-uniform addresses (broadcast reads, no lane moves)
    nb[l] = make_float4(vx, vy, vz, 0.0f);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // m is uniform: a scalar trip count, and a select instead of a branch
    const int32_t mu = __builtin_amdgcn_readfirstlane(m);
#pragma unroll 4
    for (int32_t u = 0; u < mu; u++) {
      const float4 qv = nb[u];
      const float dx = qv.x - P.px, dy = qv.y - P.py, dz = qv.z - P.pz;
      const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const uint32_t term = sterm_fix(d2, c.eps2f);
      part += d2 <= c.rs2f ? term : 0u;
    }
    acc += part;
    npair += (uint64_t)m;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return acc;
}

// Pilot tasks with more than pilot_big records in reach (2^14 unless
// PSOD_PILOT_BIG says otherwise) are split into PILOT_PIECES pieces of their
// chunks, each summed by its own wave into per-candidate totals (integers:
// the grouping does not change S).  One wave summing a 5e7-particle halo's
// densest task alone took 41 ms.
constexpr int32_t PILOT_PIECES = 64;
constexpr int32_t PILOT_BIGCAP = 1024;  // split tasks per pass; more go whole

// Pilots, one wave per task: the 64 candidates of a task (one per lane)
// against the 27 cells around their cell (every neighbour within 1 cell lies
// there).  Evaluates exactly the tasks holding their halo's largest bound
// (every task of a halo that skipped the bound) and folds their S into
// hbest.  Dense cells are in Morton order of their 1/8-cell sub-cells, so
// candidate groups and neighbour chunks are compact.  A task with more than
// pilot_big records in reach is appended to big[] (count in *nbig, its
// per-candidate totals in pacc cleared) and left to k_centre_pilot_big.
__global__ void __launch_bounds__(256)
k_centre_pilot(SodCfg c, const int32_t *__restrict__ cell_start,
               const float4 *__restrict__ rec, const HaloPlan *__restrict__ plan,
               const int32_t *__restrict__ large, const int4 *__restrict__ tasks,
               const int32_t *__restrict__ ntask_total,
               uint64_t *__restrict__ taskres,
               unsigned long long *__restrict__ hbest,
               const unsigned long long *__restrict__ hpil,
               const unsigned long long *__restrict__ ubt,
               unsigned long long *__restrict__ cstat,
               int32_t pilot_big, int32_t *__restrict__ big,
               int32_t *__restrict__ nbig,
               unsigned long long *__restrict__ pacc,
               int32_t *__restrict__ wq) {
  __shared__ float4 nbuf[4][64];  // one neighbour chunk per wave
  const int l = threadIdx.x & 63,
            w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int32_t nt = *ntask_total;
  const float cut = c.rs2f * 1.0001f + 1e-6f;  // conservative chunk culling
  // this wave's statistics, added once at the end
  unsigned long long st_done = 0, st_pairs = 0;
  for (;;) {
    const int32_t t0 = wave_next(wq) * 16;
    if (t0 >= nt) break;
    const int32_t t1 = min(nt, t0 + 16);
    for (int32_t t = t0; t < t1; t++) {
      const int4 tk = tasks[t];
      if (ubt[t] != hpil[tk.x]) continue;
      const HaloPlan pl = plan[large[tk.x]];
      PilotSetup P;
      pilot_setup(c, cell_start, rec, pl, tk, l, cut, P);
      if (wave_sum(P.qb - P.qa) > pilot_big) {
        int32_t k = 0;
        if (l == 0) k = atomicAdd(nbig, 1);
        k = __builtin_amdgcn_readfirstlane(k);
        if (k < PILOT_BIGCAP) {
          if (l == 0) big[k] = t;
          pacc[(int64_t)k * 64 + l] = 0;
          continue;
        }
      }
      uint64_t rest = __ballot(l < 27 && P.qa < P.qb);  // cells with records
      const int q = rest ? __ffsll((unsigned long long)rest) - 1 : 27;
      const int32_t s = q < 27 ? __builtin_amdgcn_readlane(P.qa, q) : 0;
      uint64_t npair = 0;  // neighbours tested per candidate
      uint64_t acc = pilot_walk(c, rec, pl, P, nbuf[w], l, cut, rest, q, s,
                                INT32_MAX, npair);
      uint64_t key = P.have
                         ? (((uint64_t)(uint32_t)tk.y << 32) | (uint32_t)P.pidx)
                         : ~0ull;
      if (!P.have) acc = 0;
      wave_argmax(acc, key);
      if (l == 0) {
        taskres[2 * (int64_t)t] = acc;
        taskres[2 * (int64_t)t + 1] = key;
        atomicMax(&hbest[tk.x], (unsigned long long)acc);
        st_done++;
        st_pairs += npair * (uint64_t)tk.w;
      }
    }
  }
  if (l == 0 && st_done) {
    atomicAdd(&cstat[0], st_done);
    atomicAdd(&cstat[2], st_pairs);
  }
}

// Pieces of the split pilot tasks, one wave per (task, piece): piece p of a
// task's C chunks in reach (walk_off order of the cells) is chunks
// [p * per, (p + 1) * per), per = ceil(C / PILOT_PIECES); its sums are
// added to the task's per-candidate totals.
__global__ void __launch_bounds__(256)
k_centre_pilot_big(SodCfg c, const int32_t *__restrict__ cell_start,
                   const float4 *__restrict__ rec,
                   const HaloPlan *__restrict__ plan,
                   const int32_t *__restrict__ large,
                   const int4 *__restrict__ tasks,
                   const int32_t *__restrict__ big,
                   const int32_t *__restrict__ nbig,
                   unsigned long long *__restrict__ pacc,
                   unsigned long long *__restrict__ cstat,
                   int32_t *__restrict__ wq) {
  __shared__ float4 nbuf[4][64];  // one neighbour chunk per wave
  const int l = threadIdx.x & 63,
            w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int32_t nu = min(*nbig, PILOT_BIGCAP) * PILOT_PIECES;
  const float cut = c.rs2f * 1.0001f + 1e-6f;  // conservative chunk culling
  unsigned long long st_pairs = 0;
  for (;;) {
    const int32_t u = wave_next(wq);
    if (u >= nu) break;
    const int32_t b = u / PILOT_PIECES, p = u % PILOT_PIECES;
    const int32_t t = big[b];
    const int4 tk = tasks[t];
    const HaloPlan pl = plan[large[tk.x]];
    PilotSetup P;
    pilot_setup(c, cell_start, rec, pl, tk, l, cut, P);
    // chunks of each cell and their prefix over the cells
    const int32_t nch = l < 27 ? (P.qb - P.qa + 63) >> 6 : 0;
    int32_t C;
    const int32_t ex = wave_excl(nch, &C);
    const int32_t per = (C + PILOT_PIECES - 1) / PILOT_PIECES;
    const int32_t k0 = p * per;
    if (k0 >= C) continue;
    // the cell holding chunk k0, and that chunk's first record
    const uint64_t at = __ballot(nch > 0 && ex <= k0 && k0 < ex + nch);
    const int q = __ffsll((unsigned long long)at) - 1;
    const int32_t s = __builtin_amdgcn_readlane(P.qa, q) +
                      (k0 - __builtin_amdgcn_readlane(ex, q)) * 64;
    const uint64_t rest =
        __ballot(l < 27 && P.qa < P.qb) & ~((1ull << q) - 1ull);
    uint64_t npair = 0;
    const uint64_t acc = pilot_walk(c, rec, pl, P, nbuf[w], l, cut, rest, q,
                                    s, min(per, C - k0), npair);
    if (P.have) atomicAdd(&pacc[(int64_t)b * 64 + l], (unsigned long long)acc);
    if (l == 0) st_pairs += npair * (uint64_t)tk.w;
  }
  if (l == 0 && st_pairs) atomicAdd(&cstat[2], st_pairs);
}

// Results of the split pilot tasks, one wave per task: the best (S, key)
// over its candidates' totals, folded into hbest.
__global__ void __launch_bounds__(256)
k_centre_pilot_fin(const float4 *__restrict__ rec,
                   const int4 *__restrict__ tasks,
                   const int32_t *__restrict__ big,
                   const int32_t *__restrict__ nbig,
                   const unsigned long long *__restrict__ pacc,
                   uint64_t *__restrict__ taskres,
                   unsigned long long *__restrict__ hbest,
                   unsigned long long *__restrict__ cstat) {
  const int l = threadIdx.x & 63;
  const int32_t b = (int32_t)(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (b >= min(*nbig, PILOT_BIGCAP)) return;
  const int32_t t = big[b];
  const int4 tk = tasks[t];
  const bool have = l < tk.w;
  uint64_t acc = have ? pacc[(int64_t)b * 64 + l] : 0;
  uint64_t key = have ? (((uint64_t)(uint32_t)tk.y << 32) |
                         (uint32_t)__float_as_int(rec[tk.z + l].w))
                      : ~0ull;
  wave_argmax(acc, key);
  if (l == 0) {
    taskres[2 * (int64_t)t] = acc;
    taskres[2 * (int64_t)t + 1] = key;
    atomicMax(&hbest[tk.x], (unsigned long long)acc);
    atomicAdd(&cstat[0], 1ull);
  }
}

// Work list of the exact phase: item t*8 + g (g even) for every box of 16
// candidates (groups g and g+1, which share the box's bound) of a task that
// was not a pilot, whose bound reaches the halo's best after the pilots.  A
// box whose bound is below the best can neither hold the maximum nor tie
// with it.
__global__ void __launch_bounds__(256)
k_task_items(const int4 *__restrict__ tasks,
             const int32_t *__restrict__ ntask_total, int32_t tcap,
             const unsigned long long *__restrict__ ubt,
             const unsigned long long *__restrict__ hpil,
             const unsigned long long *__restrict__ cub,
             const unsigned long long *__restrict__ hbest,
             int32_t *__restrict__ items, int32_t *__restrict__ nitems,
             unsigned long long *__restrict__ cstat) {
  const int32_t nt = *ntask_total;
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t mask = 0;
  bool skipped = false;
  if (t < nt) {
    const int4 tk = tasks[t];
    const unsigned long long ub = ubt[t];
    if (ub != hpil[tk.x]) {
      const unsigned long long best = hbest[tk.x];
      const int ngr = (tk.w + 7) >> 3;
      if (ub >= best)
        for (int g = 0; g < ngr; g += 2)
          if (cub[((int64_t)tcap * 27 + t) * 8 + g] >= best) mask |= 1u << g;
      skipped = mask == 0;
    }
  }
  int32_t tot;
  const int32_t ex = wave_excl(__popc(mask), &tot);
  int32_t base = 0;
  if ((threadIdx.x & 63) == 0 && tot > 0) base = atomicAdd(nitems, tot);
  int32_t p = __shfl(base, 0, 64) + ex;
  while (mask) {
    const int g = __ffs(mask) - 1;
    mask &= mask - 1;
    items[p++] = t * 8 + g;
  }
  const unsigned long long ns = wave_sum((unsigned long long)(skipped ? 1 : 0));
  if ((threadIdx.x & 63) == 0 && ns) atomicAdd(&cstat[1], ns);
}

// Sort keys of the items: the ratio of the item's bound to its halo's best
// after the pilots, largest first (the groups most likely to raise the best
// run first; those after them meet a higher best).
__global__ void k_item_keys(const int32_t *__restrict__ items,
                            const int32_t *__restrict__ nitems,
                            const int4 *__restrict__ tasks,
                            const int32_t *__restrict__ ntask_total,
                            int32_t tcap,
                            const unsigned long long *__restrict__ cub,
                            const unsigned long long *__restrict__ hbest,
                            uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *nitems) return;
  const int32_t it = items[i];
  const int32_t t = it >> 3, g = it & 7;
  const double ub = (double)cub[((int64_t)tcap * 27 + t) * 8 + g];
  const double b = (double)hbest[tasks[t].x];
  const double r = b > 0.0 ? ub / b : 1e9;
  const double q = fmin(r, 255.0) * 16777216.0;  // 2^24 per unit of ratio
  keys[i] = 0xFFFFFFFFull - (uint64_t)q;
  vals[i] = (uint32_t)it;
}

// Exact phase, one wave per item (task t, candidates 8g..8g+15).  Lane l
// takes candidate l & 15.  The neighbours within reach of the item's box are
// queued in LDS across chunks and cells and taken 64 at a time, neighbour
// u + (l >> 4) by lane l in steps of 4, so that every step evaluates 64
// pairs.  Each pair's term comes from the same operands and operations as in
// the pilots (neighbour minus candidate, the same fused squares), so S is
// the same integer.  After each cell (nearest first: the cell itself, the 6
// faces, the 12 edges, the 8 corners) the queue is drained and the item is
// abandoned when no candidate's exact sum so far plus the bound of the
// cells left reaches the halo's best.  A drain pads at most 3 entries, so
// checking after every cell rather than every shell of cells costs little
// and abandons earlier (44.4 -> 39.7 ms at 1024^3, 8% fewer pair tests).
// Items of 8 candidates (the bound's boxes were of 8 then) took 236 ms of
// SOD at c3 against 193 ms with boxes and items of 16.
// 5 waves per SIMD (94 VGPRs, a few spilled): 48.2 -> 46.5 ms at 1024^3
__global__ void __launch_bounds__(256, 5)
k_centre_group(SodCfg c, const int32_t *__restrict__ cell_start,
               const float4 *__restrict__ rec, const HaloPlan *__restrict__ plan,
               const int32_t *__restrict__ large, const int4 *__restrict__ tasks,
               const int32_t *__restrict__ ntask_total, int32_t tcap,
               const int32_t *__restrict__ items,
               const int32_t *__restrict__ nitems,
               unsigned long long *__restrict__ hbest,
               const unsigned long long *__restrict__ cub,
               uint64_t *__restrict__ itemres,
               unsigned long long *__restrict__ cstat,
               int32_t *__restrict__ wq) {
  __shared__ float4 gq[4][128];  // queued neighbours, one queue per wave
  const int l = threadIdx.x & 63,
            w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int32_t nt = *ntask_total, ni = *nitems;
  const float cut = c.rs2f * 1.0001f + 1e-6f;  // conservative culling
  const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));  // lanes < l
  // this wave's statistics, added once at the end
  unsigned long long st_done = 0, st_skip = 0, st_abort = 0, st_pairs = 0;
  // blocks of 2 items in list order (most promising first)
  for (;;) {
    const int32_t i0 = wave_next(wq) * 2;
    if (i0 >= ni) break;
    const int32_t i1 = min(ni, i0 + 2);
    for (int32_t i = i0; i < i1; i++) {
      const int32_t it = items[i];
      const int32_t t = it >> 3, g = it & 7;
      const int4 tk = tasks[t];
      uint64_t *res = itemres + 2 * (int64_t)it;
      unsigned long long best = __hip_atomic_load(
          &hbest[tk.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cub[((int64_t)tcap * 27 + t) * 8 + g] < best) {
        if (l == 0) {
          res[0] = 0;
          res[1] = ~0ull;
          st_skip++;
        }
        continue;
      }
      const HaloPlan pl = plan[large[tk.x]];
      int ci, cj, ck;
      walk_off(tk.y, 1, &ci, &cj, &ck);
      int32_t lc[3];
      key_at(c.nx, pl.cs, ci, cj, ck, lc);
      const int nc = min(16, tk.w - 8 * g);  // candidates of the item
      const int kc = l & 15;                // this lane's candidate
      // lanes of missing candidates sit beyond every cutoff
      float px = 1e30f, py = 1e30f, pz = 1e30f;
      int32_t pidx = 0;
      if (kc < nc) {
        const float4 r = rec[tk.z + 8 * g + kc];
        px = cell_rel(r.x, lc[0], ci);
        py = cell_rel(r.y, lc[1], cj);
        pz = cell_rel(r.z, lc[2], ck);
        pidx = __float_as_int(r.w);
      }
      float gb0[3], gb1[3];  // the item's box
      gb0[0] = wave_minf(kc < nc ? px : INFINITY);
      gb0[1] = wave_minf(kc < nc ? py : INFINITY);
      gb0[2] = wave_minf(kc < nc ? pz : INFINITY);
      gb1[0] = wave_maxf(kc < nc ? px : -INFINITY);
      gb1[1] = wave_maxf(kc < nc ? py : -INFINITY);
      gb1[2] = wave_maxf(kc < nc ? pz : -INFINITY);
      // the maxima in vector registers: v_med3 takes one scalar operand
      const float hb[3] = {__shfl(gb1[0], 0, 64), __shfl(gb1[1], 0, 64),
                           __shfl(gb1[2], 0, 64)};
      // record ranges of the 27 cells nearest first and the item's bound in
      // each, one per lane; a cell beyond reach of the item's box is empty
      // (its bound is 0: none of its records is within reach of the group)
      int32_t qa = 0, qb = 0, cq = 0;
      unsigned long long qbnd = 0;
      if (l < 27) {
        const int q = kCellOrder[l];
        cq = q;
        int ei, ej, ek;
        walk_off(q, 1, &ei, &ej, &ek);
        const int di = ci + ei, dj = cj + ej, dk = ck + ek;
        const float nb0[3] = {(float)di, (float)dj, (float)dk};
        const float nb1[3] = {(float)di + 1.f, (float)dj + 1.f, (float)dk + 1.f};
        if (box_gap2(gb0, gb1, nb0, nb1) <= cut) {
          int32_t ln[3];
          const int32_t key = key_at(c.nx, pl.cs, di, dj, dk, ln);
          qa = cell_start[key];
          qb = cell_start[key + 1];
        }
        qbnd = cub[(int64_t)t * 27 * 8 + q * 8 + g];
      }
      // bound of cells qi.. (nearest-first order): suffix sums over lanes
#pragma unroll
      for (int o = 1; o < 32; o <<= 1) {
        const unsigned long long v = __shfl_down(qbnd, o, 64);
        if (l + o < 27) qbnd += v;
      }
      uint64_t rest = __ballot(l < 27 && qa < qb);  // cells with records
      int qi = rest ? __ffsll((unsigned long long)rest) - 1 : 27;
      int32_t s = qi < 27 ? __builtin_amdgcn_readlane(qa, qi) : 0;
      float4 rn = make_float4(0.f, 0.f, 0.f, 0.f);
      if (qi < 27 && s + l < __builtin_amdgcn_readlane(qb, qi)) rn = rec[s + l];
      // this lane's share of its candidate's sum: u32 steps of at most 2^24,
      // folded into acc64 before 255 steps
      uint32_t part = 0;
      int nst = 0, qlen = 0;
      uint64_t acc64 = 0, npair = 0;
      bool aborted = false;
      // n queued neighbours from the queue's head, 8 per step
      // the queue's first 8 * ns entries, 8 per step (entries past the
      // queued ones are padding beyond every cutoff)
      auto pair_step = [&](int j) {
        const float4 qv = gq[w][j];
        const float dx = qv.x - px, dy = qv.y - py, dz = qv.z - pz;
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const uint32_t term = sterm_fix(d2, c.eps2f);
        part += d2 <= c.rs2f ? term : 0u;
      };
      auto fold = [&](int ns) {
        if (nst > 255 - ns) {
          acc64 += part;
          part = 0;
          nst = 0;
        }
        nst += ns;
      };
      while (qi < 27) {
        const int q = __builtin_amdgcn_readlane(cq, qi);
        const int32_t b = __builtin_amdgcn_readlane(qb, qi);
        const int32_t m = min(64, b - s);
        const float4 r = rn;
        int qn = qi;
        int32_t sn = s + 64;
        if (sn >= b) {
          rest &= ~(1ull << qi);
          qn = rest ? __ffsll((unsigned long long)rest) - 1 : 27;
          sn = qn < 27 ? __builtin_amdgcn_readlane(qa, qn) : 0;
        }
        if (qn < 27 && sn + l < __builtin_amdgcn_readlane(qb, qn)) rn = rec[sn + l];
        int ei, ej, ek;
        walk_off(q, 1, &ei, &ej, &ek);
        const int di = ci + ei, dj = cj + ej, dk = ck + ek;
        // block-local coordinates of the walked cell (8 divides nx)
        const int32_t lx = (pl.cs[0] + di) & 7, ly = (pl.cs[1] + dj) & 7,
                      lz = (pl.cs[2] + dk) & 7;
        bool near = false;
        float vx = 0.f, vy = 0.f, vz = 0.f;
        if (l < m) {
          vx = cell_rel(r.x, lx, di);
          vy = cell_rel(r.y, ly, dj);
          vz = cell_rel(r.z, lz, dk);
          // the coordinate minus its clamp to the item's box
          const float ex = vx - __builtin_amdgcn_fmed3f(vx, gb0[0], hb[0]);
          const float ey = vy - __builtin_amdgcn_fmed3f(vy, gb0[1], hb[1]);
          const float ez = vz - __builtin_amdgcn_fmed3f(vz, gb0[2], hb[2]);
          near = fmaf(ex, ex, fmaf(ey, ey, ez * ez)) <= cut;
        }
        const uint64_t nm = __ballot(near);
        const int nn = __popcll(nm);
        if (near) gq[w][qlen + __popcll(nm & below)] = make_float4(vx, vy, vz, 0.f);
        qlen += nn;
        npair += (uint64_t)nn;
        const bool cell_end = qn != qi;
        if (qlen >= 64 || (cell_end && qlen > 0)) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (qlen >= 64) {
            // 64 entries in 8 steps at fixed LDS offsets
            fold(16);
#pragma unroll
            for (int u = 0; u < 16; u++) pair_step(4 * u + (l >> 4));
            // move the entries past them to the queue's head
            qlen -= 64;
            float4 mv = make_float4(0.f, 0.f, 0.f, 0.f);
            if (l < qlen) mv = gq[w][64 + l];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (l < qlen) gq[w][l] = mv;
          }
          if (cell_end && qlen > 0) {
            // at most 63 left: pad to a multiple of 8 and take them too
            const int ns = (qlen + 3) >> 2;
            if (l >= qlen && l < 4 * ns)
              gq[w][l] = make_float4(1e30f, 1e30f, 1e30f, 0.f);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            fold(ns);
            for (int u = 0; u < ns; u++) pair_step(4 * u + (l >> 4));
            qlen = 0;
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (cell_end && qn < 27) {
          // every cell before qn is summed, the cells from qn on are bounded
          acc64 += part;
          part = 0;
          nst = 0;
          uint64_t v = acc64;  // sum over the 4 lanes of each candidate
          v += __shfl_xor(v, 16, 64);
          v += __shfl_xor(v, 32, 64);
          const uint32_t blo = __builtin_amdgcn_readlane((uint32_t)qbnd, qn);
          const uint32_t bhi =
              __builtin_amdgcn_readlane((uint32_t)(qbnd >> 32), qn);
          const unsigned long long rem = ((unsigned long long)bhi << 32) | blo;
          if (__ballot(l < nc && v + rem >= best) == 0) {
            aborted = true;
            break;
          }
          best = __hip_atomic_load(&hbest[tk.x], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        qi = qn;
        s = sn;
      }
      if (aborted) {
        if (l == 0) {
          res[0] = 0;
          res[1] = ~0ull;
          st_abort++;
          st_pairs += npair * (uint64_t)nc;
        }
        continue;
      }
      acc64 += part;
      uint64_t tot = acc64;  // sum over the 4 lanes of each candidate
      tot += __shfl_xor(tot, 16, 64);
      tot += __shfl_xor(tot, 32, 64);
      // lanes 0..nc-1 hold the candidates' S
      uint64_t acc = l < nc ? tot : 0;
      uint64_t key = l < nc ? (((uint64_t)(uint32_t)tk.y << 32) | (uint32_t)pidx)
                            : ~0ull;
      wave_argmax(acc, key);
      if (l == 0) {
        res[0] = acc;
        res[1] = key;
        atomicMax(&hbest[tk.x], (unsigned long long)acc);
        st_done++;
        st_pairs += npair * (uint64_t)nc;
      }
    }
  }
  if (l == 0) {
    if (st_done) atomicAdd(&cstat[0], st_done);
    if (st_skip) atomicAdd(&cstat[1], st_skip);
    if (st_pairs) atomicAdd(&cstat[2], st_pairs);
    if (st_abort) atomicAdd(&cstat[3], st_abort);
  }
}

// result pairs (S, key) that lose to any evaluated candidate
__global__ void k_res_init(uint64_t *__restrict__ r, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[2 * i] = 0;
  r[2 * i + 1] = ~0ull;
}

// Each halo's centre: the best (S, key) over its pilot tasks and the groups
// of its other tasks, key = (candidate cell rank << 32) | particle index.
// One wave per halo, its lanes over the tasks.
__global__ void __launch_bounds__(256)
k_centre_reduce(HaloPlan *__restrict__ plan, const int32_t *__restrict__ large,
                int32_t nlarge, const int32_t *__restrict__ dn,
                const uint64_t *__restrict__ taskres,
                const uint64_t *__restrict__ itemres) {
  const int l = threadIdx.x & 63;
  const int32_t e = blockIdx.x * 4 + (int32_t)(threadIdx.x >> 6);
  if (e >= nlarge || (dn && e >= *dn)) return;  // uniform per wave
  const int32_t h = large[e];
  const HaloPlan pl = plan[h];
  if (pl.centre >= 0 || pl.ntask == 0) return;
  uint64_t ba = 0, bk = ~0ull;
  for (int32_t t = pl.task0 + l; t < pl.task0 + pl.ntask; t += 64) {
    const uint64_t a = taskres[2 * (int64_t)t], k = taskres[2 * (int64_t)t + 1];
    if (better(a, k, ba, bk)) { ba = a; bk = k; }
    for (int g = 0; g < 8; g++) {
      const int64_t it = (int64_t)t * 8 + g;
      const uint64_t a2 = itemres[2 * it], k2 = itemres[2 * it + 1];
      if (better(a2, k2, ba, bk)) { ba = a2; bk = k2; }
    }
  }
  wave_argmax(ba, bk);
  if (l == 0) plan[h].centre = key_idx(bk);
}

// ---------------------------------------------------------------------------
// large halos: growth with radial histograms, one workgroup per halo
// ---------------------------------------------------------------------------
constexpr int GL_T = 512;
constexpr int SORT_CAP = 4096;
// segments of possible crossing bins per threshold; ranges of collected
// bins (the last slot holds the second threshold's band)
constexpr int SEGCAP = 4;
constexpr int32_t GAP_KEYS = 512;  // runs this close share a segment
constexpr int R2 = PSOD_MAXDELTA * SEGCAP;  // first slot of the delta2 segments
constexpr int MAXRANGE = R2 + SEGCAP;
// radial histogram of one workgroup: HIST_MIN..HIST_MAX bins (powers of
// two, a multiple of GL_T) and the total
constexpr int32_t HIST_MIN = 4096, HIST_MAX = 1 << 20;
// the first HIST_LDS bins (the dense inner shells) are counted in LDS
constexpr int32_t HIST_LDS = 16384;
constexpr int32_t SCAN_ROUND = 4096;  // bins per round of the scan
constexpr int64_t HIST_SLICE = HIST_MAX + 64;

struct GrowShared {
  uint64_t skey[SORT_CAP];
  int32_t pre[GL_T + 1];
  int32_t st[GL_T];
  int32_t sdi[GL_T];  // per segment: di of the cell at x = 0 of its block
  int32_t spk[GL_T];  // per segment: dj, dk (12 bits each, +2048), y & 7, z & 7
  int32_t red_i[GL_T / 64];
  int64_t red_l[GL_T / 64];
  double red_d[GL_T / 64];
  double prad[257];
  int32_t pcnt[GL_T / 64][257];  // per wave
  int32_t item;
  int32_t fill[MAXRANGE];
  int64_t off[MAXRANGE];
  int32_t rs[MAXRANGE], re[MAXRANGE];  // bin ranges, ascending, disjoint
  int64_t k0[MAXRANGE], k1[MAXRANGE];  // ranks before / at the range end
  int32_t nr;
  int32_t sg_s[PSOD_MAXDELTA][SEGCAP], sg_e[PSOD_MAXDELTA][SEGCAP],
      sg_x[PSOD_MAXDELTA][SEGCAP], nseg[PSOD_MAXDELTA];
  int64_t bump;
  int64_t need;  // slots the current halo's bands need
};

__device__ inline int32_t block_min_i32(int32_t v, GrowShared &G) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) G.red_i[threadIdx.x >> 6] = v;
  __syncthreads();
  int32_t r = G.red_i[0];
  for (int i = 1; i < GL_T / 64; i++) r = min(r, G.red_i[i]);
  __syncthreads();
  return r;
}
__device__ inline int32_t block_max_i32(int32_t v, GrowShared &G) {
  return -block_min_i32(-v, G);
}
__device__ inline int64_t block_sum_i64(int64_t v, GrowShared &G) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) G.red_l[threadIdx.x >> 6] = v;
  __syncthreads();
  int64_t r = 0;
  for (int i = 0; i < GL_T / 64; i++) r += G.red_l[i];
  __syncthreads();
  return r;
}
__device__ inline double block_sum_f64(double v, GrowShared &G) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) G.red_d[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0;
  for (int i = 0; i < GL_T / 64; i++) r += G.red_d[i];
  __syncthreads();
  return r;
}

// first b in [lo, hi] with pred(b) (INT32_MAX if none), scanning upward in
// chunks of SCAN_U * GL_T bins so that an early hit ends the scan
constexpr int SCAN_U = 16;
template <typename P>
__device__ inline int32_t block_first(int32_t lo, int32_t hi, P &&pred,
                                      GrowShared &G) {
  for (int32_t base = lo; base <= hi; base += SCAN_U * GL_T) {
    int32_t m = INT32_MAX;
    for (int u = 0; u < SCAN_U; u++) {
      const int32_t b = base + u * GL_T + (int32_t)threadIdx.x;
      if (m == INT32_MAX && b <= hi && pred(b)) m = b;
    }
    m = block_min_i32(m, G);
    if (m != INT32_MAX) return m;
  }
  return INT32_MAX;
}
// last b in [lo, hi] with pred(b) (-1 if none), scanning downward
template <typename P>
__device__ inline int32_t block_last(int32_t lo, int32_t hi, P &&pred,
                                     GrowShared &G) {
  for (int32_t top = hi; top >= lo; top -= SCAN_U * GL_T) {
    int32_t m = -1;
    for (int u = 0; u < SCAN_U; u++) {
      const int32_t b = top - u * GL_T - (int32_t)threadIdx.x;
      if (m < 0 && b >= lo && pred(b)) m = b;
    }
    m = block_max_i32(m, G);
    if (m >= 0) return m;
  }
  return -1;
}


// The cube of cells at offsets [-W, W]^3 is walked as row segments: for
// fixed (dj, dk) the x range splits at 8^3-block boundaries, and the cells of
// one block row are consecutive keys, so a segment's particles are one
// contiguous run of records.  A particle's x cell is floor(u) of its record.
struct CubeSegs {
  int32_t X0, X1, bx0, nsegx, side, nb;
  int64_t nseg;
};

__device__ inline int32_t floor8(int32_t x) { return x >= 0 ? x >> 3 : -((7 - x) >> 3); }

__device__ inline CubeSegs cube_segs(const SodCfg &c, const CellRef &cr, int W) {
  CubeSegs g;
  g.X0 = cr.c[0] - W;
  g.X1 = cr.c[0] + W;
  g.bx0 = floor8(g.X0);
  g.nsegx = floor8(g.X1) - g.bx0 + 1;
  g.side = 2 * W + 1;
  g.nb = c.nx >> 3;
  g.nseg = (int64_t)g.side * g.side * g.nsegx;
  return g;
}

// segment q: its block (unwrapped), row offsets, wrapped row cells, and its
// first cell key and cell count along x
__device__ inline void cube_seg(const SodCfg &c, const CellRef &cr,
                                const CubeSegs &g, int W, int32_t q,
                                int32_t *bx, int *dj, int *dk, int32_t *gy,
                                int32_t *gz, int32_t *ka, int32_t *ncx) {
  const int32_t row = q / g.nsegx;
  *bx = g.bx0 + (q - row * g.nsegx);
  *dj = row / g.side - W;
  *dk = row % g.side - W;
  *gy = wrapi(cr.c[1] + *dj, c.nx);
  *gz = wrapi(cr.c[2] + *dk, c.nx);
  const int32_t xa = max(g.X0, 8 * *bx), xb = min(g.X1, 8 * *bx + 7);
  const int32_t gbx = wrapi(*bx, g.nb);
  *ka = cell_key(c.nx, 8 * gbx + (xa - 8 * *bx), *gy, *gz);
  *ncx = xb - xa + 1;
}

// particles in the cells at offsets [-W, W]^3 (cell counts only), of the
// segments q0 <= q < q1; with segcnt, segment q's count to segcnt[q]
__device__ inline int64_t cube_count(const SodCfg &c, GrowShared &G,
                                     const int32_t *cell_start,
                                     const CellRef &cr, int W, int64_t q0 = 0,
                                     int64_t q1 = INT64_MAX,
                                     int32_t *segcnt = nullptr) {
  const CubeSegs g = cube_segs(c, cr, W);
  const int64_t qe = min(q1, g.nseg);
  int64_t n = 0;
  for (int64_t q = q0 + threadIdx.x; q < qe; q += GL_T) {
    int32_t bx, gy, gz, ka, ncx;
    int dj, dk;
    cube_seg(c, cr, g, W, (int32_t)q, &bx, &dj, &dk, &gy, &gz, &ka, &ncx);
    const int32_t m = cell_start[ka + ncx] - cell_start[ka];
    if (segcnt) segcnt[q] = m;
    n += m;
  }
  return block_sum_i64(n, G);
}

constexpr int CW_U = 4;

// Whether a row segment (cell offsets dia..dib along x, dj, dk) can hold a
// particle whose squared distance from the reference lies in [lo2, hi2].
// A particle of the cell at offset d lies in [d - f, d + 1 - f) along each
// axis; the intervals are widened by 1e-3 cells for the float rounding of
// the distances.
__device__ inline bool seg_meets(const CellRef &cr, int dia, int dib, int dj,
                                 int dk, float lo2, float hi2) {
  float mn = 0.0f, mx = 0.0f;
  auto axis = [&](float a, float b) {
    a -= 1e-3f;
    b += 1e-3f;
    const float m = a > 0.0f ? a : (b < 0.0f ? -b : 0.0f);
    const float M = fmaxf(fabsf(a), fabsf(b));
    mn += m * m;
    mx += M * M;
  };
  axis((float)dia - cr.f[0], (float)(dib + 1) - cr.f[0]);
  axis((float)dj - cr.f[1], (float)(dj + 1) - cr.f[1]);
  axis((float)dk - cr.f[2], (float)(dk + 1) - cr.f[2]);
  return mn <= hi2 && mx >= lo2;
}

// Band of squared distances [*lo2, *hi2] that holds every particle of
// d2 < W2 whose rbin() lies in [b0, b1], with a 1e-4 relative margin.
__device__ inline void bin_band(int32_t b0, int32_t b1, int32_t nb,
                                float scale, float W2, float *lo2,
                                float *hi2) {
  const double s = (double)scale;
  if (b0 > 0) {
    const double c0 = cbrt((double)b0 / s * (1.0 - 1e-4));
    *lo2 = (float)(c0 * c0);
  } else {
    *lo2 = -1.0f;
  }
  if (b1 >= nb - 1) {
    *hi2 = W2 * (1.0f + 1e-4f) + 1e-3f;
  } else {
    const double c1 = cbrt((double)(b1 + 1) / s * (1.0 + 1e-4));
    *hi2 = (float)(c1 * c1) + 1e-6f;
  }
}

// Calls f(d2, idx, rec, di, dj, dk, lc) for the particles of the cells at
// offsets [-W, W]^3 from the reference cell, skipping the row segments that
// cannot hold a particle with d2 in [lo2, hi2] (f sees all the others).
// Segments q0 <= q < q1 only.  An f that takes an eighth argument also gets
// the particle's ordinal among those walked (0, 1, ... in segment order).
template <typename F>
__device__ inline void cube_walk(const SodCfg &c, GrowShared &G,
                                 const int32_t *cell_start, const float4 *rec,
                                 const CellRef &cr, int W, float lo2,
                                 float hi2, F &&f, int64_t q0 = 0,
                                 int64_t q1 = INT64_MAX) {
  const CubeSegs g = cube_segs(c, cr, W);
  const int64_t qe = min(q1, g.nseg);
  int64_t done = 0;  // particles of the rounds before this one
  for (int64_t base = q0; base < qe; base += GL_T) {
    const int64_t q = base + threadIdx.x;
    int32_t cnt = 0, st = 0, sdi = 0, spk = 0;
    if (q < qe) {
      int32_t bx, gy, gz, ka, ncx;
      int dj, dk;
      cube_seg(c, cr, g, W, (int32_t)q, &bx, &dj, &dk, &gy, &gz, &ka, &ncx);
      const int dia = max(g.X0, 8 * bx) - cr.c[0];
      if (seg_meets(cr, dia, dia + ncx - 1, dj, dk, lo2, hi2)) {
        st = cell_start[ka];
        cnt = cell_start[ka + ncx] - st;
      }
      // |dj|, |dk| <= W < 2048 (2W + 1 <= nx)
      sdi = 8 * bx - cr.c[0];
      spk = (dj + 2048) | ((dk + 2048) << 12) | ((gy & 7) << 24) |
            ((gz & 7) << 27);
    }
    // block exclusive scan of cnt
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    int32_t tot;
    const int32_t ex = wave_excl(cnt, &tot);
    if (l == 0) G.red_i[w] = tot;
    __syncthreads();
    int32_t off = 0, all = 0;
    for (int i = 0; i < GL_T / 64; i++) {
      if (i < w) off += G.red_i[i];
      all += G.red_i[i];
    }
    G.pre[threadIdx.x] = off + ex;
    G.st[threadIdx.x] = st;
    G.sdi[threadIdx.x] = sdi;
    G.spk[threadIdx.x] = spk;
    if (threadIdx.x == 0) G.pre[GL_T] = all;
    __syncthreads();
    const int nq = (int)min((int64_t)GL_T, qe - base);
    // CW_U particles per thread per step: their searches and loads are
    // independent, so their latencies overlap
    for (int32_t j0 = threadIdx.x; j0 < all; j0 += CW_U * GL_T) {
      int qq[CW_U];
      float4 rv[CW_U];
#pragma unroll
      for (int u = 0; u < CW_U; u++) {
        const int32_t j = j0 + u * GL_T;
        qq[u] = 0;
        if (j < all) {
          qq[u] = bsearch_last_le(G.pre, nq, j);
          rv[u] = rec[G.st[qq[u]] + (j - G.pre[qq[u]])];
        }
      }
#pragma unroll
      for (int u = 0; u < CW_U; u++) {
        if (j0 + u * GL_T >= all) continue;
        const float4 r = rv[u];
        const int32_t pk = G.spk[qq[u]];
        const int32_t lx = (int32_t)r.x;  // u in [0, 8): the cell in the block
        const int di = G.sdi[qq[u]] + lx;
        const int dj = (pk & 4095) - 2048, dk = ((pk >> 12) & 4095) - 2048;
        int32_t lc[3] = {lx, (pk >> 24) & 7, (pk >> 27) & 7};
        const float dx = rel1(r.x, lc[0], di, cr.f[0]);
        const float dy = rel1(r.y, lc[1], dj, cr.f[1]);
        const float dz = rel1(r.z, lc[2], dk, cr.f[2]);
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if constexpr (std::is_invocable_v<F &, float, int32_t, const float4 &,
                                          int, int, int, const int32_t *,
                                          int64_t>)
          f(d2, __float_as_int(r.w), r, di, dj, dk, lc, done + j0 + u * GL_T);
        else
          f(d2, __float_as_int(r.w), r, di, dj, dk, lc);
      }
    }
    done += all;
    __syncthreads();
  }
}

// radial bin of d2 for nb bins uniform in r^3 over [0, W^3)
__device__ inline int32_t rbin(float d2, float scale, int32_t nb) {
  const float v = d2 * sqrtf(d2) * scale;
  const int32_t b = (int32_t)v;
  return b < nb ? b : nb - 1;
}

// Edges of bin b in r^3, widened by 4e-5 (relative) so that every particle
// binned by rbin() lies inside them whatever the float rounding of d2^1.5
__device__ inline double edge3_lo(int32_t b, int32_t nb, double W3) {
  return W3 * ((double)b / (double)nb) * (1.0 - 4e-5);
}
__device__ inline double edge3_hi(int32_t b, int32_t nb, double W3) {
  return W3 * ((double)(b + 1) / (double)nb) * (1.0 + 4e-5) + 1e-12;
}
// r^3 at which k particles have overdensity delta (od_k(k, m, r) = delta)
__device__ inline double rad3_for(int64_t k, double mpart, double delta) {
  return ((double)k * mpart) / (delta * (2.0 / 3.0 * TWOPI_REF));
}


// Pool slots reserved for n collected keys: n itself when the keys are
// sorted in LDS, the next power of two otherwise (padding for the bitonic
// network on global memory).
__device__ inline int64_t pool_slots(int64_t n) {
  if (n <= SORT_CAP) return n;
  int64_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// sorts the n keys at pool[0..n) ascending; pool holds pool_slots(n) slots
__device__ void block_sort_pool(uint64_t *pool, int32_t n, GrowShared &G) {
  if (n <= 1) return;
  if (n <= SORT_CAP) {
    const int np = next_pow2(n);
    for (int i = threadIdx.x; i < np; i += GL_T)
      G.skey[i] = i < n ? pool[i] : ~0ull;
    __syncthreads();
    lds_bitonic<GL_T>(G.skey, np);
    for (int i = threadIdx.x; i < n; i += GL_T) pool[i] = G.skey[i];
    __threadfence_block();
    __syncthreads();
    return;
  }
  // The bitonic network over np = next_pow2(n) keys (padding ~0 sorts
  // last).  Its steps (k, j) with j < SORT_CAP pair keys of one block of
  // SORT_CAP; they run block by block in LDS, one load and store of the
  // block per k.  Only the steps with j >= SORT_CAP pass through global
  // memory: for np = 2^20, 36 passes instead of 210.
  const int np = next_pow2(n);
  for (int i = n + threadIdx.x; i < np; i += GL_T) pool[i] = ~0ull;
  __threadfence_block();
  __syncthreads();
  // steps k in [k0, k1], j < SORT_CAP, of every block; a pair's direction
  // is that of the whole network ((i & k) == 0: ascending)
  auto blocks = [&](int k0, int k1) {
    for (int base = 0; base < np; base += SORT_CAP) {
      for (int i = threadIdx.x; i < SORT_CAP; i += GL_T)
        G.skey[i] = pool[base + i];
      __syncthreads();
      for (int k = k0; k <= k1; k <<= 1)
        for (int j = min(k >> 1, SORT_CAP >> 1); j > 0; j >>= 1) {
          for (int i = threadIdx.x; i < SORT_CAP; i += GL_T) {
            const int p = i ^ j;
            if (p > i) {
              const uint64_t x = G.skey[i], y = G.skey[p];
              const bool up = ((base + i) & k) == 0;
              if ((x > y) == up) { G.skey[i] = y; G.skey[p] = x; }
            }
          }
          __syncthreads();
        }
      for (int i = threadIdx.x; i < SORT_CAP; i += GL_T)
        pool[base + i] = G.skey[i];
      __threadfence_block();
      __syncthreads();
    }
  };
  blocks(2, SORT_CAP);
  for (int k = 2 * SORT_CAP; k <= np; k <<= 1) {
    for (int j = k >> 1; j >= SORT_CAP; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += GL_T) {
        const int p = i ^ j;
        if (p > i) {
          const uint64_t x = pool[i], y = pool[p];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { pool[i] = y; pool[p] = x; }
        }
      }
      __threadfence_block();
      __syncthreads();
    }
    blocks(k, k);
  }
}

// ---------------------------------------------------------------------------
// Growth of one large halo, in phases:
//   begin    size the radial histogram of the cube of reach W, or close the
//            halo as unclosed when the cube exceeds the box;
//   walk 1   the radial histogram (bins uniform in r^3);
//   bins     its scan, the bins that may hold each threshold's first
//            crossing and the ranges of bins to collect;
//   walk 2   the keys (d2, index) of those ranges;
//   cross    their sort, the exact crossings, the second threshold's ranges
//            and the profile radii;
//   walk 3   profile counts, boundary count, centre-of-mass sums and the
//            keys of the second threshold's ranges;
//   finish   the outputs.
// bins and cross send the halo back to begin with a larger reach when the
// crossing lies beyond the cube, or fail it when its bands exceed the
// collection slice (a re-run launch then grows it with larger slices).
// k_grow_large runs the same steps inline, each halo in one workgroup.  A
// giant halo (cube of at least c.giant_min particles) is left to the staged
// launches k_giant_*, which call the phase functions below: one workgroup
// per giant halo for begin, bins, cross and finish, every workgroup of the
// GPU for its walks, each phase a launch of its own with the halo's state
// kept in global memory between launches.
// ---------------------------------------------------------------------------
enum : int32_t {
  GS_BEGIN = 0,
  GS_HIST = 1,
  GS_COLLECT = 2,
  GS_PROFILE = 3,
  GS_DONE = 4,
  GS_FAILED = 5,
  GS_COUNT = 6  // staged path: the cube's particles counted by a walk
};

constexpr int64_t CH_SEG = 4 * GL_T;  // segments per chunk of the count walk
constexpr int64_t CH_PART = 1 << 15;  // particles per chunk of walks 1-3

// a halo's growth state between phases (scalars, kept in registers; its
// HaloResult and scaled centre are held apart: they are indexed at run time
// and live in memory)
struct GrowVars {
  int64_t h;          // halo
  int32_t e, centre;  // its position in the large list; centre particle
  CellRef cr;
  int32_t W, nb;      // reach (cells); radial bins
  float scale;        // bins per unit of r^3 (cells^3)
  float lo2, hi2;     // distance band of the next walk
  int64_t ncube;      // particles of the cube of reach W
  int64_t k0s;        // rank of the first crossing of delta[0]
  uint64_t keyR;      // key of rank k0s - 1
  double R0, rlast;   // R_delta[0]; outer profile radius
  int32_t ns2;        // ranges of the second threshold
};

// sums of walk 3
struct ProfileSums {
  double sp[3], sv[3];
  int64_t nbd;
};

__device__ inline void grow_start(const SodCfg &c, const float *pos,
                                  const int64_t *ids, const HaloPlan *plan,
                                  const int32_t *large, int32_t e,
                                  GrowVars &V, HaloResult &res,
                                  double *xcs) {
  V.e = e;
  V.h = large[e];
  const HaloPlan pl = plan[V.h];
  V.centre = pl.centre;
  V.cr = ref_of_particle(pos, V.centre, c.nx);
  centre_scaled(pos, V.centre, c.nx, xcs);
  zero_result(res);
  res.centre = V.centre;
  res.centre_id = ids ? ids[V.centre] : (int64_t)V.centre;
  V.W = pl.w0 < 2 ? 2 : pl.w0;
  V.nb = 0;
  V.scale = 0.0f;
  V.lo2 = -1.0f;
  V.hi2 = -2.0f;
  V.ncube = 0;
  V.k0s = 0;
  V.keyR = 0;
  V.R0 = 0.0;
  V.rlast = 0.0;
  V.ns2 = 0;
}

__device__ inline int64_t chunks_of(int64_t nseg) {
  return (nseg + CH_SEG - 1) / CH_SEG;
}

// false, with the halo closed as unclosed, when the cube of reach W
// exceeds the box
__device__ inline bool grow_reach(const SodCfg &c, const GrowVars &V,
                                  HaloResult &res) {
  if (2 * V.W + 1 > c.nx) {
    res.status = PSOD_H_UNCLOSED | PSOD_H_LARGE;
    res.w = V.W;
    return false;
  }
  return true;
}

// Bins of the radial histogram (uniform in r^3) of a cube of ncube
// particles: about 4 cube particles per bin up to the bins that LDS holds,
// at least 16 beyond them, so that fewer particles are counted with global
// atomics (c3 k_grow_large 14.9 -> 14.0 ms; every bin in LDS: 14.2 ms).
__device__ inline int32_t hist_bins(int64_t ncube) {
  int32_t nb = HIST_MIN;
  while ((int64_t)nb * 4 < ncube && nb < HIST_MAX &&
         (nb < HIST_LDS || (int64_t)nb * 16 < ncube))
    nb <<= 1;
  return nb;
}

// the radial histogram of a cube of ncube particles and the band of walk 1
__device__ inline void grow_size(GrowVars &V, int64_t ncube) {
  const float W2 = (float)V.W * (float)V.W;
  V.ncube = ncube;
  const int32_t nb = hist_bins(ncube);
  V.nb = nb;
  V.scale = (float)nb / ((float)V.W * (float)V.W * (float)V.W);
  // every particle of the cube is gathered; those of d2 < W2 are binned
  V.lo2 = -1.0f;
  V.hi2 = W2 * (1.0f + 1e-4f) + 1e-3f;
}

// begin: GS_HIST, or GS_DONE for a halo that does not close in the box
__device__ inline int32_t grow_begin(const SodCfg &c, GrowShared &G,
                                     const int32_t *cell_start, GrowVars &V,
                                     HaloResult &res) {
  if (!grow_reach(c, V, res)) return GS_DONE;
  const CellRef cr = V.cr;
  grow_size(V, cube_count(c, G, cell_start, cr, V.W));
  return GS_HIST;
}

// walk 1 over the segments [q0, q1) of the cube: the particles of d2 < W^2
// to their radial bins, the first HIST_LDS bins counted in lh
__device__ inline void walk_hist(const SodCfg &c, GrowShared &G, int32_t *lh,
                                 int32_t *hist, const int32_t *cell_start,
                                 const float4 *rec, const GrowVars &V,
                                 int64_t q0 = 0, int64_t q1 = INT64_MAX) {
  const float W2 = (float)V.W * (float)V.W;
  const float scale = V.scale;
  const int32_t nb = V.nb;
  const CellRef cr = V.cr;
  cube_walk(c, G, cell_start, rec, cr, V.W, V.lo2, V.hi2,
            [&](float d2, int32_t, const float4 &, int, int, int,
                const int32_t *) {
              if (d2 < W2) {
                const int32_t b = rbin(d2, scale, nb);
                // scopes differ, so that the two adds stay apart: merged
                // into one add on a selected pointer they become a flat
                // atomic, much slower on LDS
                if (b < HIST_LDS)
                  __hip_atomic_fetch_add(&lh[b], 1, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                  __hip_atomic_fetch_add(&hist[b], 1, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
              }
            },
            q0, q1);
}

// walk 2: the keys of the particles whose bin lies in one of the ranges
// [G.rs[r], G.re[r]] (r < nr) to pool[G.off[r] + fill[r]++]
__device__ inline void walk_collect(const SodCfg &c, GrowShared &G,
                                    const int32_t *cell_start,
                                    const float4 *rec, const GrowVars &V,
                                    int32_t nr, int32_t *fill, uint64_t *pool,
                                    int64_t q0 = 0, int64_t q1 = INT64_MAX) {
  const float W2 = (float)V.W * (float)V.W;
  const float scale = V.scale;
  const int32_t nb = V.nb;
  const CellRef cr = V.cr;
  cube_walk(c, G, cell_start, rec, cr, V.W, V.lo2, V.hi2,
            [&](float d2, int32_t idx, const float4 &, int, int, int,
                const int32_t *) {
              if (!(d2 < W2)) return;
              const int32_t b = rbin(d2, scale, nb);
              for (int r = 0; r < nr; r++)
                if (b >= G.rs[r] && b <= G.re[r]) {
                  const int32_t s = atomicAdd(&fill[r], 1);
                  pool[G.off[r] + s] = d2key(d2, idx);
                }
            },
            q0, q1);
}

// walk 3: profile counts per wave in G.pcnt (radii G.prad), the boundary
// count, the centre-of-mass sums of the keys up to keyR (offsets from the
// reference point, velocities when flagv), and the keys of the second
// threshold's ranges G.rs[R2 + j] to pool[G.off[R2 + j] + fill[R2 + j]++]
__device__ inline void walk_profile(const SodCfg &c, GrowShared &G,
                                    const int32_t *cell_start,
                                    const float4 *rec, const float *vel,
                                    const GrowVars &V, int flagv,
                                    int32_t *fill, uint64_t *pool,
                                    ProfileSums &S, int64_t q0 = 0,
                                    int64_t q1 = INT64_MAX) {
  const float W2 = (float)V.W * (float)V.W;
  const float scale = V.scale;
  const int32_t nb = V.nb, ns2 = V.ns2;
  const double R0 = V.R0, rlast = V.rlast;
  const uint64_t keyR = V.keyR;
  const CellRef cr = V.cr;
  // local sums (registers), added to S at the end
  double sp[3] = {0.0, 0.0, 0.0}, sv[3] = {0.0, 0.0, 0.0};
  int64_t nbd = 0;
  cube_walk(c, G, cell_start, rec, cr, V.W, V.lo2, V.hi2,
            [&](float d2, int32_t idx, const float4 &r, int di, int dj,
                int dk, const int32_t *lc) {
              if (!(d2 < W2)) return;
              const double rr = (double)sqrtf(d2);
              if (rr < rlast) {
                int lo = 0, hi = c.nbins - 1;  // first j with rr < prad[j]
                while (lo < hi) {
                  const int mid = (lo + hi) >> 1;
                  if (rr < G.prad[mid]) hi = mid; else lo = mid + 1;
                }
                atomicAdd(&G.pcnt[threadIdx.x >> 6][lo], 1);
              }
              if (fabs(rr - R0) <= c.tol_abs) nbd++;
              if (d2key(d2, idx) <= keyR) {
                sp[0] += rel1d(r.x, lc[0], di, (double)cr.f[0]);
                sp[1] += rel1d(r.y, lc[1], dj, (double)cr.f[1]);
                sp[2] += rel1d(r.z, lc[2], dk, (double)cr.f[2]);
                if (flagv)
                  for (int d = 0; d < 3; d++)
                    sv[d] += (double)vel[3 * (int64_t)idx + d];
              }
              if (ns2 > 0) {
                const int32_t b = rbin(d2, scale, nb);
                for (int j = 0; j < ns2; j++)
                  if (b >= G.rs[R2 + j] && b <= G.re[R2 + j]) {
                    const int32_t s = atomicAdd(&fill[R2 + j], 1);
                    pool[G.off[R2 + j] + s] = d2key(d2, idx);
                  }
              }
            },
            q0, q1);
  S.nbd += nbd;
  for (int d = 0; d < 3; d++) {
    S.sp[d] += sp[d];
    S.sv[d] += sv[d];
  }
}

// bins, on the histogram of walk 1 (hist[0..nb), counted): GS_COLLECT,
// GS_BEGIN (reach doubled) or GS_FAILED (G.need slots beyond pool_cap).
// lh stages the scan.
__device__ inline int32_t grow_bins(const SodCfg &c, GrowShared &G,
                                    int32_t *lh, int32_t *hist, GrowVars &V,
                                    int64_t pool_cap) {
  const int32_t nb = V.nb;
  const float W2 = (float)V.W * (float)V.W;
  // exclusive scan in place: hist[b] = #(bin < b), hist[nb] = total.
  // Rounds of SCAN_ROUND bins, read coalesced into LDS; thread t scans
  // the SCAN_ROUND / GL_T consecutive bins from t * SCAN_ROUND / GL_T.
  {
    constexpr int PER = SCAN_ROUND / GL_T;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    int32_t carry = 0;
    for (int32_t base = 0; base < nb; base += SCAN_ROUND) {
      for (int i = threadIdx.x; i < SCAN_ROUND; i += GL_T)
        lh[i] = ldcg(&hist[base + i]);
      __syncthreads();
      int32_t v[PER], sum = 0;
#pragma unroll
      for (int i = 0; i < PER; i++) {
        v[i] = lh[threadIdx.x * PER + i];
        sum += v[i];
      }
      int32_t tot;
      const int32_t ex = wave_excl(sum, &tot);
      if (l == 0) G.red_i[w] = tot;
      __syncthreads();
      int32_t run = carry + ex, all = 0;
      for (int i = 0; i < GL_T / 64; i++) {
        if (i < w) run += G.red_i[i];
        all += G.red_i[i];
      }
#pragma unroll
      for (int i = 0; i < PER; i++) {
        lh[threadIdx.x * PER + i] = run;
        run += v[i];
      }
      __syncthreads();
      for (int i = threadIdx.x; i < SCAN_ROUND; i += GL_T)
        hist[base + i] = lh[i];
      carry += all;
      __syncthreads();
    }
    if (threadIdx.x == 0) hist[nb] = carry;
    __threadfence();
    __syncthreads();
  }
  const int32_t *C = hist;
  const double Wd = (double)V.W, W3 = Wd * Wd * Wd;
  // ---- bins that may hold each threshold's first crossing --------------
  // A bin is possible for delta when the smallest admissible rank k in it
  // could have od_k <= delta at the bin's outer edge, and definite when
  // it must at the inner edge; the crossing lies in a possible bin no
  // later than the first definite one.  The possible bins are collected
  // as up to SEGCAP segments, each extended down to the previous
  // non-empty bin, which holds rank k-1 when k opens the segment.
  bool ok = true;
  auto bin_k = [&](int32_t b, int64_t *kmin) {
    const int32_t cb = ldcg(&C[b]), ce = ldcg(&C[b + 1]);
    *kmin = max((int64_t)cb + 1, (int64_t)c.nmin_count);
    return *kmin <= ce;
  };
  auto nonempty = [&](int32_t b) { return ldcg(&C[b + 1]) > ldcg(&C[b]); };
  for (int di = 0; di < c.ndelta && ok; di++) {
    const double dlt = c.delta[di];
    auto definite = [&](int32_t b) {
      int64_t k;
      return bin_k(b, &k) && rad3_for(k, c.mpart, dlt) <= edge3_lo(b, nb, W3);
    };
    auto possible = [&](int32_t b) {
      int64_t k;
      return bin_k(b, &k) && rad3_for(k, c.mpart, dlt) < edge3_hi(b, nb, W3);
    };
    const int32_t b2 = block_first(0, nb - 1, definite, G);
    if (b2 == INT32_MAX) { ok = false; break; }
    int32_t last_e = -1, x = 0, ns = 0;
    for (;;) {
      const int32_t s0 = block_first(x, b2, possible, G);
      int32_t e = b2;
      if (ns < SEGCAP - 1) {
        const int32_t q = block_first(
            s0 + 1, b2, [&](int32_t b) { return nonempty(b) && !possible(b); },
            G);
        if (q != INT32_MAX) e = q - 1;
      }
      // a run close behind the previous segment (few keys between)
      // extends it instead of opening a segment
      const bool join =
          ns > 0 && ldcg(&C[s0]) - ldcg(&C[last_e + 1]) <= GAP_KEYS;
      const int32_t xs = (!join && s0 > last_e + 1)
                             ? block_last(last_e + 1, s0 - 1, nonempty, G)
                             : -1;
      if (threadIdx.x == 0) {
        if (join) {
          G.sg_e[di][ns - 1] = e;
        } else {
          G.sg_s[di][ns] = s0;
          G.sg_e[di][ns] = e;
          G.sg_x[di][ns] = xs >= 0 ? xs : s0;
        }
      }
      if (!join) ns++;
      if (e >= b2) break;
      last_e = e;
      x = e + 1;
    }
    if (threadIdx.x == 0) G.nseg[di] = ns;
    __syncthreads();
  }
  if (!ok) {
    V.W = 2 * V.W;
    __syncthreads();
    return GS_BEGIN;
  }
  // ---- disjoint ranges of bins to collect -------------------------------
  if (threadIdx.x == 0) {
    int32_t nr = 0;
    for (int di = 0; di < c.ndelta; di++)
      for (int j = 0; j < G.nseg[di]; j++) {
        int32_t a = G.sg_x[di][j], t = G.sg_e[di][j];
        // absorb every range that overlaps or touches [a, t]
        int32_t w = 0;
        for (int r = 0; r < nr; r++) {
          if (!(t < G.rs[r] - 1 || a > G.re[r] + 1)) {
            a = min(a, G.rs[r]);
            t = max(t, G.re[r]);
          } else {
            G.rs[w] = G.rs[r];
            G.re[w] = G.re[r];
            w++;
          }
        }
        G.rs[w] = a;
        G.re[w] = t;
        nr = w + 1;
      }
    // ascending order of the ranges (ranks then ascend with them)
    for (int r = 1; r < nr; r++)
      for (int q = r; q > 0 && G.rs[q - 1] > G.rs[q]; q--) {
        const int32_t ts = G.rs[q], te = G.re[q];
        G.rs[q] = G.rs[q - 1];
        G.re[q] = G.re[q - 1];
        G.rs[q - 1] = ts;
        G.re[q - 1] = te;
      }
    G.nr = nr;
    int64_t oo = 0;
    for (int r = 0; r < nr; r++) {
      G.k0[r] = ldcg(&C[G.rs[r]]);
      G.k1[r] = ldcg(&C[G.re[r] + 1]);
      G.off[r] = oo;
      G.fill[r] = 0;
      oo += pool_slots(G.k1[r] - G.k0[r]);
    }
    G.bump = oo;  // a re-run starts the slice afresh
    G.need = oo;
  }
  __syncthreads();
  if (G.need > pool_cap) return GS_FAILED;
  // ---- band of walk 2 ----------------------------------------------------
  const int nr = G.nr;
  float blo2 = -1.0f, bhi2 = -2.0f;  // nothing when there is no range
  if (nr > 0) bin_band(G.rs[0], G.re[nr - 1], nb, V.scale, W2, &blo2, &bhi2);
  V.lo2 = blo2;
  V.hi2 = bhi2;
  return GS_COLLECT;
}

// cross, on the keys of walk 2 in pool: GS_PROFILE, GS_BEGIN (larger
// reach) or GS_FAILED (G.need slots beyond pool_cap)
// Ranges r with bit r of `presorted` set were sorted before the launch.
__device__ inline int32_t grow_cross(const SodCfg &c, GrowShared &G,
                                     const int32_t *hist, uint64_t *pool,
                                     const float *pos, GrowVars &V,
                                     HaloResult &res, const double *xcs,
                                     int64_t pool_cap, uint32_t presorted) {
  const int32_t *C = hist;
  const int32_t nb = V.nb;
  const float scale = V.scale;
  const float W2 = (float)V.W * (float)V.W;
  const double Wd = (double)V.W, W3 = Wd * Wd * Wd;
  const int nr = G.nr;
  for (int r = 0; r < nr; r++)
    if (!((presorted >> r) & 1))
      block_sort_pool(pool + G.off[r], (int32_t)(G.k1[r] - G.k0[r]), G);
  // key of rank k (1-based) among the collected ranges
  auto key_rank = [&](int64_t k) {
    int r = 0;
    while (r + 1 < nr && k > G.k1[r]) r++;
    return pool[G.off[r] + (k - G.k0[r] - 1)];
  };
  auto nonempty = [&](int32_t b) { return ldcg(&C[b + 1]) > ldcg(&C[b]); };
  // ---- exact first crossing ------------------------------------------
  int64_t kst[PSOD_MAXDELTA];
  uint64_t keyR = 0;
  bool ok = true;
  for (int di = 0; di < c.ndelta; di++) {
    int32_t firstk = INT32_MAX;
    for (int j = 0; j < G.nseg[di] && firstk == INT32_MAX; j++) {
      const int64_t klo = max((int64_t)ldcg(&C[G.sg_s[di][j]]) + 1,
                              (int64_t)c.nmin_count);
      const int64_t khi = ldcg(&C[G.sg_e[di][j] + 1]);
      int32_t fk = INT32_MAX;
      for (int64_t k = klo + threadIdx.x; k <= khi; k += GL_T) {
        const double rk = (double)sqrtf(key_d2(key_rank(k)));
        if (od_k(k, c.mpart, rk) <= c.delta[di]) {
          fk = (int32_t)k;
          break;
        }
      }
      firstk = block_min_i32(fk, G);
    }
    kst[di] = firstk == INT32_MAX ? -1 : firstk;
    if (kst[di] < 0) { ok = false; break; }
    const uint64_t k1 = key_rank(kst[di] - 1);
    const double r1 = ref_dist(pos, key_idx(k1), xcs, c.nx);
    const double r2 = ref_dist(pos, key_idx(key_rank(kst[di])), xcs, c.nx);
    finish_delta(c, di, kst[di], r1, r2, res);
    if (di == 0) keyR = k1;
  }
  if (!ok) {
    V.W = 2 * V.W;
    __syncthreads();
    return GS_BEGIN;
  }
  const double R0 = res.radius[0];
  if (c.rmax_frac * R0 > Wd) {
    const int32_t need_w = (int32_t)ceil(c.rmax_frac * R0) + 1;
    V.W = max(2 * V.W, need_w);
    __syncthreads();
    return GS_BEGIN;
  }
  // ---- second-threshold bins (last k above delta2) ----------------------
  // A bin may hold a rank k <= k0s with od_k > delta2 when its largest
  // such rank could at the bin's inner edge (possible), and must at the
  // outer edge (definite).  The answer lies in a possible bin between
  // the last definite bin (or bin 0) and the last possible bin; those
  // bins are collected as up to SEGCAP segments found downward, the
  // lowest one extended to the floor.
  const int64_t k0s = kst[0];
  int ns2 = 0;
  if (c.delta2 > 0.0) {
    auto kmax2 = [&](int32_t b, int64_t *kmax) {
      const int32_t cb = ldcg(&C[b]), ce = ldcg(&C[b + 1]);
      *kmax = min((int64_t)ce, k0s);
      return *kmax >= max((int64_t)cb + 1, (int64_t)2);
    };
    auto possible2 = [&](int32_t b) {
      int64_t k;
      return kmax2(b, &k) && rad3_for(k, c.mpart, c.delta2) > edge3_lo(b, nb, W3);
    };
    auto definite2 = [&](int32_t b) {
      int64_t k;
      return kmax2(b, &k) && rad3_for(k, c.mpart, c.delta2) >= edge3_hi(b, nb, W3);
    };
    // ranks above k0s do not count: no bin beyond the one of rank k0s
    const int32_t bk0 = rbin(key_d2(key_rank(k0s)), scale, nb);
    const int32_t b3 = block_last(0, bk0, possible2, G);
    if (b3 >= 0) {
      const int32_t b4 = block_last(0, b3, definite2, G);
      const int32_t floor2 = b4 >= 0 ? b4 : 0;
      int32_t x = b3;
      while (x >= floor2) {
        const int32_t e = block_last(floor2, x, possible2, G);
        if (e < 0) break;
        int32_t s0 = floor2;
        if (ns2 < SEGCAP - 1) {
          const int32_t q = block_last(
              floor2, e - 1,
              [&](int32_t b) { return nonempty(b) && !possible2(b); }, G);
          if (q >= 0) s0 = q + 1;
        }
        // join the segment above when few keys lie between
        const bool join = ns2 > 0 && ldcg(&C[x + 1]) - ldcg(&C[e + 1]) <= GAP_KEYS;
        if (threadIdx.x == 0) {
          if (join) {
            G.rs[R2 + ns2 - 1] = s0;
          } else {
            G.rs[R2 + ns2] = s0;
            G.re[R2 + ns2] = e;
          }
        }
        if (!join) ns2++;
        if (s0 <= floor2) break;
        x = s0 - 1;
      }
    }
    if (threadIdx.x == 0) {
      int64_t o = G.bump;
      for (int j = 0; j < ns2; j++) {
        G.k0[R2 + j] = ldcg(&C[G.rs[R2 + j]]);
        G.k1[R2 + j] = ldcg(&C[G.re[R2 + j] + 1]);
        G.off[R2 + j] = o;
        G.fill[R2 + j] = 0;
        o += pool_slots(G.k1[R2 + j] - G.k0[R2 + j]);
      }
      G.bump = o;
      G.need = o;
    }
  }
  // profile radii
  for (int j = threadIdx.x; j < c.nbins; j += GL_T) {
    G.prad[j] = prof_radius(c, R0, j);
    for (int w = 0; w < GL_T / 64; w++) G.pcnt[w][j] = 0;
  }
  __syncthreads();
  if (G.need > pool_cap) return GS_FAILED;
  // ---- band of walk 3 ----------------------------------------------------
  const double rlast = G.prad[c.nbins - 1];
  // farthest particle walk 3 uses: profile, boundary band, centre of mass
  // (keys up to keyR) and the second threshold's bins
  float phi2;
  {
    const double rm = fmax(rlast, R0 + c.tol_abs);
    float h2 = fmaxf((float)(rm * rm * (1.0 + 1e-4)) + 1e-6f, key_d2(keyR));
    for (int j = 0; j < ns2; j++) {
      float l2, u2;
      bin_band(G.rs[R2 + j], G.re[R2 + j], nb, scale, W2, &l2, &u2);
      h2 = fmaxf(h2, u2);
    }
    phi2 = fminf(h2, W2 * (1.0f + 1e-4f) + 1e-3f);
  }
  V.k0s = k0s;
  V.keyR = keyR;
  V.R0 = R0;
  V.rlast = rlast;
  V.ns2 = ns2;
  V.lo2 = -1.0f;
  V.hi2 = phi2;
  return GS_PROFILE;
}

// finish, on the sums of walk 3 (S, block totals) and its profile counts
// per bin in G.pcnt[0]: the halo's outputs in V.res, its cumulative
// profile to out
__device__ inline void grow_finish(const SodCfg &c, GrowShared &G,
                                   uint64_t *pool, const float *pos,
                                   const GrowVars &V, HaloResult &res,
                                   const double *xcs, const ProfileSums &S,
                                   psod_halo_out out) {
  const CellRef &cr = V.cr;
  const int32_t nin = res.count[0];
  const double xc[3] = {(double)cr.c[0] + (double)cr.f[0],
                        (double)cr.c[1] + (double)cr.f[1],
                        (double)cr.c[2] + (double)cr.f[2]};
  for (int d = 0; d < 3; d++) {
    double cp = (xc[d] + S.sp[d] / (double)nin) / (double)c.nx;
    cp -= floor(cp);
    res.com_pos[d] = cp;
    res.com_vel[d] = S.sv[d] / (double)nin;
  }
  // cumulative profile
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int j = 0; j < c.nbins; j++) {
      s += G.pcnt[0][j];
      out.profile[V.h * c.nbins + j] = s;
    }
  }
  // ---- delta2 --------------------------------------------------------
  if (c.delta2 > 0.0) {
    int64_t best = 0;
    uint64_t bkey = 0;
    if (V.ns2 > 0) {
      __threadfence_block();
      __syncthreads();
      // segments from the top down: the first one holding a rank above
      // delta2 holds the last such rank
      for (int j = 0; j < V.ns2 && best == 0; j++) {
        const int32_t n2 = (int32_t)(G.k1[R2 + j] - G.k0[R2 + j]);
        uint64_t *p2 = pool + G.off[R2 + j];
        block_sort_pool(p2, n2, G);
        const int64_t kb = G.k0[R2 + j];
        int32_t bestk = -1;
        for (int32_t t = threadIdx.x; t < n2; t += GL_T) {
          const int64_t k = kb + t + 1;
          if (k < 2 || k > V.k0s) continue;
          const double rk = (double)sqrtf(key_d2(p2[t]));
          if (od_k(k, c.mpart, rk) > c.delta2) bestk = max(bestk, (int32_t)k);
        }
        bestk = block_max_i32(bestk, G);
        if (bestk > 0) {
          best = bestk;
          bkey = p2[bestk - kb - 1];
        }
      }
    }
    res.count2 = (int32_t)best;
    if (best > 0) {
      const double rb = ref_dist(pos, key_idx(bkey), xcs, c.nx);
      res.overdens2 = rb > 1e-30 ? od_k(best, c.mpart, rb) : 0.0;
    }
  }
  res.w = V.W;
  res.ngather = (int32_t)V.ncube;
  res.nboundary = (int32_t)S.nbd;
  bool acc = (double)res.count[0] * c.mpart >= c.mmin;
  for (int d = 0; d < 3; d++) acc = acc && xc[d] > 0.0 && xc[d] <= c.nx;
  res.status = PSOD_H_CLOSED | PSOD_H_LARGE | (acc ? PSOD_H_ACCEPTED : 0);
}

// Stores a halo's outputs (thread 0); a failed one (its bands beyond the
// slice) goes to the retry list with the slice size it needs.
__device__ inline void grow_close(const SodCfg &c, int32_t *counters,
                                  int32_t *retry,
                                  unsigned long long *needmax, int64_t need,
                                  const GrowVars &V, HaloResult &res,
                                  bool failed, psod_halo_out out) {
  if (failed) res.status = PSOD_H_LARGE;
  if (threadIdx.x == 0) {
    if (failed) {
      const int32_t k = atomicAdd(&counters[4], 1);
      if (retry) retry[k] = V.e;
      if (needmax) atomicMax(needmax, (unsigned long long)need);
    }
    if (res.status & PSOD_H_UNCLOSED) atomicAdd(&counters[5], 1);
    store_result(out, V.h, c.nbins, res);
  }
}

// Growth of a halo whose cube of reach W holds n <= LIGHT_CAP particles:
// the keys (d2, index) of the whole cube are sorted in LDS (K), and the
// crossings, the second threshold, the profile, the boundary count and the
// centre of mass are read off the sorted keys, as the one-wave path does
// for its halos.  Rank k (1-based) is key K[k - 1]: the ranks and radii of
// the binned steps, so the outputs are theirs.  Only the reach that closes
// the halo (w, ngather) can differ: the binned steps also double it when no
// bin surely holds a crossing.  Returns 0 with res and the profile written,
// else the reach to try next.  A halo whose cube held a few thousand
// particles took 150-190 us through the bins and the three walks.
static_assert(LIGHT_CAP <= HIST_LDS / 2, "light keys live in the bins' LDS");

__device__ __forceinline__ int32_t
grow_light(const SodCfg &c, GrowShared &G, uint64_t *K,
           const int32_t *cell_start, const float4 *rec, const float *pos,
           const float *vel, const CellRef &cr, int32_t W, int32_t n,
           const double *xcs, int64_t h, const psod_halo_out &out,
           HaloResult &res) {
  cube_walk(c, G, cell_start, rec, cr, W, -1.0f, INFINITY,
            [&](float d2, int32_t idx, const float4 &, int, int, int,
                const int32_t *, int64_t j) { K[j] = d2key(d2, idx); });
  const int32_t np = next_pow2(n < 2 ? 2 : n);
  for (int32_t j = n + threadIdx.x; j < np; j += GL_T) K[j] = ~0ull;
  __syncthreads();
  lds_bitonic<GL_T>(K, np);
  // The ranks, the reach test and the growth of the reach are those of the
  // binned steps: only particles with d2 < W^2 are ranked (they sort before
  // the others), a threshold not crossed among them doubles the reach, and
  // a profile reaching beyond the cube extends it.
  const double Wd = (double)W;
  const float W2 = (float)W * (float)W;
  int64_t kst[PSOD_MAXDELTA];
  for (int di = 0; di < c.ndelta; di++) {
    // first rank k >= nmin_count with od_k <= delta, unless a particle of
    // d2 >= W^2 comes first
    int32_t fk = INT32_MAX;
    for (int32_t base = c.nmin_count; base <= n && fk == INT32_MAX;
         base += GL_T) {
      const int32_t k = base + (int32_t)threadIdx.x;
      int32_t m = INT32_MAX;
      if (k <= n) {
        const float d2 = key_d2(K[k - 1]);
        if (!(d2 < W2) || od_k(k, c.mpart, (double)sqrtf(d2)) <= c.delta[di])
          m = k;
      }
      fk = block_min_i32(m, G);
    }
    if (fk == INT32_MAX || !(key_d2(K[fk - 1]) < W2)) {
      __syncthreads();
      return 2 * W;
    }
    kst[di] = fk;
  }
  for (int di = 0; di < c.ndelta; di++) {
    const int64_t ks = kst[di];
    finish_delta(c, di, ks, ref_dist(pos, key_idx(K[ks - 2]), xcs, c.nx),
                 ref_dist(pos, key_idx(K[ks - 1]), xcs, c.nx), res);
  }
  const double R0 = res.radius[0];
  if (c.rmax_frac * R0 > Wd) {
    __syncthreads();
    return max(2 * W, (int32_t)ceil(c.rmax_frac * R0) + 1);
  }
  // second threshold: the last rank k in [2, kst[0]] above delta2
  if (c.delta2 > 0.0) {
    int32_t best = 0;
    for (int32_t top = (int32_t)kst[0]; top >= 2 && best == 0; top -= GL_T) {
      const int32_t k = top - (int32_t)threadIdx.x;
      int32_t m = -1;
      if (k >= 2) {
        const double r = (double)sqrtf(key_d2(K[k - 1]));
        if (od_k(k, c.mpart, r) > c.delta2) m = k;
      }
      best = max(0, block_max_i32(m, G));
    }
    res.count2 = best;
    if (best > 0) {
      const double rb = ref_dist(pos, key_idx(K[best - 1]), xcs, c.nx);
      res.overdens2 = rb > 1e-30 ? od_k(best, c.mpart, rb) : 0.0;
    }
  }
  // profile: particles with r below each profile radius
  for (int j = threadIdx.x; j < c.nbins; j += GL_T) {
    const double rj = prof_radius(c, R0, j);
    int32_t lo = 0, hi = n;  // first rank index with r >= rj
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if ((double)sqrtf(key_d2(K[mid])) < rj) lo = mid + 1;
      else hi = mid;
    }
    out.profile[h * c.nbins + j] = lo;
  }
  // boundary particles; centre of mass of the count[0] closest
  int64_t nbd = 0;
  double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
  const int32_t nin = res.count[0];
  const bool flagv = (c.flags & PSOD_F_VELOCITY) && vel;
  const double xc[3] = {(double)cr.c[0] + (double)cr.f[0],
                        (double)cr.c[1] + (double)cr.f[1],
                        (double)cr.c[2] + (double)cr.f[2]};
  for (int32_t k = threadIdx.x; k < n; k += GL_T) {
    const uint64_t key = K[k];
    const float d2 = key_d2(key);
    if (d2 < W2 && fabs((double)sqrtf(d2) - R0) <= c.tol_abs) nbd++;
    if (k < nin) {
      const int32_t pj = key_idx(key);
      for (int d = 0; d < 3; d++) {
        int32_t cj;
        float fj;
        split_coord(pos[3 * (int64_t)pj + d], c.nx, &cj, &fj);
        double dd = ((double)cj + (double)fj) - xc[d];
        if (dd > 0.5 * c.nx) dd -= c.nx;
        if (dd < -0.5 * c.nx) dd += c.nx;
        sp[d] += dd;
        if (flagv) sv[d] += (double)vel[3 * (int64_t)pj + d];
      }
    }
  }
  nbd = block_sum_i64(nbd, G);
  for (int d = 0; d < 3; d++) {
    const double s = block_sum_f64(sp[d], G);
    const double v = block_sum_f64(sv[d], G);
    double cp = (xc[d] + s / (double)nin) / (double)c.nx;
    cp -= floor(cp);
    res.com_pos[d] = cp;
    res.com_vel[d] = v / (double)nin;
  }
  res.w = W;
  res.ngather = n;
  res.nboundary = (int32_t)nbd;
  bool acc = (double)res.count[0] * c.mpart >= c.mmin;
  for (int d = 0; d < 3; d++) acc = acc && xc[d] > 0.0 && xc[d] <= c.nx;
  res.status = PSOD_H_CLOSED | PSOD_H_LARGE | (acc ? PSOD_H_ACCEPTED : 0);
  __syncthreads();
  return 0;
}

__global__ void __launch_bounds__(GL_T)
k_grow_large(SodCfg c, const int32_t *__restrict__ cell_start,
             const float4 *__restrict__ rec, const float *__restrict__ pos,
             const float *__restrict__ vel, const int64_t *__restrict__ ids,
             HaloPlan *__restrict__ plan, const int32_t *__restrict__ large,
             int32_t *__restrict__ counters, int32_t *__restrict__ hist_all,
             uint64_t *__restrict__ pool_all,
             int64_t pool_cap, const int32_t *__restrict__ elist,
             int32_t *__restrict__ retry,
             unsigned long long *__restrict__ needmax,
             const int32_t *__restrict__ abort_flag,
             int32_t *__restrict__ giants, int32_t giant_cap,
             psod_halo_out out) {
  // Work items are positions e of the large list (or elist[item] when a
  // list of positions is given).  A halo whose bands do not fit the slice
  // is appended to `retry` (when given) with the slice size it needs
  // folded into *needmax, for a second launch with larger slices.  With
  // `giants`, a halo whose cube reaches c.giant_min particles is appended
  // there as (e, W) instead (counted in counters[6], at most giant_cap
  // entries) and left to the staged launches.
  //
  // The phases of the growth run here inline, in one function body: the
  // same steps called as the phase functions above (as the staged launches
  // call them) made this kernel 12% slower at 1024^3 (38 against 33.5 ms,
  // same instruction counts, more wave cycles waiting).  The two forms are
  // held to identical results by the tests of the staged growth.
  __shared__ GrowShared G;
  // inner bins; staging of the scan; the keys of a cube sorted in LDS
  __shared__ __align__(16) int32_t lh[HIST_LDS];
  // a captured pass whose centre search overflowed: the host runs the pass
  // again without the graph (uniform per workgroup, before any barrier)
  if (abort_flag && *abort_flag) return;
  const int32_t nitem = counters[1];
  // this workgroup's slices of the collection pool and of the radial
  // histogram, reused for every halo
  uint64_t *pool = pool_all + (int64_t)blockIdx.x * pool_cap;
  int32_t *hist = hist_all + (int64_t)blockIdx.x * HIST_SLICE;
  for (;;) {
    if (threadIdx.x == 0) G.item = atomicAdd(&counters[3], 1);
    __syncthreads();
    const int32_t item = G.item;
    __syncthreads();
    if (item >= nitem) break;
    const int32_t e = elist ? elist[item] : item;
    const int64_t h = large[e];
    const HaloPlan pl = plan[h];
    const int32_t centre = pl.centre;
    const CellRef cr = ref_of_particle(pos, centre, c.nx);
    HaloResult res;
    zero_result(res);
    res.centre = centre;
    res.centre_id = ids ? ids[centre] : (int64_t)centre;
    int32_t W = pl.w0 < 2 ? 2 : pl.w0;
    bool failed = false;
    bool deferred = false;
    if (threadIdx.x == 0) G.bump = 0;
    double xcs[3];
    centre_scaled(pos, centre, c.nx, xcs);
    for (;;) {
      if (2 * W + 1 > c.nx) {
        res.status = PSOD_H_UNCLOSED | PSOD_H_LARGE;
        res.w = W;
        break;
      }
      const float W2 = (float)W * (float)W;
      // bins uniform in r^3, about 4 particles of the cube per bin
      const int64_t ncube = cube_count(c, G, cell_start, cr, W);
      if (giants && ncube >= c.giant_min) {
        if (threadIdx.x == 0) {
          const int32_t k = atomicAdd(&counters[6], 1);
          G.item = k < giant_cap ? 1 : 0;
          if (k < giant_cap) {
            giants[2 * k] = e;
            giants[2 * k + 1] = W;
          }
        }
        __syncthreads();
        deferred = G.item != 0;
        __syncthreads();
        if (deferred) break;
      }
      if (ncube <= c.light_cap) {
        const int32_t wn = grow_light(c, G, (uint64_t *)lh, cell_start, rec,
                                      pos, vel, cr, W, (int32_t)ncube, xcs, h,
                                      out, res);
        if (wn == 0) {
          if (threadIdx.x == 0) atomicAdd(&counters[7], 1);
          break;
        }
        W = wn;
        continue;
      }
      const int32_t nb = hist_bins(ncube);
      const float scale = (float)nb / ((float)W * (float)W * (float)W);
      const int32_t nl = min(nb, HIST_LDS);
      for (int32_t b = threadIdx.x; b < nl; b += GL_T) lh[b] = 0;
      for (int32_t b = nl + threadIdx.x; b <= nb; b += GL_T) hist[b] = 0;
      __threadfence_block();
      __syncthreads();
      // ---- pass 1: histogram -------------------------------------------
      // every particle of the cube is gathered; those of d2 < W2 are binned
      const int64_t ng = ncube;
      cube_walk(c, G, cell_start, rec, cr, W, -1.0f,
                W2 * (1.0f + 1e-4f) + 1e-3f,
                [&](float d2, int32_t, const float4 &, int, int, int,
                    const int32_t *) {
                  if (d2 < W2) {
                    const int32_t b = rbin(d2, scale, nb);
                    if (b < HIST_LDS) atomicAdd(&lh[b], 1);
                    else atomicAdd(&hist[b], 1);
                  }
                });
      for (int32_t b = threadIdx.x; b < nl; b += GL_T) hist[b] = lh[b];
      __threadfence();
      __syncthreads();
      // exclusive scan in place: hist[b] = #(bin < b), hist[nb] = total.
      // Rounds of SCAN_ROUND bins, read coalesced into LDS; thread t scans
      // the SCAN_ROUND / GL_T consecutive bins from t * SCAN_ROUND / GL_T.
      {
        constexpr int PER = SCAN_ROUND / GL_T;
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        int32_t carry = 0;
        for (int32_t base = 0; base < nb; base += SCAN_ROUND) {
          for (int i = threadIdx.x; i < SCAN_ROUND; i += GL_T)
            lh[i] = ldcg(&hist[base + i]);
          __syncthreads();
          int32_t v[PER], sum = 0;
#pragma unroll
          for (int i = 0; i < PER; i++) {
            v[i] = lh[threadIdx.x * PER + i];
            sum += v[i];
          }
          int32_t tot;
          const int32_t ex = wave_excl(sum, &tot);
          if (l == 0) G.red_i[w] = tot;
          __syncthreads();
          int32_t run = carry + ex, all = 0;
          for (int i = 0; i < GL_T / 64; i++) {
            if (i < w) run += G.red_i[i];
            all += G.red_i[i];
          }
#pragma unroll
          for (int i = 0; i < PER; i++) {
            lh[threadIdx.x * PER + i] = run;
            run += v[i];
          }
          __syncthreads();
          for (int i = threadIdx.x; i < SCAN_ROUND; i += GL_T)
            hist[base + i] = lh[i];
          carry += all;
          __syncthreads();
        }
        if (threadIdx.x == 0) hist[nb] = carry;
        __threadfence();
        __syncthreads();
      }
      const int32_t *C = hist;
      const double Wd = (double)W, W3 = Wd * Wd * Wd;
      // ---- bins that may hold each threshold's first crossing --------------
      // A bin is possible for delta when the smallest admissible rank k in it
      // could have od_k <= delta at the bin's outer edge, and definite when
      // it must at the inner edge; the crossing lies in a possible bin no
      // later than the first definite one.  The possible bins are collected
      // as up to SEGCAP segments, each extended down to the previous
      // non-empty bin, which holds rank k-1 when k opens the segment.
      bool ok = true;
      auto bin_k = [&](int32_t b, int64_t *kmin) {
        const int32_t cb = ldcg(&C[b]), ce = ldcg(&C[b + 1]);
        *kmin = max((int64_t)cb + 1, (int64_t)c.nmin_count);
        return *kmin <= ce;
      };
      auto nonempty = [&](int32_t b) { return ldcg(&C[b + 1]) > ldcg(&C[b]); };
      for (int di = 0; di < c.ndelta && ok; di++) {
        const double dlt = c.delta[di];
        auto definite = [&](int32_t b) {
          int64_t k;
          return bin_k(b, &k) && rad3_for(k, c.mpart, dlt) <= edge3_lo(b, nb, W3);
        };
        auto possible = [&](int32_t b) {
          int64_t k;
          return bin_k(b, &k) && rad3_for(k, c.mpart, dlt) < edge3_hi(b, nb, W3);
        };
        const int32_t b2 = block_first(0, nb - 1, definite, G);
        if (b2 == INT32_MAX) { ok = false; break; }
        int32_t last_e = -1, x = 0, ns = 0;
        for (;;) {
          const int32_t s0 = block_first(x, b2, possible, G);
          int32_t e = b2;
          if (ns < SEGCAP - 1) {
            const int32_t q = block_first(
                s0 + 1, b2, [&](int32_t b) { return nonempty(b) && !possible(b); },
                G);
            if (q != INT32_MAX) e = q - 1;
          }
          // a run close behind the previous segment (few keys between)
          // extends it instead of opening a segment
          const bool join =
              ns > 0 && ldcg(&C[s0]) - ldcg(&C[last_e + 1]) <= GAP_KEYS;
          const int32_t xs = (!join && s0 > last_e + 1)
                                 ? block_last(last_e + 1, s0 - 1, nonempty, G)
                                 : -1;
          if (threadIdx.x == 0) {
            if (join) {
              G.sg_e[di][ns - 1] = e;
            } else {
              G.sg_s[di][ns] = s0;
              G.sg_e[di][ns] = e;
              G.sg_x[di][ns] = xs >= 0 ? xs : s0;
            }
          }
          if (!join) ns++;
          if (e >= b2) break;
          last_e = e;
          x = e + 1;
        }
        if (threadIdx.x == 0) G.nseg[di] = ns;
        __syncthreads();
      }
      if (!ok) {
        W = 2 * W;
        __syncthreads();
        continue;
      }
      // ---- disjoint ranges of bins to collect -------------------------------
      if (threadIdx.x == 0) {
        int32_t nr = 0;
        for (int di = 0; di < c.ndelta; di++)
          for (int j = 0; j < G.nseg[di]; j++) {
            int32_t a = G.sg_x[di][j], t = G.sg_e[di][j];
            // absorb every range that overlaps or touches [a, t]
            int32_t w = 0;
            for (int r = 0; r < nr; r++) {
              if (!(t < G.rs[r] - 1 || a > G.re[r] + 1)) {
                a = min(a, G.rs[r]);
                t = max(t, G.re[r]);
              } else {
                G.rs[w] = G.rs[r];
                G.re[w] = G.re[r];
                w++;
              }
            }
            G.rs[w] = a;
            G.re[w] = t;
            nr = w + 1;
          }
        // ascending order of the ranges (ranks then ascend with them)
        for (int r = 1; r < nr; r++)
          for (int q = r; q > 0 && G.rs[q - 1] > G.rs[q]; q--) {
            const int32_t ts = G.rs[q], te = G.re[q];
            G.rs[q] = G.rs[q - 1];
            G.re[q] = G.re[q - 1];
            G.rs[q - 1] = ts;
            G.re[q - 1] = te;
          }
        G.nr = nr;
        int64_t oo = 0;
        for (int r = 0; r < nr; r++) {
          G.k0[r] = ldcg(&C[G.rs[r]]);
          G.k1[r] = ldcg(&C[G.re[r] + 1]);
          G.off[r] = oo;
          G.fill[r] = 0;
          oo += pool_slots(G.k1[r] - G.k0[r]);
        }
        G.bump = oo;  // a re-run starts the slice afresh
        G.need = oo;
      }
      __syncthreads();
      const int nr = G.nr;
      if (G.need > pool_cap) { failed = true; break; }
      // ---- pass 2: collect the ranges ---------------------------------------
      float blo2 = -1.0f, bhi2 = -2.0f;  // nothing when there is no range
      if (nr > 0) bin_band(G.rs[0], G.re[nr - 1], nb, scale, W2, &blo2, &bhi2);
      cube_walk(c, G, cell_start, rec, cr, W, blo2, bhi2,
                [&](float d2, int32_t idx, const float4 &, int, int, int,
                    const int32_t *) {
                  if (!(d2 < W2)) return;
                  const int32_t b = rbin(d2, scale, nb);
                  for (int r = 0; r < nr; r++)
                    if (b >= G.rs[r] && b <= G.re[r]) {
                      const int32_t s = atomicAdd(&G.fill[r], 1);
                      pool[G.off[r] + s] = d2key(d2, idx);
                    }
                });
      __threadfence_block();
      __syncthreads();
      for (int r = 0; r < nr; r++)
        block_sort_pool(pool + G.off[r], (int32_t)(G.k1[r] - G.k0[r]), G);
      // key of rank k (1-based) among the collected ranges
      auto key_rank = [&](int64_t k) {
        int r = 0;
        while (r + 1 < nr && k > G.k1[r]) r++;
        return pool[G.off[r] + (k - G.k0[r] - 1)];
      };
      // ---- exact first crossing ------------------------------------------
      int64_t kst[PSOD_MAXDELTA];
      uint64_t keyR = 0;
      for (int di = 0; di < c.ndelta; di++) {
        int32_t firstk = INT32_MAX;
        for (int j = 0; j < G.nseg[di] && firstk == INT32_MAX; j++) {
          const int64_t klo = max((int64_t)ldcg(&C[G.sg_s[di][j]]) + 1,
                                  (int64_t)c.nmin_count);
          const int64_t khi = ldcg(&C[G.sg_e[di][j] + 1]);
          int32_t fk = INT32_MAX;
          for (int64_t k = klo + threadIdx.x; k <= khi; k += GL_T) {
            const double rk = (double)sqrtf(key_d2(key_rank(k)));
            if (od_k(k, c.mpart, rk) <= c.delta[di]) {
              fk = (int32_t)k;
              break;
            }
          }
          firstk = block_min_i32(fk, G);
        }
        kst[di] = firstk == INT32_MAX ? -1 : firstk;
        if (kst[di] < 0) { ok = false; break; }
        const uint64_t k1 = key_rank(kst[di] - 1);
        const double r1 = ref_dist(pos, key_idx(k1), xcs, c.nx);
        const double r2 = ref_dist(pos, key_idx(key_rank(kst[di])), xcs, c.nx);
        finish_delta(c, di, kst[di], r1, r2, res);
        if (di == 0) keyR = k1;
      }
      if (!ok) {
        W = 2 * W;
        __syncthreads();
        continue;
      }
      const double R0 = res.radius[0];
      if (c.rmax_frac * R0 > Wd) {
        const int32_t need_w = (int32_t)ceil(c.rmax_frac * R0) + 1;
        W = max(2 * W, need_w);
        __syncthreads();
        continue;
      }
      // ---- second-threshold bins (last k above delta2) ----------------------
      // A bin may hold a rank k <= k0s with od_k > delta2 when its largest
      // such rank could at the bin's inner edge (possible), and must at the
      // outer edge (definite).  The answer lies in a possible bin between
      // the last definite bin (or bin 0) and the last possible bin; those
      // bins are collected as up to SEGCAP segments found downward, the
      // lowest one extended to the floor.
      const int64_t k0s = kst[0];
      int ns2 = 0;
      if (c.delta2 > 0.0) {
        auto kmax2 = [&](int32_t b, int64_t *kmax) {
          const int32_t cb = ldcg(&C[b]), ce = ldcg(&C[b + 1]);
          *kmax = min((int64_t)ce, k0s);
          return *kmax >= max((int64_t)cb + 1, (int64_t)2);
        };
        auto possible2 = [&](int32_t b) {
          int64_t k;
          return kmax2(b, &k) && rad3_for(k, c.mpart, c.delta2) > edge3_lo(b, nb, W3);
        };
        auto definite2 = [&](int32_t b) {
          int64_t k;
          return kmax2(b, &k) && rad3_for(k, c.mpart, c.delta2) >= edge3_hi(b, nb, W3);
        };
        // ranks above k0s do not count: no bin beyond the one of rank k0s
        const int32_t bk0 = rbin(key_d2(key_rank(k0s)), scale, nb);
        const int32_t b3 = block_last(0, bk0, possible2, G);
        if (b3 >= 0) {
          const int32_t b4 = block_last(0, b3, definite2, G);
          const int32_t floor2 = b4 >= 0 ? b4 : 0;
          int32_t x = b3;
          while (x >= floor2) {
            const int32_t e = block_last(floor2, x, possible2, G);
            if (e < 0) break;
            int32_t s0 = floor2;
            if (ns2 < SEGCAP - 1) {
              const int32_t q = block_last(
                  floor2, e - 1,
                  [&](int32_t b) { return nonempty(b) && !possible2(b); }, G);
              if (q >= 0) s0 = q + 1;
            }
            // join the segment above when few keys lie between
            const bool join = ns2 > 0 && ldcg(&C[x + 1]) - ldcg(&C[e + 1]) <= GAP_KEYS;
            if (threadIdx.x == 0) {
              if (join) {
                G.rs[R2 + ns2 - 1] = s0;
              } else {
                G.rs[R2 + ns2] = s0;
                G.re[R2 + ns2] = e;
              }
            }
            if (!join) ns2++;
            if (s0 <= floor2) break;
            x = s0 - 1;
          }
        }
        if (threadIdx.x == 0) {
          int64_t o = G.bump;
          for (int j = 0; j < ns2; j++) {
            G.k0[R2 + j] = ldcg(&C[G.rs[R2 + j]]);
            G.k1[R2 + j] = ldcg(&C[G.re[R2 + j] + 1]);
            G.off[R2 + j] = o;
            G.fill[R2 + j] = 0;
            o += pool_slots(G.k1[R2 + j] - G.k0[R2 + j]);
          }
          G.bump = o;
          G.need = o;
        }
      }
      // profile radii
      for (int j = threadIdx.x; j < c.nbins; j += GL_T) {
        G.prad[j] = prof_radius(c, R0, j);
        for (int w = 0; w < GL_T / 64; w++) G.pcnt[w][j] = 0;
      }
      __syncthreads();
      if (G.need > pool_cap) { failed = true; break; }
      // ---- pass 3: profile, centre of mass, boundary, band of delta2 -------
      const double rlast = G.prad[c.nbins - 1];
      double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
      int64_t nbd = 0;
      const int flagv = (c.flags & PSOD_F_VELOCITY) && vel;
      // farthest particle this pass uses: profile, boundary band, centre of
      // mass (keys up to keyR) and the second threshold's bins
      float phi2;
      {
        const double rm = fmax(rlast, R0 + c.tol_abs);
        float h2 = fmaxf((float)(rm * rm * (1.0 + 1e-4)) + 1e-6f, key_d2(keyR));
        for (int j = 0; j < ns2; j++) {
          float l2, u2;
          bin_band(G.rs[R2 + j], G.re[R2 + j], nb, scale, W2, &l2, &u2);
          h2 = fmaxf(h2, u2);
        }
        phi2 = fminf(h2, W2 * (1.0f + 1e-4f) + 1e-3f);
      }
      cube_walk(c, G, cell_start, rec, cr, W, -1.0f, phi2,
                [&](float d2, int32_t idx, const float4 &r, int di, int dj,
                    int dk, const int32_t *lc) {
                  if (!(d2 < W2)) return;
                  const double rr = (double)sqrtf(d2);
                  if (rr < rlast) {
                    int lo = 0, hi = c.nbins - 1;  // first j with rr < prad[j]
                    while (lo < hi) {
                      const int mid = (lo + hi) >> 1;
                      if (rr < G.prad[mid]) hi = mid; else lo = mid + 1;
                    }
                    atomicAdd(&G.pcnt[threadIdx.x >> 6][lo], 1);
                  }
                  if (fabs(rr - R0) <= c.tol_abs) nbd++;
                  if (d2key(d2, idx) <= keyR) {
                    sp[0] += rel1d(r.x, lc[0], di, (double)cr.f[0]);
                    sp[1] += rel1d(r.y, lc[1], dj, (double)cr.f[1]);
                    sp[2] += rel1d(r.z, lc[2], dk, (double)cr.f[2]);
                    if (flagv)
                      for (int d = 0; d < 3; d++)
                        sv[d] += (double)vel[3 * (int64_t)idx + d];
                  }
                  if (ns2 > 0) {
                    const int32_t b = rbin(d2, scale, nb);
                    for (int j = 0; j < ns2; j++)
                      if (b >= G.rs[R2 + j] && b <= G.re[R2 + j]) {
                        const int32_t s = atomicAdd(&G.fill[R2 + j], 1);
                        pool[G.off[R2 + j] + s] = d2key(d2, idx);
                      }
                  }
                });
      nbd = block_sum_i64(nbd, G);
      const int32_t nin = res.count[0];
      const double xc[3] = {(double)cr.c[0] + (double)cr.f[0],
                            (double)cr.c[1] + (double)cr.f[1],
                            (double)cr.c[2] + (double)cr.f[2]};
      for (int d = 0; d < 3; d++) {
        const double s = block_sum_f64(sp[d], G);
        const double v = block_sum_f64(sv[d], G);
        double cp = (xc[d] + s / (double)nin) / (double)c.nx;
        cp -= floor(cp);
        res.com_pos[d] = cp;
        res.com_vel[d] = v / (double)nin;
      }
      // cumulative profile
      if (threadIdx.x == 0) {
        int32_t s = 0;
        for (int j = 0; j < c.nbins; j++) {
          for (int w = 0; w < GL_T / 64; w++) s += G.pcnt[w][j];
          out.profile[h * c.nbins + j] = s;
        }
      }
      // ---- delta2 --------------------------------------------------------
      if (c.delta2 > 0.0) {
        int64_t best = 0;
        uint64_t bkey = 0;
        if (ns2 > 0) {
          __threadfence_block();
          __syncthreads();
          // segments from the top down: the first one holding a rank above
          // delta2 holds the last such rank
          for (int j = 0; j < ns2 && best == 0; j++) {
            const int32_t n2 = (int32_t)(G.k1[R2 + j] - G.k0[R2 + j]);
            uint64_t *p2 = pool + G.off[R2 + j];
            block_sort_pool(p2, n2, G);
            const int64_t kb = G.k0[R2 + j];
            int32_t bestk = -1;
            for (int32_t t = threadIdx.x; t < n2; t += GL_T) {
              const int64_t k = kb + t + 1;
              if (k < 2 || k > k0s) continue;
              const double rk = (double)sqrtf(key_d2(p2[t]));
              if (od_k(k, c.mpart, rk) > c.delta2) bestk = max(bestk, (int32_t)k);
            }
            bestk = block_max_i32(bestk, G);
            if (bestk > 0) {
              best = bestk;
              bkey = p2[bestk - kb - 1];
            }
          }
        }
        res.count2 = (int32_t)best;
        if (best > 0) {
          const double rb = ref_dist(pos, key_idx(bkey), xcs, c.nx);
          res.overdens2 = rb > 1e-30 ? od_k(best, c.mpart, rb) : 0.0;
        }
      }
      res.w = W;
      res.ngather = (int32_t)ng;
      res.nboundary = (int32_t)nbd;
      bool acc = (double)res.count[0] * c.mpart >= c.mmin;
      for (int d = 0; d < 3; d++) acc = acc && xc[d] > 0.0 && xc[d] <= c.nx;
      res.status = PSOD_H_CLOSED | PSOD_H_LARGE | (acc ? PSOD_H_ACCEPTED : 0);
      break;
    }
    if (failed) res.status = PSOD_H_LARGE;  // band beyond the slice
    if (threadIdx.x == 0 && !deferred) {
      if (failed) {
        const int32_t k = atomicAdd(&counters[4], 1);
        if (retry) retry[k] = e;
        if (needmax) atomicMax(needmax, (unsigned long long)G.need);
      }
      if (res.status & PSOD_H_UNCLOSED) atomicAdd(&counters[5], 1);
      store_result(out, h, c.nbins, res);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Staged growth of giant halos.  GiantState g holds the state of giant halo
// g between launches: its GrowVars, its next phase, and the fields of
// GrowShared that carry from one phase to the next (ranges, segments, slot
// counters, profile radii), plus the sums the walks add up over
// workgroups.  Giant g has its own histogram slice (HIST_SLICE) and
// collection slice (pool_cap keys).  Every phase kernel skips the giants
// that are not in its phase, so the host runs cycles of begin, count walk,
// size, walk 1, bins, walk 2, cross, walk 3, finish until every giant is
// done or failed; a giant whose crossing lies beyond its cube goes round
// again with a larger reach.  No launch waits on another workgroup: the
// phases meet only at launch boundaries.
// ---------------------------------------------------------------------------
constexpr int32_t GIANT_CAP = 256;  // giant halos staged per pass

struct GiantState {
  GrowVars V;
  HaloResult res;
  double xcs[3];
  int32_t step;  // next phase
  int32_t nch;   // chunks of the phase's walk
  int32_t cseg;  // segments of the cube
  int64_t segoff;  // first of its per-segment counts and their prefix sums
  unsigned long long ncube;  // particles of the cube (the count walk)
  int64_t need, bump;
  int32_t nr;
  uint32_t presorted;  // ranges sorted by the device-wide sort (bit r)
  int32_t fill[MAXRANGE], rs[MAXRANGE], re[MAXRANGE];
  int64_t off[MAXRANGE], k0[MAXRANGE], k1[MAXRANGE];
  int32_t sg_s[PSOD_MAXDELTA][SEGCAP], sg_e[PSOD_MAXDELTA][SEGCAP],
      sg_x[PSOD_MAXDELTA][SEGCAP], nseg[PSOD_MAXDELTA];
  double prad[257];
  // sums of walk 3
  int32_t pcnt[257];
  double sp[3], sv[3];
  unsigned long long nbd;
};

// GrowShared fields of giant T to LDS (all threads; ends with a barrier)
__device__ inline void giant_load(GrowShared &G, const GiantState &T) {
  for (int i = threadIdx.x; i < MAXRANGE; i += GL_T) {
    G.fill[i] = T.fill[i];
    G.rs[i] = T.rs[i];
    G.re[i] = T.re[i];
    G.off[i] = T.off[i];
    G.k0[i] = T.k0[i];
    G.k1[i] = T.k1[i];
  }
  for (int i = threadIdx.x; i < PSOD_MAXDELTA * SEGCAP; i += GL_T) {
    const int di = i / SEGCAP, j = i % SEGCAP;
    G.sg_s[di][j] = T.sg_s[di][j];
    G.sg_e[di][j] = T.sg_e[di][j];
    G.sg_x[di][j] = T.sg_x[di][j];
  }
  for (int i = threadIdx.x; i < PSOD_MAXDELTA; i += GL_T) G.nseg[i] = T.nseg[i];
  for (int j = threadIdx.x; j < 257; j += GL_T) G.prad[j] = T.prad[j];
  if (threadIdx.x == 0) {
    G.nr = T.nr;
    G.bump = T.bump;
    G.need = T.need;
  }
  __syncthreads();
}

// the same fields back to giant T (all threads, after a barrier), with
// V (thread 0) and the next phase
__device__ inline void giant_save(const GrowShared &G, GiantState &T,
                                  const GrowVars &V, int32_t step,
                                  int32_t nch) {
  for (int i = threadIdx.x; i < MAXRANGE; i += GL_T) {
    T.fill[i] = G.fill[i];
    T.rs[i] = G.rs[i];
    T.re[i] = G.re[i];
    T.off[i] = G.off[i];
    T.k0[i] = G.k0[i];
    T.k1[i] = G.k1[i];
  }
  for (int i = threadIdx.x; i < PSOD_MAXDELTA * SEGCAP; i += GL_T) {
    const int di = i / SEGCAP, j = i % SEGCAP;
    T.sg_s[di][j] = G.sg_s[di][j];
    T.sg_e[di][j] = G.sg_e[di][j];
    T.sg_x[di][j] = G.sg_x[di][j];
  }
  for (int i = threadIdx.x; i < PSOD_MAXDELTA; i += GL_T) T.nseg[i] = G.nseg[i];
  for (int j = threadIdx.x; j < 257; j += GL_T) T.prad[j] = G.prad[j];
  if (threadIdx.x == 0) {
    T.nr = G.nr;
    T.bump = G.bump;
    T.need = G.need;
    T.V = V;
    T.step = step;
    T.nch = nch;
  }
}

// giant g from the list written by k_grow_large (e, W); clears the walks'
// work counters wq[0..3]
__global__ void k_giant_init(SodCfg c, const float *__restrict__ pos,
                             const int64_t *__restrict__ ids,
                             const HaloPlan *__restrict__ plan,
                             const int32_t *__restrict__ large,
                             const int32_t *__restrict__ giants,
                             GiantState *__restrict__ S,
                             int32_t *__restrict__ wq) {
  const int32_t g = blockIdx.x;
  if (threadIdx.x == 0) {
    GrowVars V;
    HaloResult res;
    double xcs[3];
    grow_start(c, pos, ids, plan, large, giants[2 * g], V, res, xcs);
    V.W = giants[2 * g + 1];
    GiantState &T = S[g];
    T.V = V;
    T.res = res;
    for (int d = 0; d < 3; d++) T.xcs[d] = xcs[d];
    T.step = GS_BEGIN;
    T.nch = 0;
    T.need = 0;
    T.bump = 0;
    T.nr = 0;
  }
  if (g == 0 && threadIdx.x < 4) wq[threadIdx.x] = 0;
}

// begin for the giants in GS_BEGIN: close an unclosed one, else set up the
// count walk; clears the counter of its launch
__global__ void __launch_bounds__(GL_T)
k_giant_begin(SodCfg c, GiantState *__restrict__ S,
              int32_t *__restrict__ counters, int32_t *__restrict__ retry,
              unsigned long long *__restrict__ needmax, psod_halo_out out,
              int32_t *__restrict__ wq) {
  if (blockIdx.x == 0 && threadIdx.x == 0) wq[0] = 0;
  GiantState &T = S[blockIdx.x];
  const int32_t step0 = T.step;
  __syncthreads();
  if (step0 != GS_BEGIN) return;
  const GrowVars V = T.V;
  HaloResult res = T.res;
  if (!grow_reach(c, V, res)) {
    grow_close(c, counters, retry, needmax, 0, V, res, false, out);
    if (threadIdx.x == 0) T.step = GS_DONE;
    return;
  }
  if (threadIdx.x == 0) {
    const int64_t nseg = cube_segs(c, V.cr, V.W).nseg;
    T.ncube = 0;
    T.cseg = (int32_t)nseg;
    T.step = GS_COUNT;
    T.nch = (int32_t)chunks_of(nseg);
  }
}

// Layout of the per-segment counts of the giants in GS_COUNT, one after
// the other: their offsets to the states, their segment counts to
// lay[0..ngiant) (0 for the others) and the total to lay[ngiant].
__global__ void k_giant_layout(GiantState *__restrict__ S, int32_t ngiant,
                               int64_t *__restrict__ lay) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t off = 0;
  for (int32_t g = 0; g < ngiant; g++) {
    const int64_t n = S[g].step == GS_COUNT ? S[g].cseg : 0;
    S[g].segoff = off;
    lay[g] = n;
    off += n;
  }
  lay[ngiant] = off;
}

// Prefix of the chunks of the giants in phase `phase`: pre[g] = chunks of
// the giants before g that are in the phase, pre[ngiant] = all of them.
__global__ void k_giant_chunks(const GiantState *__restrict__ S,
                               int32_t ngiant, int32_t phase,
                               int32_t *__restrict__ pre) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int32_t p = 0;
  for (int32_t g = 0; g < ngiant; g++) {
    pre[g] = p;
    if (S[g].step == phase) p += S[g].nch;
  }
  pre[ngiant] = p;
}

// Exclusive prefix sums of the per-segment counts of each giant in
// GS_COUNT (one workgroup per giant, 4096 segments per round).
__global__ void __launch_bounds__(1024)
k_giant_segscan(const GiantState *__restrict__ S,
                const int32_t *__restrict__ segcnt,
                int32_t *__restrict__ segpre) {
  __shared__ int32_t wtot[16];
  const GiantState &T = S[blockIdx.x];
  if (T.step != GS_COUNT) return;  // uniform per workgroup
  const int32_t *in = segcnt + T.segoff;
  int32_t *out = segpre + T.segoff;
  const int32_t n = T.cseg;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int32_t carry = 0;
  for (int32_t base = 0; base < n; base += 4096) {
    const int32_t i0 = base + 4 * (int32_t)threadIdx.x;
    int32_t v[4], sum = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      v[u] = i0 + u < n ? in[i0 + u] : 0;
      sum += v[u];
    }
    const int32_t inc = wave_incl_scan(sum);
    if (l == 63) wtot[w] = inc;
    __syncthreads();
    int32_t run = carry + inc - sum, all = 0;
    for (int i = 0; i < 16; i++) {
      if (i < w) run += wtot[i];
      all += wtot[i];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (i0 + u < n) out[i0 + u] = run;
      run += v[u];
    }
    carry += all;
    __syncthreads();
  }
}

// chunks of CH_PART particles of a cube of ncube particles
__device__ inline int32_t part_chunks(int64_t ncube) {
  return (int32_t)max((int64_t)1, (ncube + CH_PART - 1) / CH_PART);
}

// First s in [0, n) with pre[s] >= x, or n, for a non-decreasing pre: wave
// 0 of the workgroup, 64 probes per round, the answer in all its lanes.
__device__ inline int32_t wave_lower_bound(const int32_t *pre, int32_t n,
                                           int64_t x) {
  const int l = threadIdx.x & 63;
  int32_t lo = 0, hi = n;  // the answer lies in [lo, hi]
  while (hi > lo) {
    const int32_t step = (hi - lo + 63) / 64;
    // lane l probes the last index of its part [lo + l step, lo + (l+1) step)
    const int64_t p = (int64_t)lo + (int64_t)(l + 1) * step - 1;
    const bool t = p >= hi || (int64_t)pre[p] >= x;
    const uint64_t m = __ballot(t);
    if (m == 0) {  // every probe below hi is short of x
      lo = hi;
      break;
    }
    const int f = __ffsll((unsigned long long)m) - 1;
    const int32_t nhi = (int32_t)min((int64_t)hi, (int64_t)lo + (int64_t)(f + 1) * step - 1);
    lo = lo + f * step;
    hi = nhi;
  }
  return lo;
}

// size for the giants in GS_COUNT: the histogram of the counted cube,
// cleared for walk 1; clears the counter of its launch
__global__ void __launch_bounds__(GL_T)
k_giant_size(SodCfg c, GiantState *__restrict__ S,
             int32_t *__restrict__ hist_all, int32_t *__restrict__ wq) {
  if (blockIdx.x == 0 && threadIdx.x == 0) wq[1] = 0;
  GiantState &T = S[blockIdx.x];
  const int32_t step0 = T.step;
  __syncthreads();
  if (step0 != GS_COUNT) return;
  GrowVars V = T.V;
  grow_size(V, (int64_t)T.ncube);
  int32_t *hist = hist_all + (int64_t)blockIdx.x * HIST_SLICE;
  for (int32_t b = threadIdx.x; b <= V.nb; b += GL_T) hist[b] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    T.V = V;
    T.step = GS_HIST;
    T.nch = part_chunks(V.ncube);
  }
}

// One walk of the giants in phase `phase`, spread over the launch: work
// items are (giant, chunk), taken from the counter *wq.  The count walk
// takes chunks of CH_SEG segments, records each segment's particles in
// segcnt and adds them to the giant's count.  Walks 1-3 take chunks of
// about CH_PART particles: chunk k holds the segments whose first particle
// (segpre, the prefix sums of segcnt) lies in [k CH_PART, (k+1) CH_PART),
// so that the dense rows of a giant's core are spread over many
// workgroups.  Walk 1 adds its LDS-counted inner bins to the giant's histogram with one
// atomic per bin; walks 2 and 3 take slots of the giant's collection slice
// with atomics on its fill counters; walk 3 adds its block sums to the
// giant's sums.
__global__ void __launch_bounds__(GL_T)
k_giant_walk(SodCfg c, int32_t phase, const int32_t *__restrict__ cell_start,
             const float4 *__restrict__ rec, const float *__restrict__ vel,
             GiantState *__restrict__ S, int32_t ngiant,
             int32_t *__restrict__ hist_all, uint64_t *__restrict__ pool_all,
             int64_t pool_cap, int32_t *__restrict__ segcnt,
             const int32_t *__restrict__ segpre,
             const int32_t *__restrict__ chpre, int32_t *__restrict__ wq) {
  __shared__ GrowShared G;
  __shared__ int32_t lh[HIST_LDS];
  __shared__ int32_t s_g, s_k, s_q0, s_q1;
  const int32_t total = chpre[ngiant];
  for (;;) {
    // item it: giant g with chpre[g] <= it < chpre[g + 1] (the giants out of
    // the phase hold no chunks), its chunk it - chpre[g]
    if (threadIdx.x < 64) {
      const int32_t it = __builtin_amdgcn_readfirstlane(
          threadIdx.x == 0 ? atomicAdd(wq, 1) : 0);
      int32_t g = -1, k = 0;
      if (it < total) {
        g = wave_lower_bound(chpre + 1, ngiant, (int64_t)it + 1);
        k = it - chpre[g];
      }
      if (threadIdx.x == 0) {
        s_g = g;
        s_k = k;
      }
    }
    __syncthreads();
    const int32_t g = s_g, k = s_k;
    __syncthreads();
    if (g < 0) break;
    GiantState &T = S[g];
    int64_t q0 = (int64_t)k * CH_SEG, q1 = q0 + CH_SEG;
    if (phase != GS_COUNT) {
      // the chunk's segments, from the giant's prefix sums
      if (threadIdx.x < 64) {
        const int32_t *pre = segpre + T.segoff;
        const int32_t nseg = T.cseg;
        const int32_t a = wave_lower_bound(pre, nseg, (int64_t)k * CH_PART);
        const int32_t b =
            k + 1 >= T.nch
                ? nseg
                : wave_lower_bound(pre, nseg, (int64_t)(k + 1) * CH_PART);
        if (threadIdx.x == 0) {
          s_q0 = a;
          s_q1 = b;
        }
      }
      __syncthreads();
      q0 = s_q0;
      q1 = s_q1;
    }
    const GrowVars V = T.V;
    int32_t *hist = hist_all + (int64_t)g * HIST_SLICE;
    uint64_t *pool = pool_all + (int64_t)g * pool_cap;
    if (phase == GS_COUNT) {
      const CellRef cr = V.cr;
      const int64_t n =
          cube_count(c, G, cell_start, cr, V.W, q0, q1, segcnt + T.segoff);
      if (threadIdx.x == 0 && n > 0)
        atomicAdd(&T.ncube, (unsigned long long)n);
    } else if (phase == GS_HIST) {
      const int32_t nl = min(V.nb, HIST_LDS);
      for (int32_t b = threadIdx.x; b < nl; b += GL_T) lh[b] = 0;
      __syncthreads();
      walk_hist(c, G, lh, hist, cell_start, rec, V, q0, q1);
      __syncthreads();
      for (int32_t b = threadIdx.x; b < nl; b += GL_T)
        if (lh[b]) atomicAdd(&hist[b], lh[b]);
    } else {
      for (int i = threadIdx.x; i < MAXRANGE; i += GL_T) {
        G.rs[i] = T.rs[i];
        G.re[i] = T.re[i];
        G.off[i] = T.off[i];
      }
      if (phase == GS_PROFILE)
        for (int j = threadIdx.x; j < c.nbins; j += GL_T) {
          G.prad[j] = T.prad[j];
          for (int w = 0; w < GL_T / 64; w++) G.pcnt[w][j] = 0;
        }
      __syncthreads();
      if (phase == GS_COLLECT) {
        walk_collect(c, G, cell_start, rec, V, T.nr, T.fill, pool, q0, q1);
      } else {
        const int flagv = (c.flags & PSOD_F_VELOCITY) && vel;
        ProfileSums P = {{0, 0, 0}, {0, 0, 0}, 0};
        walk_profile(c, G, cell_start, rec, vel, V, flagv, T.fill, pool, P,
                     q0, q1);
        P.nbd = block_sum_i64(P.nbd, G);
        for (int d = 0; d < 3; d++) {
          P.sp[d] = block_sum_f64(P.sp[d], G);
          P.sv[d] = block_sum_f64(P.sv[d], G);
        }
        if (threadIdx.x == 0) {
          if (P.nbd) atomicAdd(&T.nbd, (unsigned long long)P.nbd);
          for (int d = 0; d < 3; d++) {
            if (P.sp[d] != 0.0) atomicAdd(&T.sp[d], P.sp[d]);
            if (P.sv[d] != 0.0) atomicAdd(&T.sv[d], P.sv[d]);
          }
        }
        for (int j = threadIdx.x; j < c.nbins; j += GL_T) {
          int32_t s = 0;
          for (int w = 0; w < GL_T / 64; w++) s += G.pcnt[w][j];
          if (s) atomicAdd(&T.pcnt[j], s);
        }
      }
    }
    __syncthreads();
  }
}

// bins for the giants in GS_HIST; clears the counter of the walk 2 launch
__global__ void __launch_bounds__(GL_T)
k_giant_bins(SodCfg c, GiantState *__restrict__ S,
             int32_t *__restrict__ hist_all, int64_t pool_cap,
             int32_t *__restrict__ counters, int32_t *__restrict__ retry,
             unsigned long long *__restrict__ needmax, psod_halo_out out,
             int32_t *__restrict__ wq) {
  __shared__ GrowShared G;
  __shared__ int32_t lh[HIST_LDS];
  if (blockIdx.x == 0 && threadIdx.x == 0) wq[2] = 0;
  GiantState &T = S[blockIdx.x];
  const int32_t step0 = T.step;
  __syncthreads();
  if (step0 != GS_HIST) return;
  GrowVars V = T.V;
  giant_load(G, T);
  int32_t *hist = hist_all + (int64_t)blockIdx.x * HIST_SLICE;
  const int32_t step = grow_bins(c, G, lh, hist, V, pool_cap);
  __syncthreads();
  if (step == GS_FAILED) {
    HaloResult res = T.res;
    grow_close(c, counters, retry, needmax, G.need, V, res, true, out);
  }
  const int32_t nch = step == GS_COLLECT ? part_chunks(V.ncube) : 0;
  giant_save(G, T, V, step, nch);
}

// Band ranges of the staged giants with at least sort_min keys (2^16
// unless PSOD_GIANT_SORT_MIN says otherwise) are sorted by the device-wide
// radix sort before k_giant_cross, at most GIANT_SORT_LIST per cycle: one
// workgroup's bitonic network over such a range (2^20 keys, 36 passes
// through global memory) took most of a 24 ms cross.
constexpr int GIANT_SORT_LIST = 64;
static_assert(MAXRANGE <= 32, "presorted ranges are a 32-bit mask");

// The ranges to sort: list[0] = their count, then (first pool slot, keys)
// per range; marks them in each giant's presorted mask (one thread).
__global__ void k_giant_sortlist(GiantState *__restrict__ S, int32_t ngiant,
                                 int64_t pool_cap, int64_t sort_min,
                                 int64_t *__restrict__ list) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t k = 0;
  for (int32_t g = 0; g < ngiant; g++) {
    GiantState &T = S[g];
    uint32_t mask = 0;
    if (T.step == GS_COLLECT)
      for (int r = 0; r < T.nr; r++) {
        const int64_t n = T.k1[r] - T.k0[r];
        if (n >= sort_min && k < GIANT_SORT_LIST) {
          list[1 + 2 * k] = (int64_t)g * pool_cap + T.off[r];
          list[2 + 2 * k] = n;
          mask |= 1u << r;
          k++;
        }
      }
    T.presorted = mask;
  }
  list[0] = k;
}

// cross for the giants in GS_COLLECT; clears the sums of walk 3 and the
// counter of its launch
__global__ void __launch_bounds__(GL_T)
k_giant_cross(SodCfg c, GiantState *__restrict__ S,
              const int32_t *__restrict__ hist_all,
              uint64_t *__restrict__ pool_all, int64_t pool_cap,
              const float *__restrict__ pos, int32_t *__restrict__ counters,
              int32_t *__restrict__ retry,
              unsigned long long *__restrict__ needmax, psod_halo_out out,
              int32_t *__restrict__ wq) {
  __shared__ GrowShared G;
  if (blockIdx.x == 0 && threadIdx.x == 0) wq[3] = 0;
  GiantState &T = S[blockIdx.x];
  const int32_t step0 = T.step;
  __syncthreads();
  if (step0 != GS_COLLECT) return;
  GrowVars V = T.V;
  giant_load(G, T);
  const int32_t *hist = hist_all + (int64_t)blockIdx.x * HIST_SLICE;
  uint64_t *pool = pool_all + (int64_t)blockIdx.x * pool_cap;
  HaloResult res = T.res;
  double xcs[3] = {T.xcs[0], T.xcs[1], T.xcs[2]};
  const int32_t step =
      grow_cross(c, G, hist, pool, pos, V, res, xcs, pool_cap, T.presorted);
  __syncthreads();
  if (step == GS_FAILED)
    grow_close(c, counters, retry, needmax, G.need, V, res, true, out);
  if (threadIdx.x == 0) T.res = res;
  if (step == GS_PROFILE) {
    for (int j = threadIdx.x; j < 257; j += GL_T) T.pcnt[j] = 0;
    if (threadIdx.x == 0) {
      T.nbd = 0;
      for (int d = 0; d < 3; d++) T.sp[d] = T.sv[d] = 0.0;
    }
  }
  const int32_t nch = step == GS_PROFILE ? part_chunks(V.ncube) : 0;
  giant_save(G, T, V, step, nch);
}

// finish for the giants in GS_PROFILE
__global__ void __launch_bounds__(GL_T)
k_giant_finish(SodCfg c, GiantState *__restrict__ S,
               uint64_t *__restrict__ pool_all, int64_t pool_cap,
               const float *__restrict__ pos, int32_t *__restrict__ counters,
               int32_t *__restrict__ retry,
               unsigned long long *__restrict__ needmax, psod_halo_out out) {
  __shared__ GrowShared G;
  GiantState &T = S[blockIdx.x];
  const int32_t step0 = T.step;
  __syncthreads();
  if (step0 != GS_PROFILE) return;
  GrowVars V = T.V;
  giant_load(G, T);
  for (int j = threadIdx.x; j < c.nbins; j += GL_T) G.pcnt[0][j] = T.pcnt[j];
  ProfileSums P;
  for (int d = 0; d < 3; d++) {
    P.sp[d] = T.sp[d];
    P.sv[d] = T.sv[d];
  }
  P.nbd = (int64_t)T.nbd;
  __syncthreads();
  uint64_t *pool = pool_all + (int64_t)blockIdx.x * pool_cap;
  HaloResult res = T.res;
  double xcs[3] = {T.xcs[0], T.xcs[1], T.xcs[2]};
  grow_finish(c, G, pool, pos, V, res, xcs, P, out);
  grow_close(c, counters, retry, needmax, 0, V, res, false, out);
  if (threadIdx.x == 0) T.step = GS_DONE;
}

// giants still growing (not done, not failed) to *n
__global__ void k_giant_count(const GiantState *__restrict__ S,
                              int32_t ngiant, int32_t *__restrict__ n) {
  if (blockIdx.x != 0 || threadIdx.x >= 64) return;
  int32_t k = 0;
  for (int32_t g = threadIdx.x; g < ngiant; g += 64)
    k += S[g].step != GS_DONE && S[g].step != GS_FAILED;
  k = wave_sum(k);
  if (threadIdx.x == 0) *n = k;
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
// item counters of a re-run launch (layout of counters[0..7]) and its
// largest needed slice
__global__ void k_retry_reset(int32_t *__restrict__ c, int32_t nitem,
                              unsigned long long *__restrict__ need) {
  if (blockIdx.x == 0 && threadIdx.x < 8) c[threadIdx.x] = threadIdx.x == 1 ? nitem : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *need = 0ull;
}

// Sort keys of the large halos, most costly first.  For the centre search
// the cost is the particles near the seed (members plus those of the 5^3
// cells).  For the growth it is the particles of the cube of the planned
// reach, summed over the 8^3 blocks that the cube overlaps: a small halo
// beside a giant one walks the giant's particles too.
__global__ void k_large_keys(SodCfg c, const HaloPlan *__restrict__ plan,
                             const int32_t *__restrict__ large, int32_t n,
                             const int32_t *__restrict__ dn,
                             const int32_t *__restrict__ cell_start,
                             int growth,
                             uint64_t *__restrict__ keys,
                             uint32_t *__restrict__ vals) {
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || (dn && e >= *dn)) return;
  const int32_t h = large[e];
  const HaloPlan pl = plan[h];
  int64_t cost = (int64_t)max(pl.nmem, 0) + pl.n5;
  if (growth) {
    const int32_t nb = c.nx >> 3, W = max(pl.w0, 2);
    int32_t lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = floor8(pl.cs[d] - W);
      hi[d] = min(floor8(pl.cs[d] + W), lo[d] + nb - 1);  // at most nx cells
    }
    // every s-th block per axis (all of them up to 16 per axis), scaled by
    // s^3: the key only orders the halos
    const int32_t s = (hi[0] - lo[0] + 16) / 16;
    cost = 0;
    for (int32_t bz = lo[2]; bz <= hi[2]; bz += s)
      for (int32_t by = lo[1]; by <= hi[1]; by += s)
        for (int32_t bx = lo[0]; bx <= hi[0]; bx += s) {
          const int64_t b = ((int64_t)wrapi(bz, nb) * nb + wrapi(by, nb)) * nb +
                            wrapi(bx, nb);
          cost += cell_start[(b + 1) * BS3] - cell_start[b * BS3];
        }
    cost *= (int64_t)s * s * s;
    // A cube already holding more particles than a sphere of radius
    // W / rmax_frac at the first threshold would is likely to be outgrown
    // and walked again at twice the reach (about 4x the particles): such a
    // halo is costed at both walks, so that it starts among the first.
    const double rw = (double)W / c.rmax_frac;
    if ((double)cost * c.mpart > c.delta[0] * (4.0 / 3.0 * M_PI) * rw * rw * rw)
      cost *= 5;
  }
  keys[e] = (uint64_t)(0x7FFFFFFFFFFFll - min(cost, (int64_t)0x7FFFFFFFFFFFll));
  vals[e] = (uint32_t)h;
}

// Clamps the task count of a captured pass to the buffers' capacity and
// flags the overflow (the pass then computes nothing for the large halos
// and the host runs it again without the graph).
__global__ void k_task_cap(int32_t *__restrict__ ntot, int32_t tcap,
                           int32_t *__restrict__ overflow) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && *ntot > tcap) {
    *overflow = 1;
    *ntot = 0;
  }
}

__global__ void k_copy_i32(const int32_t *__restrict__ a,
                           int32_t *__restrict__ b) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *b = *a;
}

// records in reach above which a pilot task is split; PSOD_PILOT_BIG
// overrides the default for tests (at least 64).  Pilots plus split pieces
// at c3 / c4 (ms): 2^13 4.1 / 2.2, 2^14 3.0 / 2.3, 2^15 3.8 / 2.6, 2^16
// 5.0 / 2.8.
static int32_t pilot_big() {
  int32_t v = 1 << 14;
  if (const char *e = getenv("PSOD_PILOT_BIG")) {
    const long long x = atoll(e);
    if (x >= 64 && x <= INT32_MAX) v = (int32_t)x;
  }
  return v;
}

// Work counts of a pass.  Eager calls read them back on the host and size
// every launch exactly.  A pass that psod_pass_graph captures cannot read
// anything back: its launches are sized by bounds (the halo count, and a
// centre-search task capacity from an earlier eager pass) and its kernels
// read the counts on the device.  counters[36] is set when that pass's
// tasks exceed the capacity (the host then runs the pass without the
// graph), counters[37] holds the large-halo count before the small path's
// deferrals.
struct SodCounts {
  bool dev;         // counts on the device only (captured pass)
  int32_t nlarge1;  // large halos of the plan (eager)
  int32_t nlarge;   // large halos after the small path (eager)
};

// Centre search of the first nlarge halos of `large` (positions e; with
// dn, the first *dn of at most nlarge): tasks of 64 candidates, their
// bounds, pilots, groups, and each halo's centre in plan[].centre.  ntask,
// task0 and ntot are scratch of nlarge + 16 entries (ntot: 16); hbest
// holds 2 * nlarge + 4 entries.  after_setup() is called once the task
// count is known, before the search's kernels are queued.
template <typename F>
static int centre_search(psod_ctx *ctx, const SodCfg &c,
                         const psod_catalogue *cat, HaloPlan *plan,
                         const int32_t *large, int32_t nlarge,
                         const int32_t *dn, int32_t *overflow, int32_t *ntask,
                         int32_t *task0, int32_t *ntot,
                         unsigned long long *hbest, F &&after_setup) {
  hipStream_t st = ctx->stream;
  k_task_count<<<grid_for(nlarge, 256), 256, 0, st>>>(c, cat->cell_start,
                                                       plan, large, nlarge, dn,
                                                       ntask);
  PSOD_LAUNCH_CHECK();
  PSOD_TRY(psod_scan_i32(ctx, ntask, task0, nlarge, ntot));
  // tcap: tasks the buffers hold, and the stride of their layout
  int32_t nt = 0, tcap = 0;
  if (dn) {
    tcap = (int32_t)std::min<int64_t>(ctx->gcap.tasks, 1 << 27);
    k_task_cap<<<1, 64, 0, st>>>(ntot, tcap, overflow);
    PSOD_LAUNCH_CHECK();
  } else {
    PSOD_HIP(hipMemcpyAsync(&nt, ntot, 4, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    if (nt > (1 << 27))
      return psod_set_error(PSOD_E_CAPACITY,
                            "%d centre-search tasks exceed 2^27", nt);
    tcap = nt;
    ctx->sod_stat[7] = nt;
  }
  PSOD_TRY(after_setup());
  if (tcap == 0) return PSOD_OK;
  // tasks, their results, bounds, bounds per (cell, group) plus the
  // per-group totals (28 x 8 per task), the results of the groups (8 per
  // task), the work list of groups and its length, then the split pilot
  // tasks and their per-candidate totals
  int4 *tasks = (int4 *)psod_scratch(
      ctx, S_TASKRES,
      (size_t)tcap * (sizeof(int4) + 24 + 28 * 8 * 8 + 8 * 16 + 8 * 4) + 256 +
          (size_t)PILOT_BIGCAP * (4 + 64 * 8));
  if (!tasks) return PSOD_E_NOMEM;
  uint64_t *tres = (uint64_t *)(tasks + tcap);
  unsigned long long *ubt = (unsigned long long *)(tres + 2 * (int64_t)tcap);
  unsigned long long *cub = ubt + tcap;
  uint64_t *ires = (uint64_t *)(cub + (int64_t)tcap * 28 * 8);
  int32_t *items = (int32_t *)(ires + 16 * (int64_t)tcap);
  int32_t *nitems = items + 8 * (int64_t)tcap;
  // work counters of the bound, pilot, group and split-pilot launches, and
  // the count of split pilot tasks
  int32_t *wq = nitems + 4;
  int32_t *nbig = wq + 4;
  // split pilot tasks and their per-candidate totals
  int32_t *big = nitems + 64;
  unsigned long long *pacc = (unsigned long long *)(big + PILOT_BIGCAP);
  k_task_emit<<<grid_for(nlarge, 256), 256, 0, st>>>(
      c, cat->cell_start, plan, large, nlarge, dn, ntask, task0,
      dn ? overflow : nullptr, tasks);
  PSOD_LAUNCH_CHECK();
  k_res_init<<<grid_for(tcap, 256), 256, 0, st>>>(tres, tcap);
  PSOD_LAUNCH_CHECK();
  k_res_init<<<grid_for(8 * (int64_t)tcap, 256), 256, 0, st>>>(
      ires, 8 * (int64_t)tcap);
  PSOD_LAUNCH_CHECK();
  // hbest[nlarge..nlarge+3]: units evaluated, units skipped, pair tests,
  // units abandoned; then each halo's largest task bound
  unsigned long long *cstat = hbest + nlarge;
  unsigned long long *hpil = cstat + 4;
  PSOD_TRY(psod_fill_u64(ctx, (uint64_t *)hbest, 2 * (int64_t)nlarge + 4, 0));
  PSOD_TRY(psod_fill_i32(ctx, nitems, 12, 0));
  if (!ctx->capturing) PSOD_HIP(hipEventRecord(ctx->ev[0], st));
  // as many workgroups as fit at once; their waves take work from counters
  auto resident = [&](const void *k) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 256, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 4;
    return (unsigned)ctx->num_cus * (unsigned)nb;
  };
  const unsigned nwg = resident((const void *)k_centre_bound);
  k_centre_bound<<<nwg, 256, 0, st>>>(c, cat->cell_start,
                                      (const float4 *)cat->rec, plan, large,
                                      tasks, ntot, tcap, hpil, ubt, cub, wq);
  PSOD_LAUNCH_CHECK();
  k_centre_pilot<<<resident((const void *)k_centre_pilot), 256, 0, st>>>(
      c, cat->cell_start, (const float4 *)cat->rec, plan, large, tasks, ntot,
      tres, hbest, hpil, ubt, cstat, pilot_big(), big, nbig, pacc, wq + 1);
  PSOD_LAUNCH_CHECK();
  k_centre_pilot_big<<<resident((const void *)k_centre_pilot_big), 256, 0,
                       st>>>(c, cat->cell_start, (const float4 *)cat->rec,
                             plan, large, tasks, big, nbig, pacc, cstat,
                             wq + 3);
  PSOD_LAUNCH_CHECK();
  k_centre_pilot_fin<<<PILOT_BIGCAP / 4, 256, 0, st>>>(
      (const float4 *)cat->rec, tasks, big, nbig, pacc, tres, hbest, cstat);
  PSOD_LAUNCH_CHECK();
  k_task_items<<<grid_for(tcap, 256), 256, 0, st>>>(
      tasks, ntot, tcap, ubt, hpil, cub, hbest, items, nitems, cstat);
  PSOD_LAUNCH_CHECK();
  // items sorted by their bound over the halo's best, largest first
  int32_t ni = 8 * tcap;  // capacity (device count) or the exact count
  if (!dn) {
    PSOD_HIP(hipMemcpyAsync(&ni, nitems, 4, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
  }
  if (ni > 1) {
    uint64_t *keys = (uint64_t *)psod_scratch(ctx, S_KEYS, (size_t)ni * 8);
    uint32_t *vals = (uint32_t *)psod_scratch(ctx, S_VALS, (size_t)ni * 4);
    if (!keys || !vals) return PSOD_E_NOMEM;
    k_item_keys<<<grid_for(ni, 256), 256, 0, st>>>(items, nitems, tasks, ntot,
                                                   tcap, cub, hbest, keys,
                                                   vals);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(psod_radix_sort_u64(ctx, ni, keys, vals, 32, dn ? nitems : nullptr));
    PSOD_HIP(hipMemcpyAsync(items, vals, (size_t)ni * 4,
                            hipMemcpyDeviceToDevice, st));
  }
  k_centre_group<<<resident((const void *)k_centre_group), 256, 0, st>>>(
      c, cat->cell_start, (const float4 *)cat->rec, plan, large, tasks, ntot,
      tcap, items, nitems, hbest, cub, ires, cstat, wq + 2);
  PSOD_LAUNCH_CHECK();
  k_centre_reduce<<<grid_for(nlarge, 4), 256, 0, st>>>(plan, large, nlarge, dn,
                                                       tres, ires);
  PSOD_LAUNCH_CHECK();
  if (!ctx->capturing) {
    PSOD_HIP(hipEventRecord(ctx->ev[1], st));
    // read back with the growth's counters (the stream is synchronised
    // before the pass returns)
    PSOD_HIP(hipMemcpyAsync(ctx->sod_stat + 4, cstat, 3 * 8,
                            hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipMemcpyAsync(ctx->sod_stat + 9, cstat + 3, 8,
                            hipMemcpyDeviceToHost, st));
  }
  return PSOD_OK;
}

// The scratch of one pass: plan, lists, counters and the centre search's
// per-halo arrays.
struct SodScratch {
  HaloPlan *plan;
  int32_t *small, *large, *counters;
  int32_t *ntask, *task0, *ntot, *retry;
  unsigned long long *hbest;
  int32_t *giants;     // (e, W) of the giant halos left to the staged growth
  int64_t *glay;       // their per-segment count layout (k_giant_layout)
  int32_t *gchunk;     // prefix of their chunks in a walk (k_giant_chunks)
  int64_t *gsort;      // their band ranges for the device-wide sort
  GiantState *gstate;  // their states, GIANT_CAP entries
};

static int sod_scratch(psod_ctx *ctx, int64_t nhalo, SodScratch *S) {
  S->plan = (HaloPlan *)psod_scratch(ctx, S_PLAN,
                                     (size_t)nhalo * sizeof(HaloPlan));
  int32_t *lists = (int32_t *)psod_scratch(ctx, S_TASK,
                                           (size_t)(2 * nhalo + 64) * 4);
  // sized for every halo: the growth's retry list holds up to all of them
  const int32_t ncap = (int32_t)nhalo;
  int32_t *ntask = (int32_t *)psod_scratch(
      ctx, S_MISC, (size_t)(5 * ncap + 128) * 4 + (size_t)ncap * 16 + 512);
  // giant list (2 * GIANT_CAP), layout (GIANT_CAP + 1), chunk prefix
  // (GIANT_CAP + 1), states (GIANT_CAP), each 256-byte aligned
  constexpr size_t G_LIST = 0, G_LAY = 2 * GIANT_CAP * 4,
                   G_CH = G_LAY + ((GIANT_CAP + 1) * 8 + 255) / 256 * 256,
                   G_SORT = G_CH + ((GIANT_CAP + 1) * 4 + 255) / 256 * 256,
                   G_ST = G_SORT +
                          ((2 * GIANT_SORT_LIST + 1) * 8 + 255) / 256 * 256;
  char *gb = (char *)psod_scratch(
      ctx, S_GIANT, G_ST + (size_t)GIANT_CAP * sizeof(GiantState));
  if (!S->plan || !lists || !ntask || !gb) return PSOD_E_NOMEM;
  S->giants = (int32_t *)(gb + G_LIST);
  S->glay = (int64_t *)(gb + G_LAY);
  S->gchunk = (int32_t *)(gb + G_CH);
  S->gsort = (int64_t *)(gb + G_SORT);
  S->gstate = (GiantState *)(gb + G_ST);
  S->small = lists;
  S->large = lists + nhalo;
  S->counters = lists + 2 * nhalo;
  S->ntask = ntask;
  // best S found so far per large halo (branch and bound of the search)
  S->hbest = (unsigned long long *)(
      (((uintptr_t)(ntask + 5 * ncap + 128)) + 255) & ~(uintptr_t)255);
  S->task0 = ntask + ncap + 16;
  S->ntot = S->task0 + ncap + 16;
  S->retry = S->ntot + 16;  // positions of halos that need larger slices
  return PSOD_OK;
}

// Launches of the growth of large halos that failed the first launch (their
// bands exceeded the slices): fewer workgroups with larger slices.  A halo
// re-run from its first reach may need more than its first failure showed
// (a larger reach later), so the slice grows at least 4x per launch.
static int grow_retries(psod_ctx *ctx, const SodCfg &c,
                        const psod_catalogue *cat, const SodScratch &S,
                        int32_t nlarge, int32_t ngrow, int32_t *hist,
                        int64_t pool_cap, int32_t nfail, uint64_t hneed,
                        int32_t *unclosed, const psod_halo_out *out) {
  hipStream_t st = ctx->stream;
  int32_t *cur = S.retry, *nxt = S.retry + nlarge + 16;
  int64_t cap2 = pool_cap;
  for (int pass = 0; nfail > 0; pass++) {
    const int64_t budget = (int64_t)32 << 30;  // bytes for all slices
    cap2 = std::max<int64_t>((int64_t)hneed, 4 * cap2);
    if (pass >= 6 || cap2 * 8 > budget)
      return psod_set_error(PSOD_E_CAPACITY,
                            "sod: %d halos need more than %lld collected "
                            "particles in their radial bands", nfail,
                            (long long)std::min<int64_t>(cap2, budget / 8));
    if (getenv("PSOD_VERBOSE")) {
      fprintf(stderr, "psod: %d of %d large halos re-run with %lld-key "
              "slices; halos:", nfail, nlarge, (long long)cap2);
      const int32_t nshow = std::min<int32_t>(nfail, 64);
      std::vector<int32_t> pe(nshow), ph(nlarge);
      PSOD_HIP(hipMemcpy(pe.data(), cur, nshow * 4, hipMemcpyDeviceToHost));
      PSOD_HIP(hipMemcpy(ph.data(), S.large, (size_t)nlarge * 4,
                         hipMemcpyDeviceToHost));
      for (int32_t i = 0; i < nshow; i++) fprintf(stderr, " %d", ph[pe[i]]);
      fprintf(stderr, "\n");
    }
    const int64_t nfit = budget / (cap2 * 8);
    const unsigned ngrow2 = (unsigned)std::min<int64_t>(
        std::min<int64_t>(nfail, nfit), (int64_t)ngrow);  // histogram slices
    uint64_t *pool2 = (uint64_t *)psod_scratch(ctx, S_COLLECT,
                                               (size_t)cap2 * 8 * ngrow2);
    if (!pool2) return PSOD_E_NOMEM;
    if (pass == 0) ctx->sod_stat[2] = nfail;
    ctx->sod_stat[3] = cap2;
    int32_t *c2 = S.counters + 24;
    unsigned long long *need2 = (unsigned long long *)(S.counters + 32);
    k_retry_reset<<<1, 64, 0, st>>>(c2, nfail, need2);
    PSOD_LAUNCH_CHECK();
    k_grow_large<<<ngrow2, GL_T, 0, st>>>(
        c, cat->cell_start, (const float4 *)cat->rec, cat->pos, cat->vel,
        cat->ids, S.plan, S.large, c2, hist, pool2, cap2, cur, nxt, need2,
        nullptr, nullptr, 0, *out);
    PSOD_LAUNCH_CHECK();
    int32_t hf2[2];
    PSOD_HIP(hipMemcpyAsync(hf2, c2 + 4, 8, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipMemcpyAsync(&hneed, need2, 8, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    nfail = hf2[0];
    *unclosed += hf2[1];
    std::swap(cur, nxt);
  }
  return PSOD_OK;
}

// keys per workgroup slice of the growth's collection pool; PSOD_SLICE_KEYS
// lowers it (tests of the re-run launch on small catalogues)
static int64_t growth_pool_cap() {
  int64_t pool_cap = (int64_t)1 << 20;
  if (const char *e = getenv("PSOD_SLICE_KEYS")) {
    const long long v = atoll(e);
    if (v >= 16 && v < pool_cap) pool_cap = v;
  }
  return pool_cap;
}

// The batched pass up to the first growth launch.  With dev set (a pass
// that is being captured, or the eager run that sizes its buffers) nothing
// is read back and the growth's failures are left on the device for
// sod_finish_growth.
static int sod_fof_queue(psod_ctx *ctx, const psod_sod_params *p,
                         const psod_catalogue *cat, int64_t nhalo,
                         const int32_t *d_seed, const int32_t *d_nmem,
                         const psod_halo_out *out, bool dev, SodScratch *S,
                         int32_t *nlarge_out) {
  PSOD_TRY(check_sod_args(p, cat, nhalo, d_seed, out));
  if (!out->profile)
    return psod_set_error(PSOD_E_ARG, "sod: profile output is required");
  if (nhalo >= (1ll << 31) - 64)
    return psod_set_error(PSOD_E_ARG, "sod: %lld halos", (long long)nhalo);
  SodCfg c = make_cfg(p, cat->npart);
  const psod_tunables tun = psod_tunables_for(ctx, p->nx);
  // the one-wave path holds at most CAP5 particles in LDS
  c.small_max = p->small_max > 0 ? std::min(p->small_max, CAP5)
                                 : std::min(tun.small_max, CAP5);
  c.bound_min = tun.bound_min;
  c.giant_min = tun.giant_min;
  hipStream_t st = ctx->stream;
  PSOD_TRY(sod_scratch(ctx, nhalo, S));
  HaloPlan *plan = S->plan;
  int32_t *small = S->small, *large = S->large, *counters = S->counters;
  PSOD_TRY(psod_fill_i32(ctx, counters, 40, 0));
  k_plan<<<grid_for(nhalo, 4), 256, 0, st>>>(
      c, nhalo, d_seed, d_nmem, cat->pos, cat->cell_start, plan, small, large,
      counters, *out);
  PSOD_LAUNCH_CHECK();
  k_plan_lists<<<grid_for(nhalo, 256), 256, 0, st>>>(nhalo, plan, small, large,
                                                     counters);
  PSOD_LAUNCH_CHECK();
  // halos of the large path known from the plan; the small path appends the
  // ones it defers (their centre found) behind them
  int32_t nlarge1 = (int32_t)nhalo;  // a bound in the device-count mode
  const int32_t *dn1 = nullptr;
  if (dev) {
    k_copy_i32<<<1, 64, 0, st>>>(counters + 1, counters + 37);
    PSOD_LAUNCH_CHECK();
    dn1 = counters + 37;
  } else {
    int32_t hc[4];
    PSOD_HIP(hipMemcpyAsync(hc, counters, 16, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    nlarge1 = hc[1];
  }
  // Small halos, persistent (a few workgroups per CU), on the side stream:
  // they overlap the centre search of the large halos, which touches other
  // halos' plans and outputs.  They are queued once the search's setup has
  // run, so that its short kernels do not wait for the small halos' CUs.
  auto launch_small = [&]() -> int {
    PSOD_HIP(hipEventRecord(ctx->fork, st));
    PSOD_HIP(hipStreamWaitEvent(ctx->side, ctx->fork, 0));
    k_sod_small<<<(unsigned)ctx->num_cus * tun.small_wg_per_cu, 64, 0,
                  ctx->side>>>(
        c, cat->cell_start, (const float4 *)cat->rec, cat->pos, cat->vel,
        cat->ids, plan, small, large, counters, *out);
    PSOD_LAUNCH_CHECK();
    PSOD_HIP(hipEventRecord(ctx->join, ctx->side));
    return PSOD_OK;
  };
  if (!dev)
    for (int i = 0; i < 13; i++) ctx->sod_stat[i] = 0;
  // largest halos first (centre-search tasks and growth), so that the
  // persistent workgroups do not finish on a giant halo picked up last
  auto size_order = [&](int32_t n, const int32_t *dn, int growth) -> int {
    uint64_t *keys = (uint64_t *)psod_scratch(ctx, S_KEYS, (size_t)n * 8);
    uint32_t *vals = (uint32_t *)psod_scratch(ctx, S_VALS, (size_t)n * 4);
    if (!keys || !vals) return PSOD_E_NOMEM;
    k_large_keys<<<grid_for(n, 256), 256, 0, st>>>(
        c, plan, large, n, dn, cat->cell_start, growth, keys, vals);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(psod_radix_sort_u64(ctx, n, keys, vals, 48, dn));
    PSOD_HIP(hipMemcpyAsync(large, vals, (size_t)n * 4,
                            hipMemcpyDeviceToDevice, st));
    return PSOD_OK;
  };
  if (nlarge1 > 0) {
    PSOD_TRY(size_order(nlarge1, dn1, 0));
    PSOD_TRY(centre_search(ctx, c, cat, plan, large, nlarge1, dn1,
                           counters + 36, S->ntask, S->task0, S->ntot,
                           S->hbest, launch_small));
  } else {
    PSOD_TRY(launch_small());
  }
  PSOD_HIP(hipStreamWaitEvent(st, ctx->join, 0));
  int32_t nlarge = (int32_t)nhalo;  // a bound in the device-count mode
  const int32_t *dn = dev ? counters + 1 : nullptr;
  if (!dev) {
    int32_t hc[4];
    PSOD_HIP(hipMemcpyAsync(hc, counters, 16, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    nlarge = hc[1];
    ctx->sod_stat[0] = nhalo - nlarge;
    ctx->sod_stat[1] = nlarge;
  }
  *nlarge_out = nlarge;
  if (nlarge == 0) return PSOD_OK;
  PSOD_TRY(size_order(nlarge, dn, 1));
  // one 512-thread workgroup per CU (the kernel's registers allow 8 waves)
  const unsigned ngrow = (unsigned)std::min<int32_t>(nlarge, ctx->num_cus);
  int32_t *hist = (int32_t *)psod_scratch(ctx, S_BINS,
                                          (size_t)ngrow * HIST_SLICE * 4);
  const int64_t pool_cap = growth_pool_cap();
  uint64_t *pool = (uint64_t *)psod_scratch(ctx, S_COLLECT,
                                            (size_t)pool_cap * 8 * ngrow);
  if (!hist || !pool) return PSOD_E_NOMEM;
  // counters[16..17]: largest slice a failed halo needs; counters[24..31]
  // and [32..33]: the same for a re-run launch; counters[6]: giant halos
  // left to the staged growth
  unsigned long long *needmax = (unsigned long long *)(counters + 16);
  k_grow_large<<<ngrow, GL_T, 0, st>>>(
      c, cat->cell_start, (const float4 *)cat->rec, cat->pos, cat->vel,
      cat->ids, plan, large, counters, hist, pool, pool_cap, nullptr,
      S->retry, needmax, dev ? counters + 36 : nullptr, S->giants, GIANT_CAP,
      *out);
  PSOD_LAUNCH_CHECK();
  return PSOD_OK;
}

// Staged growth of the giant halos that the first growth launch listed:
// cycles of the k_giant_* launches until every giant is done or failed
// (failures join the re-run list of the first launch).  A cycle ends with
// one read-back of the giants still growing; a giant takes a second cycle
// only when its crossing lies beyond its cube.
static int grow_giants(psod_ctx *ctx, const SodCfg &c,
                       const psod_catalogue *cat, const SodScratch &S,
                       int32_t ngiant, const psod_halo_out *out) {
  hipStream_t st = ctx->stream;
  const int64_t pool_cap = growth_pool_cap();
  int32_t *hist = (int32_t *)psod_scratch(ctx, S_GIANT_H,
                                          (size_t)ngiant * HIST_SLICE * 4);
  uint64_t *pool = (uint64_t *)psod_scratch(ctx, S_GIANT_P,
                                            (size_t)ngiant * pool_cap * 8);
  if (!hist || !pool) return PSOD_E_NOMEM;
  GiantState *gs = S.gstate;
  int32_t *wq = S.counters + 8;      // counters[8..11]: the walks' queues
  int32_t *nleft = S.counters + 12;  // giants still growing
  unsigned long long *needmax = (unsigned long long *)(S.counters + 16);
  const float4 *rec = (const float4 *)cat->rec;
  // the walk kernel holds one 512-thread workgroup per CU (its LDS)
  const unsigned nwalk = (unsigned)ctx->num_cus;
  k_giant_init<<<ngiant, 64, 0, st>>>(c, cat->pos, cat->ids, S.plan, S.large,
                                      S.giants, gs, wq);
  PSOD_LAUNCH_CHECK();
  int32_t *segcnt = nullptr, *segpre = nullptr;
  auto walk = [&](int32_t phase, int32_t *q) -> int {
    k_giant_chunks<<<1, 64, 0, st>>>(gs, ngiant, phase, S.gchunk);
    PSOD_LAUNCH_CHECK();
    k_giant_walk<<<nwalk, GL_T, 0, st>>>(c, phase, cat->cell_start, rec,
                                         cat->vel, gs, ngiant, hist, pool,
                                         pool_cap, segcnt, segpre, S.gchunk,
                                         q);
    PSOD_LAUNCH_CHECK();
    return PSOD_OK;
  };
  int64_t sort_min = 1 << 16;
  if (const char *e = getenv("PSOD_GIANT_SORT_MIN")) {
    const long long v = atoll(e);
    if (v >= 2) sort_min = v;
  }
  constexpr int MAX_CYCLES = 16;  // the reach at least doubles per cycle
  int cyc = 0;
  for (;; cyc++) {
    if (cyc == MAX_CYCLES)
      return psod_set_error(PSOD_E_UNCLOSED,
                            "sod: giant halos still growing after %d cycles",
                            MAX_CYCLES);
    k_giant_begin<<<ngiant, GL_T, 0, st>>>(c, gs, S.counters, S.retry,
                                           needmax, *out, wq);
    PSOD_LAUNCH_CHECK();
    // per-segment counts of the cubes and their prefix sums, per giant
    k_giant_layout<<<1, 64, 0, st>>>(gs, ngiant, S.glay);
    PSOD_LAUNCH_CHECK();
    int64_t lay[GIANT_CAP + 1];
    PSOD_HIP(hipMemcpyAsync(lay, S.glay, (size_t)(ngiant + 1) * 8,
                            hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    segcnt = (int32_t *)psod_scratch(ctx, S_GIANT_SEG,
                                     (size_t)(2 * lay[ngiant] + 64) * 4);
    if (!segcnt) return PSOD_E_NOMEM;
    segpre = segcnt + lay[ngiant] + 32;
    PSOD_TRY(walk(GS_COUNT, wq + 0));
    k_giant_segscan<<<ngiant, 1024, 0, st>>>(gs, segcnt, segpre);
    PSOD_LAUNCH_CHECK();
    k_giant_size<<<ngiant, GL_T, 0, st>>>(c, gs, hist, wq);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(walk(GS_HIST, wq + 1));
    k_giant_bins<<<ngiant, GL_T, 0, st>>>(c, gs, hist, pool_cap, S.counters,
                                          S.retry, needmax, *out, wq);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(walk(GS_COLLECT, wq + 2));
    // the large band ranges sorted across the GPU
    k_giant_sortlist<<<1, 64, 0, st>>>(gs, ngiant, pool_cap, sort_min,
                                       S.gsort);
    PSOD_LAUNCH_CHECK();
    int64_t sl[2 * GIANT_SORT_LIST + 1];
    PSOD_HIP(hipMemcpyAsync(sl, S.gsort, sizeof(sl), hipMemcpyDeviceToHost,
                            st));
    PSOD_HIP(hipStreamSynchronize(st));
    for (int64_t k = 0; k < sl[0]; k++) {
      const int64_t n = sl[2 + 2 * k];
      uint32_t *vals =
          (uint32_t *)psod_scratch(ctx, S_GIANT_V, (size_t)n * 4);
      if (!vals) return PSOD_E_NOMEM;
      PSOD_TRY(psod_radix_sort_u64(ctx, n, pool + sl[1 + 2 * k], vals, 64));
    }
    k_giant_cross<<<ngiant, GL_T, 0, st>>>(c, gs, hist, pool, pool_cap,
                                           cat->pos, S.counters, S.retry,
                                           needmax, *out, wq);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(walk(GS_PROFILE, wq + 3));
    k_giant_finish<<<ngiant, GL_T, 0, st>>>(c, gs, pool, pool_cap, cat->pos,
                                            S.counters, S.retry, needmax,
                                            *out);
    PSOD_LAUNCH_CHECK();
    k_giant_count<<<1, 64, 0, st>>>(gs, ngiant, nleft);
    PSOD_LAUNCH_CHECK();
    int32_t left = 0;
    PSOD_HIP(hipMemcpyAsync(&left, nleft, 4, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
    if (left == 0) break;
  }
  ctx->sod_stat[10] = ngiant;
  ctx->sod_stat[11] = cyc + 1;
  return PSOD_OK;
}

// After the first growth launch: runs the staged growth of the giant
// halos, reads the failures back, runs the re-run launches they need and
// reports unclosed halos.  nlarge is the large-halo
// count (read back here when the pass ran with device counts).
static int sod_finish_growth(psod_ctx *ctx, const psod_sod_params *p,
                             const psod_catalogue *cat, int64_t nhalo,
                             const SodScratch &S, int32_t nlarge, bool dev,
                             const psod_halo_out *out) {
  hipStream_t st = ctx->stream;
  int32_t hc[8];
  unsigned long long hneed = 0;
  PSOD_HIP(hipMemcpyAsync(hc, S.counters, 32, hipMemcpyDeviceToHost, st));
  PSOD_HIP(hipMemcpyAsync(&hneed, S.counters + 16, 8, hipMemcpyDeviceToHost,
                          st));
  PSOD_HIP(hipStreamSynchronize(st));
  if (dev) {
    nlarge = hc[1];
    ctx->sod_stat[0] = nhalo - nlarge;
    ctx->sod_stat[1] = nlarge;
  }
  ctx->sod_stat[12] = hc[7];  // halos of the first growth launch sorted in LDS
  if (nlarge == 0) return PSOD_OK;
  if (ctx->sod_stat[7] > 0 && !dev) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess)
      ctx->sod_stat[8] = (int64_t)((double)ms * 1e6);
  }
  SodCfg c = make_cfg(p, cat->npart);
  const psod_tunables tun = psod_tunables_for(ctx, p->nx);
  c.small_max = p->small_max > 0 ? std::min(p->small_max, CAP5)
                                 : std::min(tun.small_max, CAP5);
  c.bound_min = tun.bound_min;
  c.giant_min = tun.giant_min;
  const int32_t ngrow = std::min<int32_t>(nlarge, ctx->num_cus);
  int32_t *hist = (int32_t *)psod_scratch(ctx, S_BINS,
                                          (size_t)ngrow * HIST_SLICE * 4);
  if (!hist) return PSOD_E_NOMEM;
  const int32_t ngiant = std::min(hc[6], GIANT_CAP);
  if (ngiant > 0) {
    PSOD_TRY(grow_giants(ctx, c, cat, S, ngiant, out));
    // the staged giants' failures and unclosed halos join the counts
    PSOD_HIP(hipMemcpyAsync(hc, S.counters, 32, hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipMemcpyAsync(&hneed, S.counters + 16, 8,
                            hipMemcpyDeviceToHost, st));
    PSOD_HIP(hipStreamSynchronize(st));
  }
  int32_t unclosed = hc[5];
  PSOD_TRY(grow_retries(ctx, c, cat, S, nlarge, ngrow, hist,
                        growth_pool_cap(), hc[4], hneed, &unclosed, out));
  if (unclosed > 0)
    return psod_set_error(PSOD_E_UNCLOSED,
                          "sod: %d halos do not close within half the box",
                          unclosed);
  return PSOD_OK;
}

static int sod_fof_run(psod_ctx *ctx, const psod_sod_params *p,
                       const psod_catalogue *cat, int64_t nhalo,
                       const int32_t *d_seed, const int32_t *d_nmem,
                       const psod_halo_out *out) {
  if (nhalo <= 0) return check_sod_args(p, cat, nhalo, d_seed, out);
  SodScratch S;
  int32_t nlarge = 0;
  PSOD_TRY(sod_fof_queue(ctx, p, cat, nhalo, d_seed, d_nmem, out, false, &S,
                         &nlarge));
  return sod_finish_growth(ctx, p, cat, nhalo, S, nlarge, false, out);
}

// The device-count form of psod_sod_fof (psod_pass_graph): queues the whole
// pass with no read-back (dev), or finishes one (reads the failures and the
// overflow flag back, runs re-runs).  *overflow (finish only) receives the
// flag of a task count beyond the capacity: the outputs are then invalid.
int psod_sod_fof_dev(psod_ctx *ctx, const psod_sod_params *p,
                     const psod_catalogue *cat, int64_t nhalo,
                     const int32_t *d_seed, const int32_t *d_nmem,
                     const psod_halo_out *out, bool finish,
                     int32_t *overflow) {
  if (nhalo <= 0) return check_sod_args(p, cat, nhalo, d_seed, out);
  SodScratch S;
  if (!finish) {
    int32_t nlarge = 0;
    return sod_fof_queue(ctx, p, cat, nhalo, d_seed, d_nmem, out, true, &S,
                         &nlarge);
  }
  PSOD_TRY(sod_scratch(ctx, nhalo, &S));
  int32_t ov = 0;
  PSOD_HIP(hipMemcpyAsync(&ov, S.counters + 36, 4, hipMemcpyDeviceToHost,
                          ctx->stream));
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  if (overflow) *overflow = ov;
  if (ov) return PSOD_OK;
  return sod_finish_growth(ctx, p, cat, nhalo, S, 0, true, out);
}

// Every return, failed ones included, waits for both streams: the small
// halos' kernel may still be writing the caller's outputs when a later step
// fails, and the caller may free them as soon as this returns.
extern "C" int psod_sod_fof(psod_ctx *ctx, const psod_sod_params *p,
                            const psod_catalogue *cat, int64_t nhalo,
                            const int32_t *d_seed, const int32_t *d_nmem,
                            const psod_halo_out *out) {
  PSOD_TRY(psod_use_device(ctx));
  const int rc = sod_fof_run(ctx, p, cat, nhalo, d_seed, d_nmem, out);
  if (rc != PSOD_OK) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamSynchronize(ctx->stream);
  }
  return rc;
}
