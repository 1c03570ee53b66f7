// gen.hip -- seeded synthetic catalogue: NFW clumps on a uniform background.
//
// No reference counterpart: the benchmark input.  Clump sizes follow
// dN/dN ~ N^-alpha on [nlo, nhi] (inverse CDF), concentrations U[clo, chi],
// clumps truncated at trunc*r200 with M200 = N mu(c)/mu(trunc c),
// mu(x) = ln(1+x) - x/(1+x), radii by 60-step bisection of mu, isotropic
// directions, periodic wrap, float32 rounding with 1.0f mapped to 0.
// Draws come from a counter-based hash (SplitMix64 finaliser) of
// (seed, stream, index), so any particle is generated independently.
#include <cmath>

#include "psod_internal.h"

__host__ __device__ inline uint64_t g_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t g_rng(uint64_t seed, uint64_t stream,
                                          uint64_t ctr) {
  return g_mix64(g_mix64(seed ^ g_mix64(stream)) + ctr);
}
__host__ __device__ inline double g_u01(uint64_t s, uint64_t st, uint64_t c) {
  return (double)(g_rng(s, st, c) >> 11) * 0x1.0p-53;
}
__host__ __device__ inline double g_u01open(uint64_t s, uint64_t st,
                                            uint64_t c) {
  return ((double)(g_rng(s, st, c) >> 11) + 0.5) * 0x1.0p-53;
}
__host__ __device__ inline double g_gauss(uint64_t s, uint64_t s1, uint64_t s2,
                                          uint64_t c) {
  const double a = g_u01open(s, s1, c), b = g_u01(s, s2, c);
  return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * b);
}
__host__ __device__ inline double g_mu(double x) {
  return log1p(x) - x / (1.0 + x);
}

enum { ST_SIZE = 1, ST_CONC = 2, ST_CX = 3, ST_BULK = 6, ST_R = 16,
       ST_MU = 17, ST_PHI = 18, ST_BG = 19, ST_V = 22 };

extern "C" int64_t psod_gen_clumps(const psod_gen_params *g, int64_t *size,
                                   int64_t *offset, double *centre,
                                   double *conc, double *r200, double *bulk) {
  const int64_t nc = g->nclump + (g->ncentral > 0 ? 1 : 0);
  const double e = 1.0 - g->alpha;
  const double ae = pow(g->nlo, e), be = pow(g->nhi, e);
  offset[0] = 0;
  for (int64_t c = 0; c < nc; c++) {
    const bool central = (g->ncentral > 0 && c == 0);
    double n, cc;
    if (central) {
      n = (double)g->ncentral;
      cc = g->ccentral;
      centre[3 * c] = centre[3 * c + 1] = centre[3 * c + 2] = 0.5;
    } else {
      const double u = g_u01(g->seed, ST_SIZE, (uint64_t)c);
      n = floor(pow(ae + u * (be - ae), 1.0 / e));
      if (n < g->nlo) n = g->nlo;
      if (n > g->nhi) n = g->nhi;
      cc = g->clo + g_u01(g->seed, ST_CONC, (uint64_t)c) * (g->chi - g->clo);
      for (int d = 0; d < 3; d++)
        centre[3 * c + d] = g_u01(g->seed, ST_CX + d, (uint64_t)c);
    }
    size[c] = (int64_t)n;
    conc[c] = cc;
    const double m200 = n * g_mu(cc) / g_mu(g->trunc * cc);
    r200[c] = cbrt(m200 * 3.0 / (200.0 * (double)g->npart * 4.0 * M_PI));
    for (int d = 0; d < 3; d++)
      bulk[3 * c + d] =
          g->sig_bulk * g_gauss(g->seed, ST_BULK + 2 * d, ST_BULK + 2 * d + 1,
                                (uint64_t)c);
    offset[c + 1] = offset[c] + size[c];
  }
  return offset[nc];
}

struct GenClump {
  double cx, cy, cz, conc, r200, bx, by, bz;
};

__global__ void __launch_bounds__(256)
k_generate(psod_gen_params g, const int64_t *__restrict__ offset,
           const GenClump *__restrict__ cl, int64_t nc, int64_t nclp,
           float *__restrict__ pos, float *__restrict__ vel,
           int32_t *__restrict__ labels, int64_t *__restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.npart) return;
  double x[3], v[3] = {0.0, 0.0, 0.0};
  int32_t lab = -1;
  const uint64_t u = (uint64_t)i;
  if (i < nclp) {
    int64_t lo = 0, hi = nc;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (offset[mid] <= i) lo = mid; else hi = mid;
    }
    const GenClump c = cl[lo];
    lab = (int32_t)lo;
    const double target = g_u01(g.seed, ST_R, u) * g_mu(g.trunc * c.conc);
    double a = 0.0, b = g.trunc * c.conc;
    for (int it = 0; it < 60; it++) {
      const double m = 0.5 * (a + b);
      if (g_mu(m) < target) a = m; else b = m;
    }
    const double r = 0.5 * (a + b) * (c.r200 / c.conc);
    const double mu = 2.0 * g_u01(g.seed, ST_MU, u) - 1.0;
    const double phi = 2.0 * M_PI * g_u01(g.seed, ST_PHI, u);
    const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
    x[0] = c.cx + r * st * cos(phi);
    x[1] = c.cy + r * st * sin(phi);
    x[2] = c.cz + r * mu;
    if (vel) {
      v[0] = c.bx + g.sig_clump * g_gauss(g.seed, ST_V + 0, ST_V + 1, u);
      v[1] = c.by + g.sig_clump * g_gauss(g.seed, ST_V + 2, ST_V + 3, u);
      v[2] = c.bz + g.sig_clump * g_gauss(g.seed, ST_V + 4, ST_V + 5, u);
    }
  } else {
    for (int d = 0; d < 3; d++) x[d] = g_u01(g.seed, ST_BG + d, u);
    if (vel)
      for (int d = 0; d < 3; d++)
        v[d] = g.sig_bg * g_gauss(g.seed, ST_V + 2 * d, ST_V + 2 * d + 1, u);
  }
  for (int d = 0; d < 3; d++) {
    const double w = x[d] - floor(x[d]);
    float f = (float)w;
    if (f >= 1.0f) f = 0.0f;
    pos[3 * i + d] = f;
    if (vel) vel[3 * i + d] = (float)v[d];
  }
  if (labels) labels[i] = lab;
  if (ids) ids[i] = i;
}

extern "C" int psod_generate(psod_ctx *ctx, const psod_gen_params *g,
                             const int64_t *offset, const double *centre,
                             const double *conc, const double *r200,
                             const double *bulk, float *d_pos, float *d_vel,
                             int32_t *d_labels, int64_t *d_ids) {
  PSOD_TRY(psod_use_device(ctx));
  if (!g || !offset || !d_pos || g->npart <= 0 || g->npart >= (1ll << 31))
    return psod_set_error(PSOD_E_ARG, "psod_generate: bad arguments");
  const int64_t nc = g->nclump + (g->ncentral > 0 ? 1 : 0);
  const int64_t nclp = offset[nc];
  if (nclp > g->npart)
    return psod_set_error(PSOD_E_ARG,
                          "psod_generate: clumps hold %lld > npart %lld",
                          (long long)nclp, (long long)g->npart);
  const size_t bo = (size_t)(nc + 1) * sizeof(int64_t);
  const size_t bc = (size_t)(nc > 0 ? nc : 1) * sizeof(GenClump);
  char *dev = (char *)psod_scratch(ctx, S_MISC, bo + bc + 256);
  if (!dev) return PSOD_E_NOMEM;
  int64_t *d_off = (int64_t *)dev;
  GenClump *d_cl = (GenClump *)(dev + ((bo + 255) & ~(size_t)255));
  GenClump *h_cl = (GenClump *)malloc(bc);
  for (int64_t c = 0; c < nc; c++) {
    h_cl[c].cx = centre[3 * c];
    h_cl[c].cy = centre[3 * c + 1];
    h_cl[c].cz = centre[3 * c + 2];
    h_cl[c].conc = conc[c];
    h_cl[c].r200 = r200[c];
    h_cl[c].bx = bulk ? bulk[3 * c] : 0.0;
    h_cl[c].by = bulk ? bulk[3 * c + 1] : 0.0;
    h_cl[c].bz = bulk ? bulk[3 * c + 2] : 0.0;
  }
  hipError_t e1 = hipMemcpyAsync(d_off, offset, bo, hipMemcpyHostToDevice,
                                 ctx->stream);
  hipError_t e2 = hipMemcpyAsync(d_cl, h_cl, bc, hipMemcpyHostToDevice,
                                 ctx->stream);
  hipError_t e3 = hipStreamSynchronize(ctx->stream);
  free(h_cl);
  PSOD_HIP(e1);
  PSOD_HIP(e2);
  PSOD_HIP(e3);
  k_generate<<<grid_for(g->npart, 256), 256, 0, ctx->stream>>>(
      *g, d_off, d_cl, nc, nclp, d_pos, d_vel, d_labels, d_ids);
  PSOD_LAUNCH_CHECK();
  PSOD_HIP(hipStreamSynchronize(ctx->stream));
  return PSOD_OK;
}
