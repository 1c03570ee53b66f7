// sod_common.h -- device code shared by the batched (split-precision) and
// the plain (FP64) SOD passes.
//
// Reference semantics (main file = sod/sod_profile_rdfof_parallel_hdf5.f90):
//   centre search  :1795-1875  S(c) = sum_{d2 <= rsearch^2} 1/sqrt(d2+eps^2)
//                              over the 5^3 cells around the seed; candidates
//                              in the 3^3 cells; first strict maximum wins in
//                              the cell walk order (x outermost).
//   growth         :2040-2100  O(k) = k m / ((2/3)(2pi) r_k^3); first
//                              k >= 10 with O(k) <= Delta; R = r_{k-1}.
//   second Delta   :2062-2069  last k with O(k) > Delta2.
//   profile        :2103-2115  counts of sqrt(d2) < r_j, r_j log spaced.
#pragma once
#include "psod_internal.h"

#define TWOPI_REF 6.283185307179586

struct SodCfg {
  int32_t nx, ndelta, nbins, nmin_count, flags;
  double delta[PSOD_MAXDELTA];
  double delta2, lr0, lr1, rmin_frac, rmax_frac, mmin, mpart;
  double rs2, eps2, tol_abs, nbar;
  float rs2f, eps2f;
  // tunables: one-wave halo cap (particles in the 5^3 cells) and the size
  // from which the centre search bounds its boxes first
  int32_t small_max, bound_min;
  int64_t giant_min;  // growth cube particles from which growth is staged
  int32_t light_cap;  // growth cube particles up to which keys sort in LDS
};

// overdensity of the k closest particles at radius r, in the reference's
// operation order (main file:2047-2050)
__device__ inline double od_k(int64_t k, double mpart, double r) {
  const double volume = 2.0 / 3.0 * TWOPI_REF * (r * r * r);
  return ((double)k * mpart) / volume;
}

// radius below which the k closest particles exceed delta: O(k) > delta iff
// r_k < rad_for(k, delta) (up to the rounding of od_k)
__device__ inline double rad_for(int64_t k, double mpart, double delta) {
  return cbrt(((double)k * mpart) / (delta * (2.0 / 3.0 * TWOPI_REF)));
}

// j-th profile radius (main file:2105), r_j = R exp(lr0 + j (lr1-lr0)/(n-1));
// the last one is R * rmax_frac exactly (R itself for the reference's
// rmax_frac = 1, which the exponential gives only up to rounding)
__device__ inline double prof_radius(const SodCfg &c, double R, int j) {
  if (j == c.nbins - 1) return R * c.rmax_frac;
  return R * exp(c.lr0 + (double)j * (c.lr1 - c.lr0) / ((double)c.nbins - 1.0));
}

// offset (di, dj, dk) of the c-th cell of a (2h+1)^3 walk, x outermost
__device__ inline void walk_off(int c, int h, int *di, int *dj, int *dk) {
  const int s = 2 * h + 1;
  *di = c / (s * s) - h;
  *dj = (c / s) % s - h;
  *dk = c % s - h;
}

// scaled coordinate of a position in [0, nx), its cell and in-cell offset;
// identical to the split performed by the cell sort
__device__ inline void split_coord(float x, int32_t nx, int32_t *cell,
                                   float *off) {
  float xs = x * (float)nx;
  if (xs >= (float)nx) xs -= (float)nx;
  if (xs < 0.0f) xs += (float)nx;
  int32_t c = (int32_t)floorf(xs);
  if (c >= nx) { c -= nx; xs -= (float)nx; }
  *cell = c;
  *off = xs - (float)c;
}

struct CellRef {
  int32_t c[3];   // reference cell
  float f[3];     // reference offset inside it
};

__device__ inline CellRef ref_of_particle(const float *pos, int32_t idx,
                                          int32_t nx) {
  CellRef r;
  for (int d = 0; d < 3; d++)
    split_coord(pos[3 * (int64_t)idx + d], nx, &r.c[d], &r.f[d]);
  return r;
}

// key of the cell at offset (di,dj,dk) from cell c (periodic)
__device__ inline int32_t key_at(int32_t nx, const int32_t *c, int di, int dj,
                                 int dk, int32_t *l) {
  const int32_t x = wrapi(c[0] + di, nx), y = wrapi(c[1] + dj, nx),
                z = wrapi(c[2] + dk, nx);
  l[0] = x & 7;
  l[1] = y & 7;
  l[2] = z & 7;
  return cell_key(nx, x, y, z);
}

// Position of a record relative to a reference point, split precision:
// (cell offset in whole cells) + (in-cell offset - reference offset).
// `loc` is the walked cell's coordinate inside its 8^3 block.
__device__ inline float rel1(float u, int32_t loc, int dcell, float fref) {
  return (float)dcell + ((u - (float)loc) - fref);
}

// The same with fref = 0, written u + (dcell - loc): the record lies in its
// cell loc of the block (u in [loc, loc + 1), or loc = 0), so u - loc is
// exact and both forms round the same real number once.  This one costs a
// conversion of the wave-uniform offset and one add per record.
__device__ inline float cell_rel(float u, int32_t loc, int dcell) {
  return u + (float)(dcell - loc);
}

__device__ inline double rel1d(float u, int32_t loc, int dcell, double fref) {
  return (double)dcell + (((double)u - (double)loc) - fref);
}

// Distance of particle j from the centre in the reference's FP64 arithmetic
// (main file:2014-2020): float32 positions scaled by nx, minimum image per
// axis, d2 = dx*dx + dy*dy + dz*dz without fused operations.  The batched
// pass ranks particles with split float32 distances and uses this for the
// radii that end up in the outputs.
__device__ inline double ref_dist(const float *pos, int32_t j,
                                  const double *xc, int32_t nx) {
  const double taille = (double)nx;
  double d[3];
  for (int a = 0; a < 3; a++) {
    const double xj = (double)(pos[3 * (int64_t)j + a] * (float)nx);
    double dd = fabs(xj - xc[a]);
    d[a] = fmin(dd, taille - dd);
  }
  const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]),
                                        __dmul_rn(d[1], d[1])),
                              __dmul_rn(d[2], d[2]));
  return sqrt(d2);
}

__device__ inline void centre_scaled(const float *pos, int32_t c, int32_t nx,
                                     double *xc) {
  for (int a = 0; a < 3; a++)
    xc[a] = (double)(pos[3 * (int64_t)c + a] * (float)nx);
}

// Fixed point of one centre-search term (value <= 1/eps = 64 = 2^6): units
// of 2^-18, so 64 terms fit a 32-bit partial sum.  Sums of fixed-point terms
// do not depend on the order of the neighbours.
constexpr float SFIX = 262144.0f;  // 2^18
constexpr double SFIX_INV = 1.0 / 262144.0;
__device__ inline uint32_t sterm_fix(float d2, float eps2) {
  // v_rsq_f32 without the denormal-input rescaling of rsqrtf: the argument
  // is at least eps2, a normal float, so both give the same value
  return (uint32_t)(__builtin_amdgcn_rsqf(d2 + eps2) * SFIX + 0.5f);
}

// per-halo plan
struct HaloPlan {
  int32_t seed;
  int32_t nmem;
  int32_t cs[3];
  int32_t w0;       // first reach to try
  int32_t centre;   // -1 until known
  int32_t n5;       // particles in the 5^3 cells around the seed cell
  int32_t klass;    // 0 small, 1 large
  int32_t ntask;    // centre-search tasks (large)
  int32_t task0;
  int32_t pad;
};

__device__ inline int32_t reach_estimate(const SodCfg &c, int32_t nmem) {
  const double nest = (double)(nmem > 0 ? nmem : 100);
  const double rest = cbrt(nest * c.mpart * 3.0 / (4.0 * M_PI * c.delta[0]));
  int32_t w = (int32_t)ceil(c.rmax_frac * rest * 1.25) + 1;
  return w < 2 ? 2 : w;
}

// writes one halo's outputs
struct HaloResult {
  int32_t centre, status, w, count2, ngather, nboundary;
  int64_t centre_id;
  int32_t count[PSOD_MAXDELTA];
  double radius[PSOD_MAXDELTA], overdens[PSOD_MAXDELTA], rint[PSOD_MAXDELTA],
      mint[PSOD_MAXDELTA];
  double overdens2, com_pos[3], com_vel[3];
};

__device__ inline void finish_delta(const SodCfg &c, int di, int64_t ks,
                                    double r1, double r2, HaloResult &res) {
  const double volume = 2.0 / 3.0 * TWOPI_REF * (r1 * r1 * r1);
  res.count[di] = (int32_t)(ks - 1);
  res.radius[di] = r1;
  res.overdens[di] = volume > 0.0 ? ((double)(ks - 1) * c.mpart) / volume : 0.0;
  // log-log interpolation of the crossing (haloanalyzer_new/src/
  // modfunctions.f90:556-558) on the particle-resolved cumulative profile
  const double o1 = od_k(ks - 1, c.mpart, r1), o2 = od_k(ks, c.mpart, r2);
  double lr;
  if (r1 <= 0.0 || !(o1 > o2))
    lr = log(r2);
  else
    lr = log(r1) + (log(c.delta[di]) - log(o1)) * (log(r2) - log(r1)) /
                       (log(o2) - log(o1));
  res.rint[di] = exp(lr);
  res.mint[di] = c.delta[di] * (4.0 / 3.0 * M_PI) * res.rint[di] *
                 res.rint[di] * res.rint[di];
}

__device__ inline void store_result(const psod_halo_out &o, int64_t h,
                                    int nbins, const HaloResult &r) {
  if (o.centre) o.centre[h] = r.centre;
  if (o.centre_id) o.centre_id[h] = r.centre_id;
  if (o.status) o.status[h] = r.status;
  if (o.w) o.w[h] = r.w;
  for (int d = 0; d < PSOD_MAXDELTA; d++) {
    if (o.count) o.count[h * PSOD_MAXDELTA + d] = r.count[d];
    if (o.radius) o.radius[h * PSOD_MAXDELTA + d] = r.radius[d];
    if (o.overdens) o.overdens[h * PSOD_MAXDELTA + d] = r.overdens[d];
    if (o.rint) o.rint[h * PSOD_MAXDELTA + d] = r.rint[d];
    if (o.mint) o.mint[h * PSOD_MAXDELTA + d] = r.mint[d];
  }
  if (o.count2) o.count2[h] = r.count2;
  if (o.overdens2) o.overdens2[h] = r.overdens2;
  for (int d = 0; d < 3; d++) {
    if (o.com_pos) o.com_pos[h * 3 + d] = r.com_pos[d];
    if (o.com_vel) o.com_vel[h * 3 + d] = r.com_vel[d];
  }
  if (o.ngather) o.ngather[h] = r.ngather;
  if (o.nboundary) o.nboundary[h] = r.nboundary;
}

__device__ inline void zero_result(HaloResult &r) {
  r.centre = -1;
  r.status = 0;
  r.w = 0;
  r.count2 = 0;
  r.ngather = 0;
  r.nboundary = 0;
  r.centre_id = -1;
  for (int d = 0; d < PSOD_MAXDELTA; d++) {
    r.count[d] = 0;
    r.radius[d] = r.overdens[d] = r.rint[d] = r.mint[d] = 0.0;
  }
  r.overdens2 = 0.0;
  for (int d = 0; d < 3; d++) r.com_pos[d] = r.com_vel[d] = 0.0;
}

SodCfg make_cfg(const psod_sod_params *p, int64_t npart);
int check_sod_args(const psod_sod_params *p, const psod_catalogue *cat,
                   int64_t nhalo, const int32_t *seed,
                   const psod_halo_out *out);
