/*
 * psod_oracle.c -- FP64 CPU restatement of the Psod SOD pass.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 *
 * "main file" = sod/sod_profile_rdfof_parallel_hdf5.f90 of the reference.
 * Built with -O2 -ffp-contract=off so that every FP64 expression is
 * evaluated in the order written here (the reference's order).
 */
#include "psod_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TWOPI_REF 6.283185307179586 /* main file:1425 */

/* ------------------------------------------------------------------------ */
/* a2: CIC deposit, periodic branch of `cic` (main file:1133-1215, 1263).   */
/* The zone of a single process covers the box, so every particle passes   */
/* the ok_part selection (:1140-1148) and every wrapped node lies inside   */
/* the (local_nx+1)^3 cube (:1194-1213); only nodes 0..ng-1 receive mass.   */
/* ------------------------------------------------------------------------ */
static void cic_index(double u, int32_t ng, int32_t *i0, int32_t *i1,
                      double *f) {
  /* ix = ddx truncated, floor correction for negatives (main file:1154-1164) */
  int32_t ix = (int32_t)u;
  if (u < 0.0) ix = ix - 1;
  *f = u - (double)ix;
  int32_t ixp1 = ix + 1;
  /* periodic wrap (main file:1177-1190) */
  if (ixp1 < 0) ixp1 += ng;
  if (ix < 0) ix += ng;
  if (ixp1 >= ng) ixp1 -= ng;
  if (ix >= ng) ix -= ng;
  *i0 = ix;
  *i1 = ixp1;
}

void or_cic(int64_t npart, const float *pos, const float *mass, double xxmin,
            int32_t ng, double *grid) {
  const int64_t ng3 = (int64_t)ng * ng * ng;
  const double dx = 1.0 / (double)ng;
  /* mpf = mpf/nparttotal in real(4) (main file:662) */
  const float mdef = (float)(1.0 / (double)npart);
  memset(grid, 0, sizeof(double) * (size_t)ng3);
  for (int64_t i = 0; i < npart; i++) {
    const double m = (double)(mass ? mass[i] : mdef);
    int32_t ix, iy, iz, ixp1, iyp1, izp1;
    double ddx, ddy, ddz;
    cic_index(((double)pos[3 * i + 0] - xxmin) / dx, ng, &ix, &ixp1, &ddx);
    cic_index(((double)pos[3 * i + 1] - xxmin) / dx, ng, &iy, &iyp1, &ddy);
    cic_index(((double)pos[3 * i + 2] - xxmin) / dx, ng, &iz, &izp1, &ddz);
    const double dex = 1.0 - ddx, dey = 1.0 - ddy, dez = 1.0 - ddz;
#define G(a, b, c) grid[((int64_t)(c) * ng + (b)) * ng + (a)]
    /* same corner order and product order as main file:1194-1213 */
    G(ix, iy, iz) = G(ix, iy, iz) + m * dex * dey * dez;
    G(ixp1, iy, iz) = G(ixp1, iy, iz) + m * ddx * dey * dez;
    G(ix, iyp1, iz) = G(ix, iyp1, iz) + m * dex * ddy * dez;
    G(ixp1, iyp1, iz) = G(ixp1, iyp1, iz) + m * ddx * ddy * dez;
    G(ix, iy, izp1) = G(ix, iy, izp1) + m * dex * dey * ddz;
    G(ixp1, iy, izp1) = G(ixp1, iy, izp1) + m * ddx * dey * ddz;
    G(ix, iyp1, izp1) = G(ix, iyp1, izp1) + m * dex * ddy * ddz;
    G(ixp1, iyp1, izp1) = G(ixp1, iyp1, izp1) + m * ddx * ddy * ddz;
#undef G
  }
  /* cube = cube/(dx*dy*dz) (main file:1263) */
  const double vol = dx * dx * dx;
  for (int64_t k = 0; k < ng3; k++) grid[k] = grid[k] / vol;
}

/* a3: `cicm1` (main file:1318-1385), same indices and weights as `cic`.    */
void or_cicm1(int64_t npart, const float *pos, double xxmin, int32_t ng,
              const double *grid, double *rho) {
  const double dx = 1.0 / (double)ng;
  for (int64_t i = 0; i < npart; i++) {
    int32_t ix, iy, iz, ixp1, iyp1, izp1;
    double ddx, ddy, ddz;
    cic_index(((double)pos[3 * i + 0] - xxmin) / dx, ng, &ix, &ixp1, &ddx);
    cic_index(((double)pos[3 * i + 1] - xxmin) / dx, ng, &iy, &iyp1, &ddy);
    cic_index(((double)pos[3 * i + 2] - xxmin) / dx, ng, &iz, &izp1, &ddz);
    const double dex = 1.0 - ddx, dey = 1.0 - ddy, dez = 1.0 - ddz;
#define G(a, b, c) grid[((int64_t)(c) * ng + (b)) * ng + (a)]
    double r = 0.0;
    r = r + G(ix, iy, iz) * dex * dey * dez;
    r = r + G(ixp1, iy, iz) * ddx * dey * dez;
    r = r + G(ix, iyp1, iz) * dex * ddy * dez;
    r = r + G(ixp1, iyp1, iz) * ddx * ddy * dez;
    r = r + G(ix, iy, izp1) * dex * dey * ddz;
    r = r + G(ixp1, iy, izp1) * ddx * dey * ddz;
    r = r + G(ix, iyp1, izp1) * dex * ddy * ddz;
    r = r + G(ixp1, iyp1, izp1) * ddx * ddy * ddz;
#undef G
    rho[i] = r;
  }
}

/* ------------------------------------------------------------------------ */
/* a11: quick_sort (main file:2477-2571).  The published method of         */
/* Brainerd, Goldberg & Adams (1990) pp. 149-150 with A. Miller's index     */
/* array: middle-element pivot, Hoare scans, interchange sort below 7       */
/* elements.  Written with an explicit stack; the two sub-ranges of a       */
/* partition are disjoint, so the visiting order does not change the        */
/* result.  Arrays are addressed 1-based through L()/O().                   */
/* ------------------------------------------------------------------------ */
void or_quick_sort(int64_t n, double *list, int32_t *order) {
#define L(k) list[(k)-1]
#define O(k) order[(k)-1]
  for (int64_t i = 1; i <= n; i++) O(i) = (int32_t)i;
  if (n <= 1) return;
  int64_t cap = 256, top = 0;
  int64_t *stk = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)cap);
  stk[0] = 1;
  stk[1] = n;
  top = 1;
  while (top > 0) {
    top--;
    const int64_t left = stk[2 * top], right = stk[2 * top + 1];
    if (right < left + 6) {
      /* interchange_sort (main file:2551-2567) */
      for (int64_t i = left; i <= right - 1; i++)
        for (int64_t j = i + 1; j <= right; j++)
          if (L(i) > L(j)) {
            double t = L(i); L(i) = L(j); L(j) = t;
            int32_t it = O(i); O(i) = O(j); O(j) = it;
          }
      continue;
    }
    const double ref = L((left + right) / 2);
    int64_t i = left - 1, j = right + 1;
    for (;;) {
      do { i++; } while (!(L(i) >= ref));
      do { j--; } while (!(L(j) <= ref));
      if (i < j) {
        double t = L(i); L(i) = L(j); L(j) = t;
        int32_t it = O(i); O(i) = O(j); O(j) = it;
      } else if (i == j) {
        i = i + 1;
        break;
      } else {
        break;
      }
    }
    if (top + 2 > cap) {
      cap *= 2;
      stk = (int64_t *)realloc(stk, sizeof(int64_t) * 2 * (size_t)cap);
    }
    if (left < j) { stk[2 * top] = left; stk[2 * top + 1] = j; top++; }
    if (i < right) { stk[2 * top] = i; stk[2 * top + 1] = right; top++; }
  }
  free(stk);
#undef L
#undef O
}

/* a4: tri_cic (main file:1022-1032).  pdens = 1/rho, sorted ascending, so  */
/* isort lists the particles by decreasing density; npart_new is the first  */
/* 1-based rank whose pdens >= 1/threshold (reference threshold 100).  The  */
/* reference scan has no upper guard (:1029); here it stops at npart.       */
int64_t or_tri_cic(int64_t npart, const double *rho, double *pdens,
                   int32_t *isort, double threshold) {
  for (int64_t i = 0; i < npart; i++) pdens[i] = 1.0 / rho[i];
  or_quick_sort(npart, pdens, isort);
  const double lim = 1.0 / threshold;
  int64_t i = 1;
  while (i < npart && pdens[i - 1] < lim) i++;
  return i;
}

/* ------------------------------------------------------------------------ */
/* a5: cells                                                                 */
/* ------------------------------------------------------------------------ */
int64_t or_cell_key(int32_t nx, int32_t bs, int32_t ix, int32_t iy,
                    int32_t iz) {
  const int64_t nb = nx / bs;
  const int64_t blk = ((int64_t)(iz / bs) * nb + (iy / bs)) * nb + (ix / bs);
  const int64_t loc = ((int64_t)(iz % bs) * bs + (iy % bs)) * bs + (ix % bs);
  return blk * (int64_t)bs * bs * bs + loc;
}

void or_cell_of(float x, int32_t nx, int32_t *ic) {
  /* positions are scaled by nx in real(4) (main file:689-690): exact for a
   * power-of-two nx; cell = floor, wrapped like main file:1564-1576.        */
  const float xs = x * (float)nx;
  int32_t i = (int32_t)floorf(xs);
  if (i < 0) i += nx;
  if (i >= nx) i -= nx;
  *ic = i;
}

int or_cell_sort(int64_t npart, const float *pos, int32_t nx, int32_t bs,
                 int64_t *cell_start, int32_t *perm) {
  if (nx <= 0 || bs <= 0 || nx % bs) return OR_E_ARG;
  const int64_t ncell = (int64_t)nx * nx * nx;
  int64_t *key = (int64_t *)malloc(sizeof(int64_t) * (size_t)(npart ? npart : 1));
  memset(cell_start, 0, sizeof(int64_t) * (size_t)(ncell + 1));
  for (int64_t i = 0; i < npart; i++) {
    int32_t c[3];
    for (int d = 0; d < 3; d++) or_cell_of(pos[3 * i + d], nx, &c[d]);
    key[i] = or_cell_key(nx, bs, c[0], c[1], c[2]);
    cell_start[key[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cell_start[c + 1] += cell_start[c];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
  memcpy(fill, cell_start, sizeof(int64_t) * (size_t)ncell);
  for (int64_t i = 0; i < npart; i++) perm[fill[key[i]]++] = (int32_t)i;
  free(fill);
  free(key);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* a6-a10: greedy spherover (main file:1561-2166)                           */
/* ------------------------------------------------------------------------ */
static inline double min_image(double a, double b, double taille) {
  double d = fabs(a - b);   /* main file:2014-2020 */
  return fmin(d, taille - d);
}

static int32_t wrap(int32_t i, int32_t n) {
  if (i >= n) i -= n;
  if (i < 0) i += n;
  return i;
}

typedef struct {
  int32_t nx, L;
  const int64_t *first; /* 0-based start in isort_ngp of each key (key-1)  */
  const int64_t *num;
  const int32_t *isort_ngp; /* 0-based particle index in cell order        */
} greedy_cells;

int or_sod_greedy(int64_t npart, const float *pos, const int64_t *ids,
                  const float *mass, const int32_t *isort, int64_t npart_new,
                  const or_greedy_params *p, int64_t maxrec, int64_t *nrec,
                  or_greedy_halo *rec, double *rad_bins, int64_t *mass_bins,
                  uint8_t *alive, int64_t *structure) {
  const int32_t nx = p->nx, L = p->local_nx;
  const double scale = (double)nx, taille = (double)nx;
  *nrec = 0;
  if (nx <= 0 || L < nx || p->nbins < 2 || p->ncut < 1) return OR_E_ARG;
  const int64_t nkey = (int64_t)L * L * L;

  /* scaled positions: posf = posf*scale (main file:689-690)               */
  float *xs = (float *)malloc(sizeof(float) * 3 * (size_t)npart);
  for (int64_t i = 0; i < 3 * npart; i++)
    xs[i] = (float)((double)pos[i] * scale);

  /* cell key per particle (main file:1562-1577)                           */
  double *keyd = (double *)malloc(sizeof(double) * (size_t)npart);
  int32_t *iso = (int32_t *)malloc(sizeof(int32_t) * (size_t)npart);
  int32_t *ckey = (int32_t *)malloc(sizeof(int32_t) * (size_t)npart);
  for (int64_t i = 0; i < npart; i++) {
    int32_t c[3];
    for (int d = 0; d < 3; d++) {
      int32_t v = (int32_t)floor((double)xs[3 * i + d] - p->xxmin * scale);
      if (v < 0) v += nx;
      if (v >= nx) v -= nx;
      c[d] = v;
    }
    ckey[i] = 1 + c[0] + c[1] * L + c[2] * L * L;
    keyd[i] = (double)ckey[i];
  }
  if (p->ref_order) {
    /* quick_sort of the keys (main file:1580-1582) */
    or_quick_sort(npart, keyd, iso);
    for (int64_t k = 0; k < npart; k++) iso[k] -= 1;
  } else {
    /* stable order: by key, then by index                                  */
    int64_t *cnt = (int64_t *)calloc((size_t)nkey + 2, sizeof(int64_t));
    for (int64_t i = 0; i < npart; i++) cnt[ckey[i] + 1]++;
    for (int64_t k = 0; k <= nkey; k++) cnt[k + 1] += cnt[k];
    for (int64_t i = 0; i < npart; i++) iso[cnt[ckey[i]]++] = (int32_t)i;
    free(cnt);
  }
  /* first_ngp / num_ngp by run length (main file:1589-1604)               */
  int64_t *first = (int64_t *)calloc((size_t)nkey + 1, sizeof(int64_t));
  int64_t *num = (int64_t *)calloc((size_t)nkey + 1, sizeof(int64_t));
  for (int64_t k = 0; k < npart;) {
    const int32_t key = ckey[iso[k]];
    int64_t e = k;
    while (e < npart && ckey[iso[e]] == key) e++;
    first[key] = k;
    num[key] = e - k;
    k = e;
  }
  free(keyd);

  for (int64_t i = 0; i < npart; i++) {
    alive[i] = 1; /* xsort = 1.0d0 (main file:1618) */
    structure[i] = 0;
  }
  int64_t cap = p->nsel;
  int32_t *sel = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  int32_t *sel2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  double *dist = (double *)malloc(sizeof(double) * (size_t)cap);
  int32_t *isel = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
  const float mdef = 1.0f;
  int status = OR_OK;

#define XS(i, d) ((double)xs[3 * (int64_t)(i) + (d)])
#define MASS(i) ((double)(mass ? mass[i] : mdef))
  for (int64_t ipart = 1; ipart <= npart_new && ipart <= npart; ipart++) {
    const int32_t s = isort[ipart - 1] - 1;
    if (!alive[s]) continue; /* main file:1628 */
    double xc = XS(s, 0), yc = XS(s, 1), zc = XS(s, 2);
    int64_t iamas = ids ? ids[s] : s;
    int64_t centre = s;
    int32_t cc[3];
    for (int d = 0; d < 3; d++) {
      int32_t v = (int32_t)floor(XS(s, d) - p->xxmin * scale);
      if (v < 0) v += nx;
      if (v >= nx) v -= nx;
      cc[d] = v;
    }
    /* S1: alive particles of the (2ncut+1)^3 cells (main file:1658-1716)  */
    int64_t nsel1 = 0;
    double mtotsel = 0.0;
    for (int i = -p->ncut; i <= p->ncut && status == OR_OK; i++) {
      const int32_t ii = wrap(i + cc[0], nx);
      for (int j = -p->ncut; j <= p->ncut && status == OR_OK; j++) {
        const int32_t jj = wrap(j + cc[1], nx);
        for (int k = -p->ncut; k <= p->ncut; k++) {
          const int32_t kk = wrap(k + cc[2], nx);
          const int32_t ind = 1 + ii + jj * L + kk * L * L;
          for (int64_t q = first[ind]; q < first[ind] + num[ind]; q++) {
            const int32_t pi = iso[q];
            if (!alive[pi]) continue;
            if (nsel1 + 1 > cap) { status = OR_E_CAPACITY; break; }
            sel[nsel1++] = pi;
            mtotsel = mtotsel + MASS(pi);
          }
          if (status != OR_OK) break;
        }
      }
    }
    if (status != OR_OK) break;
    /* S2: alive particles of the (2ncut2+1)^3 cells (main file:1720-1755)  */
    int64_t nsel2 = 0;
    for (int i = -p->ncut2; i <= p->ncut2; i++) {
      const int32_t ii = wrap(i + cc[0], nx);
      for (int j = -p->ncut2; j <= p->ncut2; j++) {
        const int32_t jj = wrap(j + cc[1], nx);
        for (int k = -p->ncut2; k <= p->ncut2; k++) {
          const int32_t kk = wrap(k + cc[2], nx);
          const int32_t ind = 1 + ii + jj * L + kk * L * L;
          for (int64_t q = first[ind]; q < first[ind] + num[ind]; q++)
            if (alive[iso[q]]) sel2[nsel2++] = iso[q];
        }
      }
    }
    if (!(mtotsel >= p->mmin)) continue; /* main file:1759 */

    /* centre search, search_max_npart_in_sphere with self-binding weight  */
    /* (main file:1799-1875): rayon_min = 1 cell, smoothing = 1/64 cell     */
    const double rmin2 = 1.0 * 1.0;
    const double eps = (scale / nx) / 64.0;
    const double eps2 = eps * eps;
    double smax = 0.0;
    for (int i = -p->ncut3; i <= p->ncut3; i++) {
      const int32_t ii = wrap(i + cc[0], nx);
      for (int j = -p->ncut3; j <= p->ncut3; j++) {
        const int32_t jj = wrap(j + cc[1], nx);
        for (int k = -p->ncut3; k <= p->ncut3; k++) {
          const int32_t kk = wrap(k + cc[2], nx);
          const int32_t ind = 1 + ii + jj * L + kk * L * L;
          for (int64_t q = first[ind]; q < first[ind] + num[ind]; q++) {
            const int32_t pc = iso[q];
            if (!alive[pc]) continue;
            const double xn = XS(pc, 0), yn = XS(pc, 1), zn = XS(pc, 2);
            double sacc = 0.0;
            for (int64_t t = 0; t < nsel2; t++) {
              const int32_t pj = sel2[t];
              const double ddx = min_image(XS(pj, 0), xn, taille);
              const double ddy = min_image(XS(pj, 1), yn, taille);
              const double ddz = min_image(XS(pj, 2), zn, taille);
              const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
              if (d2 <= rmin2) sacc = sacc + 1.0 / sqrt(d2 + eps2);
            }
            if (sacc > smax) {
              xc = xn; yc = yn; zc = zn;
              iamas = ids ? ids[pc] : pc;
              centre = pc;
              smax = sacc;
            }
          }
        }
      }
    }
    /* distances to the centre and sort (main file:2002-2030)              */
    for (int64_t t = 0; t < nsel1; t++) {
      const int32_t pj = sel[t];
      const double ddx = min_image(XS(pj, 0), xc, taille);
      const double ddy = min_image(XS(pj, 1), yc, taille);
      const double ddz = min_image(XS(pj, 2), zc, taille);
      dist[t] = ddx * ddx + ddy * ddy + ddz * ddz;
    }
    or_quick_sort(nsel1, dist, isel);
#define D(k) dist[(k)-1]
#define MSEL(k) MASS(sel[isel[(k)-1] - 1])
    /* sphere growth (main file:2040-2069)                                 */
    int64_t i = 1;
    double overdens = 2.0 * p->seuil, amass = MSEL(1), rayon = 0.0;
    int64_t masse2 = 0;
    double overdens2 = 0.0;
    const double dc0 = 0.0; /* main file:1765-1771: xc0 == xc here */
    while (i < 10 || (overdens > p->seuil && i < nsel1 &&
                      rayon < ((double)p->ncut - sqrt(dc0)))) {
      i = i + 1;
      if (i > nsel1) { status = OR_E_UNCLOSED; break; }
      rayon = sqrt(D(i));
      const double volume = 2.0 / 3.0 * TWOPI_REF * (rayon * rayon * rayon);
      amass = amass + MSEL(i);
      overdens = amass / volume;
      if (overdens > p->seuil2) {
        masse2 = i;
        overdens2 = (rayon > 1e-30) ? overdens : 0.0;
      }
    }
    if (status != OR_OK) break;
    if (overdens > p->seuil) { status = OR_E_UNCLOSED; break; } /* :2072 */
    rayon = sqrt(D(i - 1));
    {
      const double volume = 2.0 / 3.0 * TWOPI_REF * (rayon * rayon * rayon);
      amass = amass - MSEL(i);
      overdens = (volume > 0.0) ? amass / volume : 0.0;
    }
    const int64_t masse = i - 1;

    if (*nrec >= maxrec) { status = OR_E_CAPACITY; break; }
    or_greedy_halo *h = &rec[*nrec];
    /* log profile (main file:2103-2115)                                   */
    double *rb = rad_bins + (*nrec) * p->nbins;
    int64_t *mb = mass_bins + (*nrec) * p->nbins;
    for (int j = 1; j <= p->nbins; j++) {
      /* the last radius is rayon * rmax_frac exactly (R itself when
         rmax_frac = 1, which main file:2105 reaches only up to rounding) */
      rb[j - 1] = j == p->nbins
                      ? rayon * p->rmax_frac
                      : rayon * exp(log(p->rmin_frac) +
                                    (double)(j - 1) *
                                        (log(p->rmax_frac) - log(p->rmin_frac)) /
                                        ((double)p->nbins - 1.0));
      int64_t k = (j == 1) ? 1 : mb[j - 2];
      if (j > 1) mb[j - 1] = mb[j - 2];
      else mb[0] = 0;
      while (k <= nsel1 && sqrt(D(k)) < rb[j - 1]) {
        mb[j - 1] = k;
        k = k + 1;
      }
    }
    /* acceptance (main file:2118-2120)                                    */
    const int acc = amass >= p->mmin && xc <= p->zmax[0] && xc > p->zmin[0] &&
                    yc <= p->zmax[1] && yc > p->zmin[1] && zc <= p->zmax[2] &&
                    zc > p->zmin[2];
    h->ipart = ipart;
    h->seed = s;
    h->centre = centre;
    h->iamas = iamas;
    h->masse = masse;
    h->masse2 = masse2;
    h->npart_sel = nsel1;
    h->amass = amass;
    h->xc = xc; h->yc = yc; h->zc = zc;
    h->rayon = rayon;
    h->overdens = overdens;
    h->overdens2 = overdens2;
    h->accepted = acc;
    (*nrec)++;
    /* exclusion, accepted or not (main file:2148-2166)                    */
    for (int64_t t = 1; t <= masse; t++) {
      const int32_t pj = sel[isel[t - 1] - 1];
      structure[pj] = acc ? iamas : 0;
      alive[pj] = 0;
    }
#undef D
#undef MSEL
  }
#undef XS
#undef MASS
  free(sel); free(sel2); free(dist); free(isel);
  free(first); free(num); free(iso); free(ckey); free(xs);
  return status;
}

/* ------------------------------------------------------------------------ */
/* FoF-seeded independent mode                                               */
/* ------------------------------------------------------------------------ */
void or_fof_seeds(int64_t npart, const int32_t *labels, const double *rho,
                  int64_t nhalo, int64_t *seeds, int64_t *nmem) {
  for (int64_t h = 0; h < nhalo; h++) { seeds[h] = -1; nmem[h] = 0; }
  for (int64_t i = 0; i < npart; i++) {
    const int32_t h = labels[i];
    if (h < 0 || h >= nhalo) continue;
    nmem[h]++;
    if (seeds[h] < 0 || rho[i] > rho[seeds[h]]) seeds[h] = i;
  }
}

typedef struct { double d2; int32_t idx; } d2idx;
static int cmp_d2idx(const void *a, const void *b) {
  const d2idx *x = (const d2idx *)a, *y = (const d2idx *)b;
  if (x->d2 < y->d2) return -1;
  if (x->d2 > y->d2) return 1;
  return (x->idx > y->idx) - (x->idx < y->idx);
}

/* O(k) of the growth loop: k particle masses over (2/3)(2pi) r^3            */
static inline double overdens_k(int64_t k, double mpart, double r) {
  const double volume = 2.0 / 3.0 * TWOPI_REF * (r * r * r);
  return ((double)k * mpart) / volume;
}

/* centres == NULL: the centre search of each seed (main file:1795-1875);  */
/* otherwise the given centre particle of each halo, and only the growth.  */
static int sod_fof_impl(int64_t npart, const float *pos, const float *vel,
                        const int64_t *ids, const int64_t *cell_start,
                        const int32_t *perm, int32_t bs, int64_t nhalo,
                        const int64_t *seeds, const int64_t *nmem,
                        const int64_t *centres, const or_fof_params *p,
                        or_fof_halo *out, int64_t *profile) {
  const int32_t nx = p->nx;
  const double taille = (double)nx;
  if (nx < 4 || p->ndelta < 1 || p->ndelta > OR_MAXDELTA || p->nbins < 2)
    return OR_E_ARG;
  const double rs2 = p->rsearch * p->rsearch, eps2 = p->eps * p->eps;
  int64_t capn = 1 << 16, capc = 1 << 12, capg = 1 << 16;
  int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)capn);
  int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)capc);
  double *sc = (double *)malloc(sizeof(double) * (size_t)capc);
  double *pc = (double *)malloc(sizeof(double) * (size_t)capc);
  d2idx *g = (d2idx *)malloc(sizeof(d2idx) * (size_t)capg);
  double *rr = (double *)malloc(sizeof(double) * (size_t)capg);
#define XS(i, d) ((double)(pos[3 * (int64_t)(i) + (d)] * (float)nx))
  for (int64_t h = 0; h < nhalo; h++) {
    or_fof_halo *o = &out[h];
    memset(o, 0, sizeof(*o));
    int64_t *prof = profile + h * p->nbins;
    for (int j = 0; j < p->nbins; j++) prof[j] = 0;
    const int64_t s = seeds[h];
    o->seed = s;
    o->status = OR_OK;
    if (s < 0) { o->status = OR_E_ARG; continue; }
    int64_t c0;
    if (centres) {
      c0 = centres[h];
      if (c0 < 0 || c0 >= npart) { o->status = OR_E_ARG; continue; }
      o->robust_centre = 1;
      goto grow;
    }
    int32_t cs[3];
    for (int d = 0; d < 3; d++) or_cell_of(pos[3 * s + d], nx, &cs[d]);
    /* neighbours (5^3 cells) and candidates (3^3 cells), cell walk with x  */
    /* outermost as main file:1720-1755 and :1807-1836                      */
    int64_t nn = 0, nc = 0;
    for (int i = -2; i <= 2; i++)
      for (int j = -2; j <= 2; j++)
        for (int k = -2; k <= 2; k++) {
          const int64_t key = or_cell_key(nx, bs, wrap(cs[0] + i, nx),
                                          wrap(cs[1] + j, nx),
                                          wrap(cs[2] + k, nx));
          for (int64_t q = cell_start[key]; q < cell_start[key + 1]; q++) {
            if (nn >= capn) { capn *= 2; nb = (int32_t *)realloc(nb, sizeof(int32_t) * (size_t)capn); }
            nb[nn++] = perm[q];
          }
        }
    for (int i = -1; i <= 1; i++)
      for (int j = -1; j <= 1; j++)
        for (int k = -1; k <= 1; k++) {
          const int64_t key = or_cell_key(nx, bs, wrap(cs[0] + i, nx),
                                          wrap(cs[1] + j, nx),
                                          wrap(cs[2] + k, nx));
          for (int64_t q = cell_start[key]; q < cell_start[key + 1]; q++) {
            if (nc >= capc) {
              capc *= 2;
              cand = (int32_t *)realloc(cand, sizeof(int32_t) * (size_t)capc);
              sc = (double *)realloc(sc, sizeof(double) * (size_t)capc);
              pc = (double *)realloc(pc, sizeof(double) * (size_t)capc);
            }
            cand[nc++] = perm[q];
          }
        }
    o->ncand = nc;
    o->nneigh = nn;
    /* S(c) = sum over neighbours within rsearch of 1/sqrt(d2+eps2), FP64,  */
    /* first strict maximum wins (main file:1837-1875).  P(c) sums the      */
    /* terms whose d2 lies within tol of the cut: their inclusion can flip  */
    /* under float32 rounding.                                              */
    int64_t best = -1;
    double smax = 0.0;
    for (int64_t c = 0; c < nc; c++) {
      const int32_t pcand = cand[c];
      const double xn = XS(pcand, 0), yn = XS(pcand, 1), zn = XS(pcand, 2);
      double sacc = 0.0, pacc = 0.0;
      for (int64_t t = 0; t < nn; t++) {
        const int32_t pj = nb[t];
        const double ddx = min_image(XS(pj, 0), xn, taille);
        const double ddy = min_image(XS(pj, 1), yn, taille);
        const double ddz = min_image(XS(pj, 2), zn, taille);
        const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (d2 <= rs2) sacc = sacc + 1.0 / sqrt(d2 + eps2);
        if (fabs(d2 - rs2) <= p->tol * rs2) pacc += 1.0 / sqrt(d2 + eps2);
      }
      sc[c] = sacc;
      pc[c] = pacc;
      if (sacc > smax) { smax = sacc; best = c; }
    }
    if (best < 0) { o->status = OR_E_ARG; continue; }
    double second = 0.0, worst_other = -1.0;
    for (int64_t c = 0; c < nc; c++) {
      if (c == best) continue;
      if (sc[c] > second) second = sc[c];
      const double hi = sc[c] + pc[c] + p->tol * sc[c];
      if (hi > worst_other) worst_other = hi;
    }
    o->s_best = smax;
    o->s_second = second;
    o->robust_centre = (smax - pc[best] - p->tol * smax) > worst_other;
    c0 = cand[best];
  grow:
    o->centre = c0;
    o->centre_id = ids ? ids[c0] : c0;
    const double xc = XS(c0, 0), yc = XS(c0, 1), zc = XS(c0, 2);
    int32_t ccell[3];
    for (int d = 0; d < 3; d++) or_cell_of(pos[3 * c0 + d], nx, &ccell[d]);

    /* reach: start from the member count, double until complete            */
    const double nest = (double)(nmem ? nmem[h] : 100);
    const double rest = cbrt(nest * p->mpart * 3.0 /
                             (4.0 * M_PI * p->delta[0]));
    int32_t w = (int32_t)ceil(p->rmax_frac * rest * 1.25) + 1;
    if (w < 2) w = 2;
    for (;;) {
      if (2 * w + 1 > nx) { o->status = OR_E_UNCLOSED; break; }
      int64_t ng = 0;
      for (int i = -w; i <= w; i++)
        for (int j = -w; j <= w; j++)
          for (int k = -w; k <= w; k++) {
            const int64_t key = or_cell_key(nx, bs, wrap(ccell[0] + i, nx),
                                            wrap(ccell[1] + j, nx),
                                            wrap(ccell[2] + k, nx));
            for (int64_t q = cell_start[key]; q < cell_start[key + 1]; q++) {
              if (ng >= capg) {
                capg *= 2;
                g = (d2idx *)realloc(g, sizeof(d2idx) * (size_t)capg);
                rr = (double *)realloc(rr, sizeof(double) * (size_t)capg);
              }
              const int32_t pj = perm[q];
              const double ddx = min_image(XS(pj, 0), xc, taille);
              const double ddy = min_image(XS(pj, 1), yc, taille);
              const double ddz = min_image(XS(pj, 2), zc, taille);
              g[ng].d2 = ddx * ddx + ddy * ddy + ddz * ddz;
              g[ng].idx = pj;
              ng++;
            }
          }
      qsort(g, (size_t)ng, sizeof(d2idx), cmp_d2idx);
      for (int64_t k = 0; k < ng; k++) rr[k] = sqrt(g[k].d2);
      o->ngather = ng;
      /* first crossing per Delta: k* = min{k >= nmin_count: O(k) <= Delta} */
      /* (main file:2046 loop condition), R = r_{k*-1} (:2091)               */
      int complete = 1;
      int64_t kstar[OR_MAXDELTA];
      for (int di = 0; di < p->ndelta; di++) {
        kstar[di] = -1;
        for (int64_t k = p->nmin_count; k <= ng; k++) {
          if (rr[k - 1] >= (double)w) break;
          if (overdens_k(k, p->mpart, rr[k - 1]) <= p->delta[di]) {
            kstar[di] = k;
            break;
          }
        }
        if (kstar[di] < 0) complete = 0;
      }
      if (complete) {
        const double R0 = rr[kstar[0] - 2];
        if (p->rmax_frac * R0 > (double)w) complete = 0;
      }
      if (!complete) {
        int32_t wn = 2 * w;
        if (kstar[0] > 0) {
          const int32_t need =
              (int32_t)ceil(p->rmax_frac * rr[kstar[0] - 2]) + 1;
          if (need > wn) wn = need;
        }
        w = wn;
        continue;
      }
      o->w = w;
      for (int di = 0; di < p->ndelta; di++) {
        const int64_t ks = kstar[di];
        const double r1 = rr[ks - 2], r2 = rr[ks - 1];
        const double volume = 2.0 / 3.0 * TWOPI_REF * (r1 * r1 * r1);
        o->count[di] = ks - 1;
        o->radius[di] = r1;
        o->overdens[di] =
            (volume > 0.0) ? ((double)(ks - 1) * p->mpart) / volume : 0.0;
        /* log-log interpolation of the crossing between (r1, O(k*-1)) and   */
        /* (r2, O(k*)), the rule of haloanalyzer_new/src/modfunctions.f90:   */
        /* 556-558 applied to the particle-resolved cumulative profile.      */
        const double o1 = overdens_k(ks - 1, p->mpart, r1);
        const double o2 = overdens_k(ks, p->mpart, r2);
        double lr;
        if (r1 <= 0.0 || !(o1 > o2)) lr = log(r2);
        else
          lr = log(r1) + (log(p->delta[di]) - log(o1)) * (log(r2) - log(r1)) /
                             (log(o2) - log(o1));
        o->rint[di] = exp(lr);
        o->mint[di] = p->delta[di] * (4.0 / 3.0 * M_PI) * o->rint[di] *
                      o->rint[di] * o->rint[di];
        double mg = INFINITY;
        for (int64_t k = p->nmin_count; k <= ks; k++) {
          const double rel =
              fabs(overdens_k(k, p->mpart, rr[k - 1]) / p->delta[di] - 1.0);
          if (rel < mg) mg = rel;
        }
        o->margin[di] = mg;
        double gp = (r2 - r1) / r1;
        if (ks >= 3) gp = fmin(gp, (r1 - rr[ks - 3]) / r1);
        o->rgap[di] = r1 > 0.0 ? gp : INFINITY;
      }
      /* second threshold, last count above delta2 (main file:2062-2069)    */
      o->count2 = 0;
      o->overdens2 = 0.0;
      o->margin2 = INFINITY;
      if (p->delta2 > 0.0) {
        for (int64_t k = 2; k <= kstar[0]; k++) {
          const double ok = overdens_k(k, p->mpart, rr[k - 1]);
          const double rel = fabs(ok / p->delta2 - 1.0);
          if (rel < o->margin2) o->margin2 = rel;
          if (ok > p->delta2) {
            o->count2 = k;
            o->overdens2 = (rr[k - 1] > 1e-30) ? ok : 0.0;
          }
        }
      }
      /* profile: counts of sqrt(d2) < r_j (main file:2103-2115)            */
      const double R0 = o->radius[0];
      double pm = INFINITY;
      for (int j = 0; j < p->nbins; j++) {
        const double rj =
            j == p->nbins - 1
                ? R0 * p->rmax_frac
                : R0 * exp(log(p->rmin_frac) + (double)j *
                                                   (log(p->rmax_frac) -
                                                    log(p->rmin_frac)) /
                                                   ((double)p->nbins - 1.0));
        int64_t lo = 0, hi = ng; /* first index with rr >= rj */
        while (lo < hi) {
          const int64_t mid = (lo + hi) / 2;
          if (rr[mid] < rj) lo = mid + 1; else hi = mid;
        }
        prof[j] = lo;
        if (lo < ng) pm = fmin(pm, fabs(rr[lo] / rj - 1.0));
        if (lo > 0) pm = fmin(pm, fabs(rr[lo - 1] / rj - 1.0));
      }
      o->profile_margin = pm;
      /* boundary particles and centre of mass over the count[0] closest    */
      int64_t nbd = 0;
      for (int64_t k = 0; k < ng; k++)
        if (fabs(rr[k] - R0) <= p->tol_abs) nbd++;
      o->nboundary = nbd;
      double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
      const int64_t nin = o->count[0];
      for (int64_t k = 0; k < nin; k++) {
        const int32_t pj = g[k].idx;
        const double c3[3] = {xc, yc, zc};
        for (int d = 0; d < 3; d++) {
          double dd = XS(pj, d) - c3[d];
          if (dd > 0.5 * taille) dd -= taille;
          if (dd < -0.5 * taille) dd += taille;
          sp[d] += dd;
          if (vel) sv[d] += (double)vel[3 * (int64_t)pj + d];
        }
      }
      const double c3[3] = {xc, yc, zc};
      for (int d = 0; d < 3; d++) {
        double cp = (c3[d] + sp[d] / (double)nin) / taille;
        cp -= floor(cp);
        o->com_pos[d] = cp;
        o->com_vel[d] = sv[d] / (double)nin;
      }
      o->accepted = (o->count[0] * p->mpart >= p->mmin) && xc > 0.0 &&
                    xc <= taille && yc > 0.0 && yc <= taille && zc > 0.0 &&
                    zc <= taille;
      break;
    }
  }
#undef XS
  free(nb); free(cand); free(sc); free(pc); free(g); free(rr);
  return OR_OK;
}

int or_sod_fof(int64_t npart, const float *pos, const float *vel,
               const int64_t *ids, const int64_t *cell_start,
               const int32_t *perm, int32_t bs, int64_t nhalo,
               const int64_t *seeds, const int64_t *nmem,
               const or_fof_params *p, or_fof_halo *out, int64_t *profile) {
  return sod_fof_impl(npart, pos, vel, ids, cell_start, perm, bs, nhalo,
                      seeds, nmem, NULL, p, out, profile);
}

int or_sod_fof_at(int64_t npart, const float *pos, const float *vel,
                  const int64_t *ids, const int64_t *cell_start,
                  const int32_t *perm, int32_t bs, int64_t nhalo,
                  const int64_t *seeds, const int64_t *nmem,
                  const int64_t *centres, const or_fof_params *p,
                  or_fof_halo *out, int64_t *profile) {
  return sod_fof_impl(npart, pos, vel, ids, cell_start, perm, bs, nhalo,
                      seeds, nmem, centres, p, out, profile);
}

/* a12: computecom (Pfof/trunk/common/src/modhalo.f90:400-432), periodic:    */
/* running sums in FP64, a coordinate is unwrapped by one box when it is    */
/* more than half a box from the running mean; final wrap into [0,1].       */
void or_computecom(int64_t n, const float *pos, const float *vel,
                   double *compos, double *comvel) {
  double cp[3], cv[3];
  for (int d = 0; d < 3; d++) { cp[d] = pos[d]; cv[d] = vel ? vel[d] : 0.0; }
  int64_t halom = 1;
  for (int64_t i = 1; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      const double delta = cp[d] / (double)halom - (double)pos[3 * i + d];
      if (fabs(delta) > 0.5) {
        if (delta > 0.0) cp[d] = cp[d] + 1.0;
        if (delta < 0.0) cp[d] = cp[d] - 1.0;
      }
    }
    halom++;
    for (int d = 0; d < 3; d++) {
      cp[d] = cp[d] + (double)pos[3 * i + d];
      if (vel) cv[d] = cv[d] + (double)vel[3 * i + d];
    }
  }
  for (int d = 0; d < 3; d++) {
    compos[d] = cp[d] / (double)n;
    comvel[d] = cv[d] / (double)n;
    if (compos[d] > 1.0) compos[d] -= 1.0;
    if (compos[d] < 0.0) compos[d] += 1.0;
  }
}

/* Delta_crit -> Delta_mean at z=0: Delta_m = Delta_c / Omega_m            */
/* (haloanalyzer_new/src/modanalyzeengines.f90:311-317 with H(z=0)=1).      */
double or_delta_crit_to_mean(double delta_c, double omega_m) {
  return delta_c / omega_m;
}

/* Bryan & Norman (1998), flat LCDM: Delta_c = 18pi^2 + 82x - 39x^2 with    */
/* x = Omega_m(z) - 1; returned relative to the mean density at z = 0.      */
double or_bryan_norman_mean(double omega_m) {
  const double x = omega_m - 1.0;
  const double dc = 18.0 * M_PI * M_PI + 82.0 * x - 39.0 * x * x;
  return dc / omega_m;
}

/* ------------------------------------------------------------------------ */
/* Synthetic catalogue.  Counter-based hash RNG (SplitMix64 finaliser), so  */
/* every particle's draws depend only on (seed, stream, index).             */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rng64(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return mix64(mix64(seed ^ mix64(stream)) + ctr);
}
static inline double u01(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return (double)(rng64(seed, stream, ctr) >> 11) * 0x1.0p-53;
}
static inline double u01open(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return ((double)(rng64(seed, stream, ctr) >> 11) + 0.5) * 0x1.0p-53;
}
static inline double gauss(uint64_t seed, uint64_t s1, uint64_t s2,
                           uint64_t ctr) {
  const double a = u01open(seed, s1, ctr), b = u01(seed, s2, ctr);
  return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * b);
}
static inline double nfw_mu(double x) { return log1p(x) - x / (1.0 + x); }

enum { ST_SIZE = 1, ST_CONC = 2, ST_CX = 3, ST_BULK = 6, ST_R = 16,
       ST_MU = 17, ST_PHI = 18, ST_BG = 19, ST_V = 22 };

int64_t or_gen_clumps(const or_gen_params *g, int64_t *size, int64_t *offset,
                      double *centre, double *conc, double *r200,
                      double *bulk) {
  const int64_t nc = g->nclump + (g->ncentral > 0 ? 1 : 0);
  const double e = 1.0 - g->alpha;
  const double ae = pow(g->nlo, e), be = pow(g->nhi, e);
  offset[0] = 0;
  for (int64_t c = 0; c < nc; c++) {
    const int central = (g->ncentral > 0 && c == 0);
    double n, cc;
    if (central) {
      n = (double)g->ncentral;
      cc = g->ccentral;
      centre[3 * c] = centre[3 * c + 1] = centre[3 * c + 2] = 0.5;
    } else {
      const double u = u01(g->seed, ST_SIZE, (uint64_t)c);
      n = floor(pow(ae + u * (be - ae), 1.0 / e));
      if (n < g->nlo) n = g->nlo;
      if (n > g->nhi) n = g->nhi;
      cc = g->clo + u01(g->seed, ST_CONC, (uint64_t)c) * (g->chi - g->clo);
      for (int d = 0; d < 3; d++)
        centre[3 * c + d] = u01(g->seed, ST_CX + d, (uint64_t)c);
    }
    size[c] = (int64_t)n;
    conc[c] = cc;
    const double m200 = n * nfw_mu(cc) / nfw_mu(g->trunc * cc);
    r200[c] = cbrt(m200 * 3.0 / (200.0 * (double)g->npart * 4.0 * M_PI));
    for (int d = 0; d < 3; d++)
      bulk[3 * c + d] =
          g->sig_bulk * gauss(g->seed, ST_BULK + 2 * d, ST_BULK + 2 * d + 1,
                              (uint64_t)c);
    offset[c + 1] = offset[c] + size[c];
  }
  return offset[nc];
}

void or_gen_particles(const or_gen_params *g, const int64_t *offset,
                      const double *centre, const double *conc,
                      const double *r200, const double *bulk, int64_t i0,
                      int64_t i1, float *pos, float *vel, int32_t *labels,
                      int64_t *ids) {
  const int64_t nc = g->nclump + (g->ncentral > 0 ? 1 : 0);
  const int64_t nclp = offset[nc];
  for (int64_t i = i0; i < i1; i++) {
    double x[3];
    double sig;
    int32_t lab = -1;
    if (i < nclp) {
      int64_t lo = 0, hi = nc; /* last c with offset[c] <= i */
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (offset[mid] <= i) lo = mid; else hi = mid;
      }
      const int64_t c = lo;
      lab = (int32_t)c;
      const double cc = conc[c];
      const double target = u01(g->seed, ST_R, (uint64_t)i) *
                            nfw_mu(g->trunc * cc);
      double a = 0.0, b = g->trunc * cc;
      for (int it = 0; it < 60; it++) {
        const double m = 0.5 * (a + b);
        if (nfw_mu(m) < target) a = m; else b = m;
      }
      const double r = 0.5 * (a + b) * (r200[c] / cc);
      const double mu = 2.0 * u01(g->seed, ST_MU, (uint64_t)i) - 1.0;
      const double phi = 2.0 * M_PI * u01(g->seed, ST_PHI, (uint64_t)i);
      const double st = sqrt(fmax(0.0, 1.0 - mu * mu));
      x[0] = centre[3 * c + 0] + r * st * cos(phi);
      x[1] = centre[3 * c + 1] + r * st * sin(phi);
      x[2] = centre[3 * c + 2] + r * mu;
      sig = g->sig_clump;
      if (vel)
        for (int d = 0; d < 3; d++)
          vel[3 * i + d] = (float)(bulk[3 * c + d] +
                                   sig * gauss(g->seed, ST_V + 2 * d,
                                               ST_V + 2 * d + 1, (uint64_t)i));
    } else {
      for (int d = 0; d < 3; d++) x[d] = u01(g->seed, ST_BG + d, (uint64_t)i);
      sig = g->sig_bg;
      if (vel)
        for (int d = 0; d < 3; d++)
          vel[3 * i + d] =
              (float)(sig * gauss(g->seed, ST_V + 2 * d, ST_V + 2 * d + 1,
                                  (uint64_t)i));
    }
    for (int d = 0; d < 3; d++) {
      double v = x[d] - floor(x[d]);
      float f = (float)v;
      if (f >= 1.0f) f = 0.0f;
      pos[3 * i + d] = f;
    }
    if (labels) labels[i] = lab;
    if (ids) ids[i] = i;
  }
}
