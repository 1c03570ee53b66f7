/*
 * psod_oracle.h -- FP64 CPU restatement of the Psod SOD pass.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in psod_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * Every routine cites the reference routine it restates
 * (sod/sod_profile_rdfof_parallel_hdf5.f90 = "main file").
 *
 * Conventions
 *   - positions: float32 xyz triplets (Fortran posf(3,N)), box units [0,1)
 *   - grids: x fastest, index (iz*ng + iy)*ng + ix
 *   - cell keys of the blocked cell order: see or_cell_key()
 *   - particle indices are 0-based unless a comment says "1-based"
 */
#ifndef PSOD_ORACLE_H
#define PSOD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_OK = 0, OR_E_ARG = 1, OR_E_CAPACITY = 3, OR_E_UNCLOSED = 4 };

/* ---- a2: CIC deposit (main file:1077-1265, periodic branch) ------------- */
void or_cic(int64_t npart, const float *pos, const float *mass, double xxmin,
            int32_t ng, double *grid);

/* ---- a3: CIC^-1 (main file:1268-1390) ------------------------------------ */
void or_cicm1(int64_t npart, const float *pos, double xxmin, int32_t ng,
              const double *grid, double *rho);

/* ---- a11: quick_sort (main file:2477-2571); order is 1-based ------------- */
void or_quick_sort(int64_t n, double *list, int32_t *order);

/* ---- a4: tri_cic (main file:998-1037); returns npart_new (1-based count) - */
int64_t or_tri_cic(int64_t npart, const double *rho, double *pdens,
                   int32_t *isort, double threshold);

/* ---- a5: cell bucketing as a stable counting sort -------------------------
 * Cell of a particle: floor(x*nx) per axis (x*nx is exact in float32 for
 * power-of-two nx), wrapped into [0,nx).  Cells are laid out in the blocked
 * order of or_cell_key().  cell_start has ncells+1 entries; perm[k] is the
 * original index of the k-th particle in cell order, ties by index.        */
int64_t or_cell_key(int32_t nx, int32_t bs, int32_t ix, int32_t iy, int32_t iz);
void or_cell_of(float x, int32_t nx, int32_t *ic);
int or_cell_sort(int64_t npart, const float *pos, int32_t nx, int32_t bs,
                 int64_t *cell_start, int32_t *perm);

/* ---- a6-a10: greedy spherover (main file:1396-2201) ---------------------- */
typedef struct {
  int32_t nx;          /* SOD cells per side; positions are scaled by nx      */
  int32_t ncut;        /* -sel: gather half-width in cells (default 4)        */
  int32_t ncut2;       /* neighbour half-width for the centre search (2)      */
  int32_t ncut3;       /* candidate half-width for the centre search (1)      */
  double seuil;        /* -del: Delta (default 200)                           */
  double seuil2;       /* -dl2: second threshold (default 800)                */
  double mmin;         /* -min: minimum mass (default 100)                    */
  int64_t nsel;        /* -nse: selection capacity                            */
  double xxmin;        /* zone origin in box units (single zone: -0.5)        */
  int32_t local_nx;    /* key stride of the reference cell key (single: 2nx)  */
  double zmin[3];      /* working zone in scaled units: accept (zmin, zmax]   */
  double zmax[3];
  int32_t nbins;       /* profile points (60)                                 */
  double rmin_frac;    /* first profile radius / R (0.01)                     */
  double rmax_frac;    /* last profile radius / R (1)                         */
  int32_t ref_order;   /* 1: within-cell order from quick_sort of the keys    */
} or_greedy_params;

typedef struct {
  int64_t ipart;       /* 1-based rank of the seed in the density order       */
  int64_t seed;        /* seed particle (0-based)                             */
  int64_t centre;      /* centre particle (0-based)                           */
  int64_t iamas;       /* id of the centre particle                           */
  int64_t masse;       /* particles inside R                                  */
  int64_t masse2;      /* last count above seuil2 (0 if never above)          */
  int64_t npart_sel;   /* gathered alive particles                            */
  double amass;        /* mass inside R                                       */
  double xc, yc, zc;   /* centre, scaled units                                */
  double rayon;        /* R, scaled units (cells)                             */
  double overdens, overdens2;
  int32_t accepted;
} or_greedy_halo;

/* Returns OR_OK, OR_E_CAPACITY (reference: "Increase Nsel") or
 * OR_E_UNCLOSED (reference: "Increase selection radius").  *nrec halos are
 * written (accepted or not); rad_bins/mass_bins are nbins per record.
 * alive[] (0/1) and structure[] (id of the accepted halo, 0 otherwise)
 * return the final exclusion state.                                          */
int or_sod_greedy(int64_t npart, const float *pos, const int64_t *ids,
                  const float *mass, const int32_t *isort, int64_t npart_new,
                  const or_greedy_params *p, int64_t maxrec, int64_t *nrec,
                  or_greedy_halo *rec, double *rad_bins, int64_t *mass_bins,
                  uint8_t *alive, int64_t *structure);

/* ---- FoF-seeded independent mode ----------------------------------------- */
#define OR_MAXDELTA 4
typedef struct {
  int32_t nx;
  int32_t ndelta;
  double delta[OR_MAXDELTA];
  double delta2;        /* <= 0: no second threshold                          */
  int32_t nbins;
  double rmin_frac, rmax_frac;
  double mmin;          /* acceptance mass                                     */
  double mpart;         /* particle mass in units of rho_bar*cell^3 (nx^3/N)   */
  int32_t nmin_count;   /* first crossing is searched from this count (10)     */
  double rsearch;       /* centre-search radius, cells (1)                     */
  double eps;           /* centre-search softening, cells (1/64)               */
  double tol;           /* relative margin below which a decision is near-tie  */
  double tol_abs;       /* |r - R| below which a particle is a boundary one   */
} or_fof_params;

typedef struct {
  int64_t seed, centre, centre_id;
  int32_t status;       /* OR_OK or OR_E_UNCLOSED                              */
  int32_t robust_centre;/* 1 if the argmax cannot flip under float32 error     */
  int32_t w;            /* reach (half-width in cells) that closed the sphere  */
  int32_t accepted;
  int64_t ncand, nneigh, ngather;
  double s_best, s_second;
  int64_t count[OR_MAXDELTA];
  double radius[OR_MAXDELTA];    /* R (cells), reference discrete rule        */
  double overdens[OR_MAXDELTA];
  double rint[OR_MAXDELTA];      /* log-log interpolated crossing (cells)     */
  double mint[OR_MAXDELTA];      /* Delta * 4pi/3 * rint^3 (particle masses)  */
  double margin[OR_MAXDELTA];    /* min_k |O(k)/Delta - 1| over the walk      */
  double rgap[OR_MAXDELTA];      /* relative gap between R = r_{k*-1} and the */
                                 /* sorted distances beside it (r_{k*-2},     */
                                 /* r_{k*}): below float resolution another   */
                                 /* particle may rank as the boundary one     */
  int64_t count2;
  double overdens2, margin2;
  double profile_margin;
  int64_t nboundary;
  double com_pos[3];             /* box units                                 */
  double com_vel[3];
} or_fof_halo;

/* seeds[h] = member of label h with the largest rho, ties to the smallest
 * index; nmem[h] = member count.  Labels < 0 are background.                */
void or_fof_seeds(int64_t npart, const int32_t *labels, const double *rho,
                  int64_t nhalo, int64_t *seeds, int64_t *nmem);

/* cell_start/perm from or_cell_sort with block side bs.                     */
int or_sod_fof(int64_t npart, const float *pos, const float *vel,
               const int64_t *ids, const int64_t *cell_start,
               const int32_t *perm, int32_t bs, int64_t nhalo,
               const int64_t *seeds, const int64_t *nmem,
               const or_fof_params *p, or_fof_halo *out, int64_t *profile);

/* The growth of or_sod_fof (main file:2000-2166 and the FoF-mode outputs)
 * around a given centre particle per halo: centres[h] replaces the centre
 * search, whose O(n^2) cost is out of reach for halos of 10^6 members.    */
int or_sod_fof_at(int64_t npart, const float *pos, const float *vel,
                  const int64_t *ids, const int64_t *cell_start,
                  const int32_t *perm, int32_t bs, int64_t nhalo,
                  const int64_t *seeds, const int64_t *nmem,
                  const int64_t *centres, const or_fof_params *p,
                  or_fof_halo *out, int64_t *profile);

/* ---- a12: centre of mass, Pfof/trunk/common/src/modhalo.f90:377-437 ------- */
void or_computecom(int64_t n, const float *pos, const float *vel,
                   double *compos, double *comvel);

/* ---- Delta conversions (haloanalyzer_new/src/modanalyzeengines.f90:311-317)*/
double or_delta_crit_to_mean(double delta_c, double omega_m);
double or_bryan_norman_mean(double omega_m);

/* ---- synthetic catalogue (same spec as the device generator) -------------- */
typedef struct {
  int64_t npart;
  int64_t nclump;
  double nlo, nhi, alpha;      /* clump sizes: N^-alpha on [nlo, nhi]        */
  double clo, chi;             /* concentration U[clo, chi]                  */
  double trunc;                /* truncation radius in units of r200 (2)     */
  uint64_t seed;
  int64_t ncentral;            /* 0 or members of a central clump at 0.5^3   */
  double ccentral;
  double sig_clump, sig_bulk, sig_bg;  /* velocity dispersions, km/s         */
} or_gen_params;

/* Clump table: sizes, prefix offsets (nclump+1), centres (3*nclump),
 * concentrations, r200 (box units).  Returns the number of clump particles. */
int64_t or_gen_clumps(const or_gen_params *g, int64_t *size, int64_t *offset,
                      double *centre, double *conc, double *r200,
                      double *bulk);
void or_gen_particles(const or_gen_params *g, const int64_t *offset,
                      const double *centre, const double *conc,
                      const double *r200, const double *bulk, int64_t i0,
                      int64_t i1, float *pos, float *vel, int32_t *labels,
                      int64_t *ids);

#ifdef __cplusplus
}
#endif
#endif
