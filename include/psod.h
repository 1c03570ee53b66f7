/*
 * psod.h -- C ABI of the MI355X spherical-overdensity (SOD) pass.
 *
 * Reference: fif092/Psod, sod/sod_profile_rdfof_parallel_hdf5.f90 ("main
 * file").  The reference exposes Fortran external subroutines with implicit
 * interfaces (all arguments by reference, bulk data in module globals).
 * This header replaces them with plain C entry points: scalars by value,
 * arrays as pointers, an int status per call.  psod_amd/fortran/psod_sod.f90
 * binds them with ISO_C_BINDING under the reference's subroutine names.
 *
 * Two families of entry points:
 *   psod_host_*   caller-owned HOST arrays (the Fortran drop-in); they upload,
 *                 run the kernels on the context's device and download.
 *   psod_*        DEVICE pointers already resident in HBM (hot path, bench);
 *                 stream-ordered on the context's stream and blocking unless
 *                 the name ends in _async.
 *
 * Conventions: positions are float32 xyz triplets (Fortran posf(3,N)) in box
 * units [0,1).  Grids are x fastest: (iz*ng + iy)*ng + ix.  nx (SOD cells per
 * side) and ng (CIC nodes per side) are powers of two; nx >= 8.  npart <
 * 2^31.  Indices returned are 0-based except where a comment says 1-based.
 * A context is used by one host thread at a time.
 */
#ifndef PSOD_H
#define PSOD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference prints and returns (main file:1681-1693,
 * 2072-2089); here each call returns one of these and psod_last_error()
 * holds the message. */
#define PSOD_OK 0
#define PSOD_E_ARG 1        /* invalid argument                              */
#define PSOD_E_DEVICE 2     /* HIP runtime error / no device                 */
#define PSOD_E_CAPACITY 3   /* reference "Increase Nsel" (main file:1681)    */
#define PSOD_E_UNCLOSED 4   /* reference "Increase selection radius" (:2072) */
#define PSOD_E_NOMEM 5      /* device allocation failed                      */

#define PSOD_MAXDELTA 4
#define PSOD_BLOCK 8        /* cells are stored in 8^3-cell blocks            */

typedef struct psod_ctx psod_ctx;

/* ---- context ------------------------------------------------------------ */
int psod_create(int device, psod_ctx **ctx);
int psod_destroy(psod_ctx *ctx);
const char *psod_last_error(void);
int psod_synchronize(psod_ctx *ctx);
/* hipStream_t of the context, for callers that order their own work on it */
void *psod_stream(psod_ctx *ctx);
/* bytes of device scratch the context currently holds                       */
int64_t psod_scratch_bytes(psod_ctx *ctx);
/* counts of the last psod_sod_fof call, first min(n, 13) of: out[0] halos
 * grown by one wave, out[1] halos grown by one workgroup, out[2] of these
 * re-run with larger collection slices, out[3] the keys per slice of that
 * re-run (0 if none), out[4] centre-search units summed to the end (pilot
 * tasks of 64 candidates and items of 16), out[5] units skipped by their
 * upper bound, out[6] candidate-neighbour pair tests, out[7] centre-search
 * tasks in all, out[8] centre-search time in ns, out[9] items abandoned
 * part way (exact sums plus the bound of the remaining neighbour cells
 * below the best S), out[10] giant halos grown by the staged launches,
 * out[11] their cycles (a cycle per reach tried), out[12] halos of the
 * workgroup path whose growth cube was sorted whole in LDS.
 * No reference counterpart: diagnostics of the execution paths. */
int psod_sod_stats(psod_ctx *ctx, int64_t *out, int32_t n);

/* ---- tunables and the autotuner (no reference counterpart) ----------------
 * Thresholds of the batched pass that the autotuner chooses per device and
 * workload.  0 in a field means the built-in default (shown in brackets).
 *   small_max        particles in the 5^3 cells around the seed up to which a
 *                    halo is done by one wavefront [1024, at most 1024]
 *   bound_min        particles in the 5^3 cells from which the centre search
 *                    bounds its boxes of 16 candidates before summing [1024]
 *   dense_tile       own particles from which a CIC tile is dispatched first
 *                    [16384]
 *   small_wg_per_cu  one-wavefront halo workgroups per CU [8, at most 32]
 *   giant_min        particles in a growth cube from which the halo is grown
 *                    by the staged launches, each of its walks spread over
 *                    the whole GPU [1048576] */
typedef struct {
  int32_t small_max, bound_min, dense_tile, small_wg_per_cu, giant_min;
} psod_tunables;

/* Sets / reads the context's tunables (those of psod_set_tunables override
 * any loaded from a tuning file). */
int psod_set_tunables(psod_ctx *ctx, const psod_tunables *t);
int psod_get_tunables(psod_ctx *ctx, int32_t nx, psod_tunables *t);

/* ---- a5: cell bucketing (main file:1561-1604) ----------------------------
 * Replaces the quick_sort of cell keys + run-length encoding with an integer
 * counting sort.  Cell of a particle: floor(x*nx) per axis, wrapped.  Cells
 * are ordered by 8^3 blocks (key = block*512 + (lz*8+ly)*8+lx, blocks in
 * (bz,by,bx) row-major order).
 *   d_cell_start [nx^3+1]  first sorted slot of each cell key
 *   d_perm       [npart]   original index of each sorted slot
 *   d_rec        [4*npart] sorted records {u,v,w, idx-as-int-bits}: u,v,w is
 *                          the position in cells relative to the origin of the
 *                          particle's 8^3 block (exact in float32)
 *   d_lab_sorted [npart]   labels in sorted order (NULL if d_labels is NULL)
 * The set of particles in each cell is exact; the order inside a cell is not
 * specified (no downstream result depends on it). */
int psod_cell_sort(psod_ctx *ctx, int64_t npart, const float *d_pos,
                   const int32_t *d_labels, int32_t nx, int32_t *d_cell_start,
                   int32_t *d_perm, float *d_rec, int32_t *d_lab_sorted);

/* ---- a2: CIC deposit (main file:1077-1265) ---------------------------------
 * Vertex-centred periodic CIC of equal-mass particles onto ng^3 nodes
 * (ng = nx or 2*nx), normalised to rho/rho_bar like the reference's
 * mpf = 1/N followed by cube/dx^3.  shift = -xxmin*ng (mod ng) relabels the
 * nodes as the reference's zone origin xxmin does (single zone: xxmin=-0.5).
 * Sums are exact 64-bit fixed point, so the result does not depend on the
 * order of the particles. */
int psod_cic_deposit(psod_ctx *ctx, int64_t npart, int32_t nx,
                     const int32_t *d_cell_start, const float *d_rec,
                     int32_t ng, int32_t shift, float *d_grid);

/* No counterpart in the reference.  psod_cic_deposit for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments and result (bit-identical
 * grid).  Its tile list lives in a fixed array of each device, sized for the
 * largest valid grid (2^21 + 8256 int32, about 8.4 MB, held until the
 * process ends; sizes beyond it return PSOD_E_CAPACITY).  It allocates
 * nothing and does not synchronise: the grid is complete when the context's
 * stream (psod_stream) reaches this point.  Every call on one device shares
 * that array, so calls that overlap on one GPU (from two contexts, or two
 * streams) race. */
int psod_cic_deposit_noalloc(psod_ctx *ctx, int64_t npart, int32_t nx,
                             const int32_t *d_cell_start, const float *d_rec,
                             int32_t ng, int32_t shift, float *d_grid);

/* ---- a3 (+ FoF seeds): CIC^-1 (main file:1268-1390) -----------------------
 * Interpolates the grid back to every particle.  Optional outputs:
 *   d_rho_sorted [npart] in sorted order, d_rho [npart] in original order,
 *   d_seed_key [nhalo]: for each label h >= 0 the member with the largest rho
 *   (ties: smallest index) packed as (float bits of rho << 32 | ~index), 0
 *   for a label without members.  The call clears the keys itself on its
 *   own stream before it fills them. */
int psod_cic_interp(psod_ctx *ctx, int64_t npart, int32_t nx,
                    const int32_t *d_cell_start, const float *d_rec,
                    int32_t ng, int32_t shift, const float *d_grid,
                    float *d_rho_sorted, float *d_rho,
                    const int32_t *d_lab_sorted, int64_t nhalo,
                    uint64_t *d_seed_key);

/* Unpacks seed keys into particle indices (-1 for an empty label).          */
int psod_seed_index(psod_ctx *ctx, int64_t nhalo, const uint64_t *d_seed_key,
                    int32_t *d_seed);

/* ---- a4: tri_cic (main file:998-1037) ---------------------------------------
 * Orders particles by decreasing rho (ties: smallest index) -- the order of
 * quick_sort on 1/rho -- and counts seeds: *npart_new = #(rho > threshold)+1
 * (reference threshold 100, main file:1029).  d_isort is 1-based like the
 * reference's isort. */
int psod_density_order(psod_ctx *ctx, int64_t npart, const float *d_rho,
                       double threshold, int32_t *d_isort,
                       int64_t *npart_new);

/* ---- a6-a12: SOD of FoF-seeded halos ------------------------------------- */
typedef struct {
  int32_t nx;            /* SOD cells per side; N = nx^3 gives rho_bar = 1    */
  int32_t ndelta;        /* thresholds, first-crossing rule (main file:2046)  */
  double delta[PSOD_MAXDELTA];
  double delta2;         /* reference -dl2 (last-above rule); <= 0: off       */
  int32_t nbins;         /* profile points                                    */
  double rmin_frac;      /* first profile radius / R_delta[0]                 */
  double rmax_frac;      /* last profile radius / R_delta[0]                  */
  double mmin;           /* acceptance mass (reference -min)                  */
  double mpart;          /* particle mass in rho_bar*cell^3 units (nx^3/N)    */
  int32_t nmin_count;    /* crossing searched from this count (10, :2046)     */
  double rsearch;        /* centre-search radius in cells (1, :1525)          */
  double eps;            /* centre-search softening in cells (1/64, :1526)    */
  double tol_abs;        /* boundary listing: |r - R| <= tol_abs (cells)      */
  int32_t small_max;     /* overrides the small_max tunable when > 0         */
  int32_t flags;         /* PSOD_F_* below                                    */
} psod_sod_params;

#define PSOD_F_VELOCITY 1  /* d_vel is given: centre-of-mass velocity        */
#define PSOD_F_CENTRE_ONLY 2 /* psod_sod_fof_plain: the centre search alone;
                              * growth outputs are zero (no reference
                              * counterpart: the FP64 centre of halos whose
                              * plain growth would take minutes)            */

/* Per-halo outputs (device arrays, nhalo rows). */
typedef struct {
  int32_t *centre;       /* [nhalo]    centre particle (0-based)              */
  int64_t *centre_id;    /* [nhalo]    its id (reference iamas)               */
  int32_t *status;       /* [nhalo]    PSOD_H_* flags                         */
  int32_t *w;            /* [nhalo]    reach in cells that closed the sphere  */
  int32_t *count;        /* [nhalo*4]  particles inside R (reference masse)   */
  double *radius;        /* [nhalo*4]  R in cells (discrete, :2091)           */
  double *overdens;      /* [nhalo*4]  mean overdensity inside R              */
  double *rint;          /* [nhalo*4]  log-log interpolated crossing, cells   */
  double *mint;          /* [nhalo*4]  delta*4pi/3*rint^3, particle masses    */
  int32_t *count2;       /* [nhalo]    reference masse2                       */
  double *overdens2;     /* [nhalo]                                           */
  int32_t *profile;      /* [nhalo*nbins] counts with r < r_j (:2103-2115)    */
  double *com_pos;       /* [nhalo*3]  centre of mass of the R[0] sphere, box */
  double *com_vel;       /* [nhalo*3]  mean velocity inside R[0]              */
  int32_t *ngather;      /* [nhalo]    particles examined by the growth       */
  int32_t *nboundary;    /* [nhalo]    particles with |r - R[0]| <= tol_abs   */
} psod_halo_out;

#define PSOD_H_CLOSED 1    /* every threshold crossed inside the reach        */
#define PSOD_H_ACCEPTED 2  /* M >= mmin and centre in (0, nx] (:2118-2120)    */
#define PSOD_H_LARGE 4     /* processed by the multi-wave path                */
#define PSOD_H_UNCLOSED 8  /* reach would exceed nx/2                         */

/* Catalogue already cell-sorted by psod_cell_sort.                          */
typedef struct {
  int64_t npart;
  int32_t nx;
  const float *pos;          /* [3*npart] original order                      */
  const float *vel;          /* [3*npart] or NULL                             */
  const int64_t *ids;        /* [npart] or NULL (ids = index)                 */
  const int32_t *cell_start; /* [nx^3+1]                                      */
  const float *rec;          /* [4*npart]                                     */
} psod_catalogue;

/* Batched split-precision pass: every halo in one stream of launches,
 * small halos one wavefront each, large halos split over many waves.
 * d_seed [nhalo] (from psod_seed_index), d_nmem [nhalo] member counts (a
 * reach estimate only; may be NULL). */
int psod_sod_fof(psod_ctx *ctx, const psod_sod_params *p,
                 const psod_catalogue *cat, int64_t nhalo,
                 const int32_t *d_seed, const int32_t *d_nmem,
                 const psod_halo_out *out);

/* Plain method: all-FP64 arithmetic, one launch sequence per halo.  Same
 * outputs; the comparator of the batched pass. */
int psod_sod_fof_plain(psod_ctx *ctx, const psod_sod_params *p,
                       const psod_catalogue *cat, int64_t nhalo,
                       const int32_t *d_seed, const int32_t *d_nmem,
                       const psod_halo_out *out);

/* Member counts per label (labels < 0 ignored).                             */
int psod_label_counts(psod_ctx *ctx, int64_t npart, const int32_t *d_labels,
                      int64_t nhalo, int32_t *d_nmem);

/* The whole FoF-seeded pass in one call, the chain of the reference's zone
 * loop (main file:659-744) for labelled halos: psod_cell_sort ->
 * psod_cic_deposit (shift 0) -> psod_cic_interp with seed keys ->
 * psod_seed_index -> psod_label_counts -> psod_sod_fof.  Device pointers;
 * d_vel and d_ids may be NULL; d_grid [ng^3] receives rho/rho_mean when not
 * NULL (else the context keeps it).  Intermediates are context-owned and
 * reused by the next call.  Same status codes as psod_sod_fof. */
int psod_pass(psod_ctx *ctx, const psod_sod_params *p, int64_t npart,
              int32_t nx, int32_t ng, const float *d_pos, const float *d_vel,
              const int64_t *d_ids, const int32_t *d_labels, int64_t nhalo,
              float *d_grid, const psod_halo_out *out);

/* ---- autotuner (no reference counterpart) -------------------------------
 * Chooses the tunables (psod_tunables above) for one workload on the
 * context's device: the FoF-seeded pass is prepared once on the caller's
 * catalogue (as psod_pass does), then each tunable in turn is swept over its candidates with the
 * others at their best so far, timing the stage it affects (CIC deposit or
 * batched SOD, `reps` runs each, the fastest counts).  The choice becomes
 * the context's tunables for grids of this nx; with cache_file not NULL it
 * is also appended to that file as one line
 *   psod-tune <device name> <CUs> <nx> <npart> <nhalo> <5 tunables> <ms>
 * which psod_load_tunables reads back.  *chosen (may be NULL) receives it;
 * stage_ms (may be NULL) receives [best deposit ms, best SOD ms]. */
int psod_autotune(psod_ctx *ctx, const psod_sod_params *p, int64_t npart,
                  int32_t nx, const float *d_pos, const float *d_vel,
                  const int64_t *d_ids, const int32_t *d_labels,
                  int64_t nhalo, int32_t reps, const char *cache_file,
                  psod_tunables *chosen, double *stage_ms);

/* Loads the lines of a tuning file written on a device with the same name
 * and CU count: the last line for each nx applies to later calls on grids of
 * that nx.  psod_create calls it with $PSOD_TUNE_FILE when that is set.
 * Returns PSOD_OK with *nloaded = 0 when the file has no matching line. */
int psod_load_tunables(psod_ctx *ctx, const char *cache_file,
                       int32_t *nloaded);

/* The same pass as psod_pass (same arguments, outputs and status codes),
 * replayed as one HIP graph.  The reference's zone loop (main file:659-744)
 * calls its routines one after the other; here the whole chain -- cell sort,
 * CIC deposit, CIC^-1 and seeds, member counts, batched SOD with its second
 * stream -- is captured once and then launched as one graph.
 * The first call with a given set of arguments (pointers, sizes, params and
 * the tunables in force) runs the pass eagerly, sizes every buffer, and
 * captures the graph; later calls with the same arguments replay it (the
 * inputs' contents may change between calls).  A replay reads back only the
 * growth's failures; halos that need larger slices are then re-run
 * eagerly, and a replay whose centre search needs more tasks than the
 * capture allowed (twice those of the eager pass) is run again eagerly and
 * captured anew on the next call.  *replayed (may be NULL) is 1 when this
 * call replayed the graph.  Blocking, like psod_pass.  No reference
 * counterpart for the graph itself. */
int psod_pass_graph(psod_ctx *ctx, const psod_sod_params *p, int64_t npart,
                    int32_t nx, int32_t ng, const float *d_pos,
                    const float *d_vel, const int64_t *d_ids,
                    const int32_t *d_labels, int64_t nhalo, float *d_grid,
                    const psod_halo_out *out, int32_t *replayed);

/* ---- synthetic catalogue (no reference counterpart: benchmark input) ------
 * Seeded NFW clumps on a uniform background; see or_gen_* in the oracle for
 * the same specification on the CPU. */
typedef struct {
  int64_t npart, nclump;
  double nlo, nhi, alpha, clo, chi, trunc;
  uint64_t seed;
  int64_t ncentral;
  double ccentral;
  double sig_clump, sig_bulk, sig_bg;
} psod_gen_params;

/* Host part: clump table (nclump[+1] rows).  Returns the clumped particles. */
int64_t psod_gen_clumps(const psod_gen_params *g, int64_t *size,
                        int64_t *offset, double *centre, double *conc,
                        double *r200, double *bulk);
/* Device part: particles [0, npart) from a clump table (host arrays).       */
int psod_generate(psod_ctx *ctx, const psod_gen_params *g,
                  const int64_t *offset, const double *centre,
                  const double *conc, const double *r200, const double *bulk,
                  float *d_pos, float *d_vel, int32_t *d_labels,
                  int64_t *d_ids);

/* ---- host-pointer entry points (Fortran drop-in) ---------------------------
 * Each replaces one reference subroutine for a single zone covering the box.
 */

/* The reference's cube and rho are real(8) (main file:22-23): these three
 * compute in FP64 with the reference's expressions and operation order.
 * Node sums are FP64 atomic adds (they differ from the reference's
 * particle-order sums by the rounding of the additions only). */

/* `cic` (main file:1077-1109, deposit :1133-1215, normalise :1263): equal
 * masses (mass may be NULL, giving the reference's mpf = 1/N in real(4),
 * or all equal), nodes at xxmin + i/ng, periodic wrap at ng.
 * grid [ng^3] FP64, x fastest. */
int psod_host_cic(psod_ctx *ctx, int64_t npart, const float *pos,
                  const float *mass, double xxmin, int32_t ng, double *grid);

/* `cicm1` (main file:1268-1390): rho [npart] FP64 at each particle. */
int psod_host_cicm1(psod_ctx *ctx, int64_t npart, const float *pos,
                    double xxmin, int32_t ng, const double *grid,
                    double *rho);

/* `tri_cic` (main file:998-1037): isort [npart] 1-based by ascending
 * pdens = 1/rho (FP64), ties by index; *npart_new = first 1-based rank
 * with pdens >= 1/threshold (at most npart). */
int psod_host_tri_cic(psod_ctx *ctx, int64_t npart, const double *rho,
                      double threshold, int32_t *isort, int64_t *npart_new);

/* Device form of `tri_cic` on FP64 densities (main file:1022-1032), the
 * order psod_host_tri_cic returns. */
int psod_density_order_f64(psod_ctx *ctx, int64_t npart, const double *d_rho,
                           double threshold, int32_t *d_isort,
                           int64_t *npart_new);

/* `spherover` (main file:1396-1400), greedy seeds-with-exclusion semantics.
 * Positions are UNscaled box units; nx, ncut, seuil, seuil2, mmin, nsel as the
 * reference flags; nbins/rmin_frac/rmax_frac set the profile (60, 0.01, 1).
 * Outputs: one row per processed seed (accepted or not), *nrec rows:
 *   rec_i64 [maxrec*6]: ipart(1-based), seed, centre, iamas, masse, masse2
 *   rec_f64 [maxrec*8]: amass, xc, yc, zc (cells), rayon, overdens,
 *                       overdens2, accepted(0/1)
 *   rad_bins [maxrec*nbins], mass_bins [maxrec*nbins]
 *   alive [npart] final exclusion flags; structure [npart] accepted iamas.
 * Returns PSOD_E_CAPACITY / PSOD_E_UNCLOSED like the reference's zone abort. */
typedef struct {
  int32_t nx, ncut;
  double seuil, seuil2, mmin;
  int64_t nsel;
  int32_t nbins;
  double rmin_frac, rmax_frac;
} psod_greedy_params;

int psod_host_spherover(psod_ctx *ctx, int64_t npart, const float *pos,
                        const int64_t *ids, const int32_t *isort,
                        int64_t npart_new, const psod_greedy_params *p,
                        int64_t maxrec, int64_t *nrec, int64_t *rec_i64,
                        double *rec_f64, double *rad_bins, int32_t *mass_bins,
                        uint8_t *alive, int64_t *structure);

#ifdef __cplusplus
}
#endif
#endif
