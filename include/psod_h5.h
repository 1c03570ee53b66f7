/*
 * psod_h5.h -- HDF5 halo output of the SOD pass (host code, library
 * psod_amd/libpsod_h5.so, linked against the HDF5 C library).
 *
 * Reference: fif092/Psod sod/modwritehalo.f90 (h5writehalopart :21-209,
 * mpih5writehalomass :503-643) as called by the main file (:2264-2410), and
 * the metadata of Pfof/trunk/common/src/modiocommons.f90 (h5write_meta_common
 * :14-75).  The files have the reference's names, groups, dataset names,
 * types and shapes, so haloanalyzer reads them unchanged:
 *   <dir>/psod_halo_snap_part_data_<sim>_<process_id, 5 digits>.h5
 *     metadata/  attributes created_by, svn_version, simulation_code,
 *                particle_type, process_id, constant_mass, units,
 *                pfof_cells_order, halo_finder, file_type = "halo", nfile,
 *                nhalo_file; datasets npart_simulation, npart_file
 *     data/      identity_halo [nhalo] int64, identity_halo_minmax [2],
 *                halo_<id, 19 digits>/ (attribute npart_halo) with
 *                position_part [n][3] float32, velocity_part [n][3] float32,
 *                identity_part [n] int64
 *   <dir>/psod_halo_snap_part_hfprop_<sim>.h5
 *     metadata/  the same common attributes, file_type = "halomass",
 *                center_type = "center_of_mass", nhalo_file
 *     data/      position_halo [nhalo][3] float64, velocity_halo [nhalo][3]
 *                float64, identity_halo [nhalo] int64, npart_halo [nhalo]
 *                int32, rmax_halo [nhalo] float64
 * One process writes both (the reference's process 0 of one zone).
 * String attributes are 32-byte blank-padded strings like the reference's
 * Character(len=H5STRLEN).  The reference writes velocity_halo and rmax_halo
 * as zeros (main file:2271-2276); callers here pass the centre-of-mass
 * velocity and R_Delta.  The reference's pfof parameter and RAMSES info
 * groups (h5write_meta_pfof_parameter, h5write_meta_info_ramses) describe an
 * input snapshot and a pfof run this pass does not have; they are not
 * written.  Returns 0, or -1 with psod_h5_last_error() set.
 */
#ifndef PSOD_H5_H
#define PSOD_H5_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const char *simulation_name;  /* <sim> in the file names               */
  int64_t npart_simulation;     /* metadata/npart_simulation             */
  int32_t process_id;           /* metadata process_id, file name suffix */
  int32_t nfile;                /* metadata nfile (processes writing)    */
} psod_h5_meta;

/* h5writehalopart (modwritehalo.f90:21-209): the particles of each halo.
 * npart_halo [nhalo] (reference haloMass), id_halo [nhalo] (haloID); the
 * particles of halo i are rows off_i .. off_i + npart_halo[i] - 1 of
 * pos/vel [npart][3] and id_part [npart], halos one after the other. */
int psod_h5_write_halo_part(const char *dir, const psod_h5_meta *meta,
                            int32_t nhalo, const int32_t *npart_halo,
                            const int64_t *id_halo, int64_t npart,
                            const float *pos, const float *vel,
                            const int64_t *id_part);

/* mpih5writehalomass (modwritehalo.f90:503-643), one process: per-halo
 * properties.  com_pos, com_vel [nhalo][3]; rmax [nhalo]. */
int psod_h5_write_halo_mass(const char *dir, const psod_h5_meta *meta,
                            int32_t nhalo, const int32_t *npart_halo,
                            const double *com_pos, const double *com_vel,
                            const int64_t *id_halo, const double *rmax);

const char *psod_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
