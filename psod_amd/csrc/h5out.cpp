// h5out.cpp -- HDF5 halo files of the SOD pass in the reference's layout
// (sod/modwritehalo.f90 h5writehalopart and mpih5writehalomass; common
// metadata of Pfof/trunk/common/src/modiocommons.f90 h5write_meta_common).
// Host code on the HDF5 C API; the layout is documented in psod_h5.h.
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/psod_h5.h"

namespace {

thread_local char g_err[512] = "";

int fail(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return -1;
}

// the reference's string attributes are Character(len=H5STRLEN = 32)
constexpr size_t H5STRLEN = 32;

// Holds an HDF5 identifier and closes it with the matching call.
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id() = default;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id &) = delete;
  Id &operator=(const Id &) = delete;
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
};

bool attr_str(hid_t loc, const char *name, const char *value) {
  char buf[H5STRLEN];
  memset(buf, ' ', sizeof(buf));
  const size_t n = strlen(value);
  memcpy(buf, value, n < H5STRLEN ? n : H5STRLEN);
  Id t(H5Tcopy(H5T_C_S1), H5Tclose);
  if (!t.ok() || H5Tset_size(t.id, H5STRLEN) < 0) return false;
  Id s(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(loc, name, t.id, s.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return s.ok() && a.ok() && H5Awrite(a.id, t.id, buf) >= 0;
}

// integer attributes: one element (hdf5_write_int4_attr0D)
bool attr_i32(hid_t loc, const char *name, int32_t v) {
  const hsize_t dim = 1;
  Id s(H5Screate_simple(1, &dim, nullptr), H5Sclose);
  Id a(H5Acreate2(loc, name, H5T_NATIVE_INT32, s.id, H5P_DEFAULT, H5P_DEFAULT),
       H5Aclose);
  return s.ok() && a.ok() && H5Awrite(a.id, H5T_NATIVE_INT32, &v) >= 0;
}

bool dset(hid_t loc, const char *name, hid_t type, int rank,
          const hsize_t *dims, const void *data) {
  Id s(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  Id d(H5Dcreate2(loc, name, type, s.id, H5P_DEFAULT, H5P_DEFAULT,
                  H5P_DEFAULT),
       H5Dclose);
  if (!s.ok() || !d.ok()) return false;
  // an empty dataset has nothing to write
  hssize_t n = H5Sget_simple_extent_npoints(s.id);
  return n == 0 || H5Dwrite(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

// metadata group with the common attributes (modiocommons.f90:14-75) and
// the halo finder's own
bool meta_common(hid_t file, const psod_h5_meta *m, const char *file_type,
                 hid_t *group) {
  *group = H5Gcreate2(file, "metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (*group < 0) return false;
  hid_t g = *group;
  const hsize_t one = 1;
  return attr_str(g, "created_by", "psod_snap") &&
         attr_str(g, "svn_version", "psod_amd") &&
         attr_str(g, "simulation_code", "ramses") &&
         attr_str(g, "particle_type", "dark_matter") &&
         attr_i32(g, "process_id", m->process_id) &&
         attr_i32(g, "constant_mass", 1) && attr_str(g, "units", "ramses") &&
         attr_str(g, "pfof_cells_order", "none") &&
         dset(g, "npart_simulation", H5T_NATIVE_INT64, 1, &one,
              &m->npart_simulation) &&
         attr_str(g, "halo_finder", "psod") &&
         attr_str(g, "file_type", file_type);
}

int check_meta(const char *dir, const psod_h5_meta *m) {
  if (!dir || !m || !m->simulation_name)
    return fail("psod_h5: null directory or metadata");
  if (m->process_id < 0 || m->process_id > 99999 || m->nfile < 1)
    return fail("psod_h5: process_id %d, nfile %d", m->process_id, m->nfile);
  return 0;
}

}  // namespace

extern "C" const char *psod_h5_last_error(void) { return g_err; }

extern "C" int psod_h5_write_halo_part(const char *dir, const psod_h5_meta *m,
                                       int32_t nhalo,
                                       const int32_t *npart_halo,
                                       const int64_t *id_halo, int64_t npart,
                                       const float *pos, const float *vel,
                                       const int64_t *id_part) {
  if (check_meta(dir, m)) return -1;
  if (nhalo < 0 || npart < 0 || (nhalo > 0 && (!npart_halo || !id_halo)) ||
      (npart > 0 && (!pos || !vel || !id_part)))
    return fail("psod_h5_write_halo_part: bad arguments");
  int64_t tot = 0;
  for (int32_t i = 0; i < nhalo; i++) {
    if (npart_halo[i] < 0) return fail("npart_halo[%d] = %d", i, npart_halo[i]);
    tot += npart_halo[i];
  }
  if (tot != npart)
    return fail("psod_h5_write_halo_part: halos hold %lld particles, npart "
                "%lld", (long long)tot, (long long)npart);
  char pid[8];
  snprintf(pid, sizeof(pid), "%05d", m->process_id);
  const std::string path = std::string(dir) + "/psod_halo_snap_part_data_" +
                           m->simulation_name + "_" + pid + ".h5";
  Id f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
       H5Fclose);
  if (!f.ok()) return fail("cannot create %s", path.c_str());
  hid_t gm = -1;
  const bool okm = meta_common(f.id, m, "halo", &gm);
  Id meta(gm, H5Gclose);
  const hsize_t one = 1;
  if (!okm || !attr_i32(meta.id, "nfile", m->nfile) ||
      !attr_i32(meta.id, "nhalo_file", nhalo) ||
      !dset(meta.id, "npart_file", H5T_NATIVE_INT64, 1, &one, &npart))
    return fail("%s: metadata", path.c_str());
  Id data(H5Gcreate2(f.id, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
          H5Gclose);
  if (!data.ok()) return fail("%s: data group", path.c_str());
  // the reference writes the halo datasets only when there is a halo
  // (modwritehalo.f90:137)
  if (nhalo == 0) return 0;
  const hsize_t nh = (hsize_t)nhalo, two = 2;
  const int64_t mm[2] = {id_halo[0], id_halo[nhalo - 1]};
  if (!dset(data.id, "identity_halo", H5T_NATIVE_INT64, 1, &nh, id_halo) ||
      !dset(data.id, "identity_halo_minmax", H5T_NATIVE_INT64, 1, &two, mm))
    return fail("%s: halo identities", path.c_str());
  int64_t off = 0;
  for (int32_t i = 0; i < nhalo; i++) {
    char gname[32];
    snprintf(gname, sizeof(gname), "halo_%019lld", (long long)id_halo[i]);
    Id g(H5Gcreate2(data.id, gname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
         H5Gclose);
    const hsize_t n = (hsize_t)npart_halo[i];
    const hsize_t d2[2] = {n, 3};
    if (!g.ok() || !attr_i32(g.id, "npart_halo", npart_halo[i]) ||
        !dset(g.id, "position_part", H5T_NATIVE_FLOAT, 2, d2, pos + 3 * off) ||
        !dset(g.id, "velocity_part", H5T_NATIVE_FLOAT, 2, d2, vel + 3 * off) ||
        !dset(g.id, "identity_part", H5T_NATIVE_INT64, 1, &n, id_part + off))
      return fail("%s: group %s", path.c_str(), gname);
    off += npart_halo[i];
  }
  return 0;
}

extern "C" int psod_h5_write_halo_mass(const char *dir, const psod_h5_meta *m,
                                       int32_t nhalo,
                                       const int32_t *npart_halo,
                                       const double *com_pos,
                                       const double *com_vel,
                                       const int64_t *id_halo,
                                       const double *rmax) {
  if (check_meta(dir, m)) return -1;
  if (nhalo < 0 || (nhalo > 0 && (!npart_halo || !com_pos || !com_vel ||
                                  !id_halo || !rmax)))
    return fail("psod_h5_write_halo_mass: bad arguments");
  const std::string path = std::string(dir) + "/psod_halo_snap_part_hfprop_" +
                           m->simulation_name + ".h5";
  Id f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
       H5Fclose);
  if (!f.ok()) return fail("cannot create %s", path.c_str());
  hid_t gm = -1;
  const bool okm = meta_common(f.id, m, "halomass", &gm);
  Id meta(gm, H5Gclose);
  if (!okm || !attr_str(meta.id, "center_type", "center_of_mass") ||
      !attr_i32(meta.id, "nhalo_file", nhalo))
    return fail("%s: metadata", path.c_str());
  Id data(H5Gcreate2(f.id, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
          H5Gclose);
  const hsize_t nh = (hsize_t)nhalo, d2[2] = {nh, 3};
  if (!data.ok() ||
      !dset(data.id, "position_halo", H5T_NATIVE_DOUBLE, 2, d2, com_pos) ||
      !dset(data.id, "velocity_halo", H5T_NATIVE_DOUBLE, 2, d2, com_vel) ||
      !dset(data.id, "identity_halo", H5T_NATIVE_INT64, 1, &nh, id_halo) ||
      !dset(data.id, "npart_halo", H5T_NATIVE_INT32, 1, &nh, npart_halo) ||
      !dset(data.id, "rmax_halo", H5T_NATIVE_DOUBLE, 1, &nh, rmax))
    return fail("%s: data", path.c_str());
  return 0;
}
