"""HDF5 halo files (psod_amd/libpsod_h5.so) in the layout of the reference's
writers (sod/modwritehalo.f90:21-209, 503-643), read back through the HDF5
C library itself (ctypes; no h5py in this image), not through the writer.
CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

HDF5 = os.path.join(os.environ.get("PSOD_HDF5", "/opt/conda"), "lib",
                    "libhdf5.so")
pytestmark = pytest.mark.skipif(not os.path.exists(HDF5),
                                reason="HDF5 C library absent")


class H5:
    """Minimal reader on the HDF5 C API."""

    def __init__(self, path):
        L = C.CDLL(HDF5)
        L.H5open()
        hid = C.c_int64
        for f, res, args in (
                ("H5Fopen", hid, [C.c_char_p, C.c_uint, hid]),
                ("H5Fclose", C.c_int, [hid]),
                ("H5Dopen2", hid, [hid, C.c_char_p, hid]),
                ("H5Dclose", C.c_int, [hid]),
                ("H5Dget_space", hid, [hid]),
                ("H5Dget_type", hid, [hid]),
                ("H5Dread", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
                ("H5Sget_simple_extent_ndims", C.c_int, [hid]),
                ("H5Sget_simple_extent_dims", C.c_int,
                 [hid, C.POINTER(C.c_uint64), C.c_void_p]),
                ("H5Sget_simple_extent_type", C.c_int, [hid]),
                ("H5Sclose", C.c_int, [hid]),
                ("H5Aopen", hid, [hid, C.c_char_p, hid]),
                ("H5Aclose", C.c_int, [hid]),
                ("H5Aget_type", hid, [hid]),
                ("H5Aread", C.c_int, [hid, hid, C.c_void_p]),
                ("H5Tget_size", C.c_size_t, [hid]),
                ("H5Tget_class", C.c_int, [hid]),
                ("H5Tclose", C.c_int, [hid]),
                ("H5Gopen2", hid, [hid, C.c_char_p, hid]),
                ("H5Gclose", C.c_int, [hid]),
                ("H5Lexists", C.c_int, [hid, C.c_char_p, hid])):
            getattr(L, f).restype = res
            getattr(L, f).argtypes = args
        self.L = L
        self.native = {np.dtype(k): C.c_int64.in_dll(L, v).value for k, v in (
            ("int32", "H5T_NATIVE_INT_g"), ("int64", "H5T_NATIVE_LLONG_g"),
            ("float32", "H5T_NATIVE_FLOAT_g"),
            ("float64", "H5T_NATIVE_DOUBLE_g"))}
        self.f = L.H5Fopen(path.encode(), 0, 0)
        assert self.f >= 0, path

    def close(self):
        self.L.H5Fclose(self.f)

    def exists(self, path):
        return self.L.H5Lexists(self.f, path.encode(), 0) > 0

    def data(self, path, dtype):
        L = self.L
        d = L.H5Dopen2(self.f, path.encode(), 0)
        assert d >= 0, path
        s = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(s)
        dims = (C.c_uint64 * max(nd, 1))()
        L.H5Sget_simple_extent_dims(s, dims, None)
        shape = tuple(dims[i] for i in range(nd))
        t = L.H5Dget_type(d)
        size = L.H5Tget_size(t)
        L.H5Tclose(t)
        out = np.zeros(shape, dtype)
        if out.size:
            assert L.H5Dread(d, self.native[np.dtype(dtype)], 0, 0, 0,
                             out.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Sclose(s)
        L.H5Dclose(d)
        return out, size

    def attr(self, group, name, dtype=None):
        L = self.L
        g = L.H5Gopen2(self.f, group.encode(), 0)
        a = L.H5Aopen(g, name.encode(), 0)
        assert a >= 0, (group, name)
        t = L.H5Aget_type(a)
        size = L.H5Tget_size(t)
        if dtype is None:  # fixed-length string in the file's own type
            buf = C.create_string_buffer(size)
            assert L.H5Aread(a, t, buf) >= 0
            val = buf.raw.decode()
        else:
            v = np.zeros(1, dtype)
            assert L.H5Aread(a, self.native[np.dtype(dtype)],
                             v.ctypes.data_as(C.c_void_p)) >= 0
            val = v[0]
        L.H5Tclose(t)
        L.H5Aclose(a)
        L.H5Gclose(g)
        return val, size


def test_halo_part_and_mass_files(tmp_path):
    from psod_amd import h5out
    rng = np.random.default_rng(3)
    npart = np.array([5, 1, 7], np.int32)
    hid = np.array([11, 402, 77000000000], np.int64)
    n = int(npart.sum())
    pos = rng.random((n, 3), dtype=np.float32)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    pid = rng.integers(1, 1 << 40, n)
    f1 = h5out.h5writehalopart(str(tmp_path), "box82", npart, hid, pos, vel,
                               pid, 128 ** 3, process_id=3)
    assert os.path.basename(f1) == "psod_halo_snap_part_data_box82_00003.h5"
    r = H5(f1)
    try:
        assert r.attr("metadata", "file_type") == ("halo" + " " * 28, 32)
        assert r.attr("metadata", "halo_finder")[0].strip() == "psod"
        assert r.attr("metadata", "created_by")[0].strip() == "psod_snap"
        assert r.attr("metadata", "nhalo_file", np.int32)[0] == 3
        assert r.attr("metadata", "process_id", np.int32)[0] == 3
        assert r.attr("metadata", "constant_mass", np.int32)[0] == 1
        ns, _ = r.data("metadata/npart_simulation", np.int64)
        assert ns.tolist() == [128 ** 3]
        nf, _ = r.data("metadata/npart_file", np.int64)
        assert nf.tolist() == [n]
        ih, size = r.data("data/identity_halo", np.int64)
        assert size == 8 and ih.tolist() == hid.tolist()
        mm, _ = r.data("data/identity_halo_minmax", np.int64)
        assert mm.tolist() == [11, 77000000000]
        off = 0
        for i, h in enumerate(hid):
            g = f"data/halo_{h:019d}"
            assert r.exists(g)
            assert r.attr(g, "npart_halo", np.int32)[0] == npart[i]
            p, size = r.data(g + "/position_part", np.float32)
            assert size == 4 and p.shape == (npart[i], 3)
            np.testing.assert_array_equal(p, pos[off:off + npart[i]])
            v, _ = r.data(g + "/velocity_part", np.float32)
            np.testing.assert_array_equal(v, vel[off:off + npart[i]])
            q, size = r.data(g + "/identity_part", np.int64)
            assert size == 8
            np.testing.assert_array_equal(q, pid[off:off + npart[i]])
            off += npart[i]
    finally:
        r.close()
    cp = rng.random((3, 3))
    cv = rng.normal(size=(3, 3))
    rr = rng.random(3)
    f2 = h5out.h5writehalomass(str(tmp_path), "box82", npart, cp, cv, hid,
                               rr, 128 ** 3)
    assert os.path.basename(f2) == "psod_halo_snap_part_hfprop_box82.h5"
    r = H5(f2)
    try:
        assert r.attr("metadata", "file_type")[0].strip() == "halomass"
        assert r.attr("metadata", "center_type")[0].strip() == "center_of_mass"
        for name, ref, dt, size in (("position_halo", cp, np.float64, 8),
                                    ("velocity_halo", cv, np.float64, 8),
                                    ("identity_halo", hid, np.int64, 8),
                                    ("npart_halo", npart, np.int32, 4),
                                    ("rmax_halo", rr, np.float64, 8)):
            got, sz = r.data("data/" + name, dt)
            assert sz == size, name
            np.testing.assert_array_equal(got, ref, err_msg=name)
    finally:
        r.close()


def test_halo_writer_rejects_inconsistent_counts(tmp_path):
    from psod_amd import h5out
    with pytest.raises(OSError, match="halos hold 3 particles"):
        h5out.h5writehalopart(str(tmp_path), "x", [1, 2], [1, 2],
                              np.zeros((4, 3)), np.zeros((4, 3)),
                              np.arange(4), 8)


def test_greedy_catalogue_files(tmp_path, orc):
    """The greedy pass's output on a 32^3 catalogue (oracle records, the
    format the drop-in returns): every accepted halo is written with exactly
    its `masse` particles, those `structure` marks, nearest first."""
    from psod_amd import h5out
    nx = 32
    pos, vel, lab, ids, tab = orc.generate(orc.gen_params(nx ** 3, 40, nlo=20,
                                                          nhi=3000, seed=7))
    ids = ids + 1  # RAMSES ids start at 1; structure uses 0 for no halo
    rho = orc.cicm1(pos, orc.cic(pos, 2 * nx, xxmin=-0.5), xxmin=-0.5)
    _, isort, nnew = orc.tri_cic(rho)
    g = orc.sod_greedy(pos, ids, isort, nnew, orc.greedy_params(nx,
                                                                ref_order=0))
    h = g["halos"]
    acc = h["accepted"] == 1
    assert acc.sum() >= 5
    f1, f2 = h5out.write_greedy(str(tmp_path), "c32", g, pos, vel, ids, nx)
    r = H5(f1)
    try:
        ih, _ = r.data("data/identity_halo", np.int64)
        np.testing.assert_array_equal(ih, h["iamas"][acc])
        for k, (hid, m) in enumerate(zip(h["iamas"][acc], h["masse"][acc])):
            g_ = f"data/halo_{hid:019d}"
            assert r.attr(g_, "npart_halo", np.int32)[0] == m
            q, _ = r.data(g_ + "/identity_part", np.int64)
            np.testing.assert_array_equal(np.sort(q),
                                          np.sort(ids[g["structure"] == hid]))
            p, _ = r.data(g_ + "/position_part", np.float32)
            c = np.array([h["xc"][acc][k], h["yc"][acc][k],
                          h["zc"][acc][k]]) / nx
            d = np.abs(p.astype(np.float64) - c)
            d = np.minimum(d, 1 - d)
            r2 = (d * d).sum(axis=1)
            assert np.all(np.diff(r2) >= 0)
            # every member lies inside R
            assert r2.max() <= (h["rayon"][acc][k] / nx) ** 2 * (1 + 1e-9)
    finally:
        r.close()
    r = H5(f2)
    try:
        rm, _ = r.data("data/rmax_halo", np.float64)
        np.testing.assert_array_equal(rm, h["rayon"][acc] / nx)
        nh, _ = r.data("data/npart_halo", np.int32)
        np.testing.assert_array_equal(nh, h["masse"][acc])
    finally:
        r.close()


def test_h5_library_exports_every_declared_symbol():
    import re
    from psod_amd import h5out
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "include", "psod_h5.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(psod_h5_[a-z0-9_]+)\s*\(", src)))
    assert len(names) == 3
    lib = C.CDLL(h5out.LIBPATH)
    assert all(hasattr(lib, n) for n in names), names
