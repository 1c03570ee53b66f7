"""HDF5 halo files of the SOD pass in the reference's layout (include/psod_h5.h).

``h5writehalopart`` and ``h5writehalomass`` keep the names and argument
meaning of the reference writers (sod/modwritehalo.f90:21-209, 503-643);
``write_greedy`` and ``write_fof`` gather what they need from the outputs of
``psod_amd.spherover`` (the reference-semantics pass, main file:2264-2410)
and of ``Context.sod_fof`` / ``Context.run_pass`` (the FoF-seeded pass).

The writer is host code in psod_amd/libpsod_h5.so, linked against the HDF5 C
library; importing this module without it raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpsod_h5.so")


class Meta(C.Structure):
    _fields_ = [("simulation_name", C.c_char_p),
                ("npart_simulation", C.c_int64), ("process_id", C.c_int32),
                ("nfile", C.c_int32)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing (built by "
                              "__graft_entry__.build() when HDF5 is present)")
        L = C.CDLL(LIBPATH)
        v = C.c_void_p
        L.psod_h5_write_halo_part.restype = C.c_int
        L.psod_h5_write_halo_part.argtypes = [C.c_char_p, C.POINTER(Meta),
                                              C.c_int32, v, v, C.c_int64, v,
                                              v, v]
        L.psod_h5_write_halo_mass.restype = C.c_int
        L.psod_h5_write_halo_mass.argtypes = [C.c_char_p, C.POINTER(Meta),
                                              C.c_int32, v, v, v, v, v]
        L.psod_h5_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise OSError(_load().psod_h5_last_error().decode(errors="replace"))


def _meta(simulation_name, npart_simulation, process_id, nfile):
    return Meta(simulation_name.encode(), int(npart_simulation),
                int(process_id), int(nfile))


def h5writehalopart(outdir, simulation_name, halo_mass, halo_id, part_pos,
                    part_vel, part_id, npart_simulation, process_id=0,
                    nfile=1):
    """h5writehalopart: halo i's particles are the halo_mass[i] rows after
    those of halos 0..i-1.  Returns the file path."""
    hm = np.ascontiguousarray(halo_mass, np.int32)
    hid = np.ascontiguousarray(halo_id, np.int64)
    pp = np.ascontiguousarray(part_pos, np.float32).reshape(-1, 3)
    pv = np.ascontiguousarray(part_vel, np.float32).reshape(-1, 3)
    pid = np.ascontiguousarray(part_id, np.int64)
    m = _meta(simulation_name, npart_simulation, process_id, nfile)
    _check(_load().psod_h5_write_halo_part(
        outdir.encode(), C.byref(m), len(hm), _p(hm), _p(hid), len(pid),
        _p(pp), _p(pv), _p(pid)))
    return os.path.join(outdir, f"psod_halo_snap_part_data_{simulation_name}_"
                                f"{process_id:05d}.h5")


def h5writehalomass(outdir, simulation_name, halo_mass, com_pos, com_vel,
                    halo_id, radius, npart_simulation, process_id=0, nfile=1):
    """mpih5writehalomass on one process.  Returns the file path."""
    hm = np.ascontiguousarray(halo_mass, np.int32)
    cp = np.ascontiguousarray(com_pos, np.float64).reshape(-1, 3)
    cv = np.ascontiguousarray(com_vel, np.float64).reshape(-1, 3)
    hid = np.ascontiguousarray(halo_id, np.int64)
    rr = np.ascontiguousarray(radius, np.float64)
    m = _meta(simulation_name, npart_simulation, process_id, nfile)
    _check(_load().psod_h5_write_halo_mass(
        outdir.encode(), C.byref(m), len(hm), _p(hm), _p(cp), _p(cv), _p(hid),
        _p(rr)))
    return os.path.join(outdir,
                        f"psod_halo_snap_part_hfprop_{simulation_name}.h5")


def _periodic_d2(x, c):
    d = np.abs(x - c)
    d = np.minimum(d, 1.0 - d)
    return (d * d).sum(axis=1)


def write_greedy(outdir, simulation_name, g, pos, vel, ids, nx,
                 process_id=0):
    """The reference's output of a greedy pass (main file:2264-2410): the
    accepted halos in processing order, each with its particles (those the
    pass marked in `structure`) in order of distance from the centre (ties
    by index; the reference lists them in its sort order, main file:2030,
    2154), the halo mass as a particle count, the centre / nx as
    position_halo (the reference's xctab), the mean velocity of the members
    as velocity_halo and R / nx as rmax_halo (both zero in the reference).
    Returns (part file, hfprop file)."""
    h = g["halos"]
    acc = h["accepted"] == 1
    iamas = h["iamas"][acc]
    if np.any(iamas == 0):
        raise ValueError("an accepted halo has id 0: `structure` marks "
                         "particles of no halo with 0 (the reference's "
                         "convention), so particle ids must not be 0")
    cen = np.stack([h["xc"][acc], h["yc"][acc], h["zc"][acc]], axis=1) / nx
    st = np.asarray(g["structure"])
    order = np.argsort(st, kind="stable")
    stv = st[order]
    pos = np.asarray(pos, np.float32)
    vel = np.asarray(vel, np.float32)
    ids = np.asarray(ids, np.int64)
    members, mass, cvel = [], [], []
    for k, hid in enumerate(iamas):
        a, b = np.searchsorted(stv, [hid, hid + 1])
        mem = order[a:b]
        d2 = _periodic_d2(pos[mem].astype(np.float64), cen[k])
        mem = mem[np.lexsort((mem, d2))]
        members.append(mem)
        mass.append(len(mem))
        cvel.append(vel[mem].astype(np.float64).mean(axis=0) if len(mem)
                    else np.zeros(3))
    allm = (np.concatenate(members) if members else np.zeros(0, np.int64))
    npart_sim = len(pos)
    f1 = h5writehalopart(outdir, simulation_name, mass, iamas, pos[allm],
                         vel[allm], ids[allm], npart_sim, process_id)
    f2 = h5writehalomass(outdir, simulation_name, mass, cen,
                         np.array(cvel).reshape(-1, 3), iamas,
                         h["rayon"][acc] / nx, npart_sim, process_id)
    return f1, f2


def write_fof(outdir, simulation_name, out, nx, npart_simulation,
              process_id=0):
    """hfprop file of the FoF-seeded pass (device outputs of sod_fof /
    run_pass): the accepted halos with their count inside R_Delta[0], centre
    of mass (box units), centre-of-mass velocity, centre id and R_Delta[0]
    / nx.  Returns the file path."""
    o = {k: (v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v))
         for k, v in out.items()}
    acc = (o["status"] & 2) != 0
    return h5writehalomass(outdir, simulation_name, o["count"][acc, 0],
                           o["com_pos"][acc], o["com_vel"][acc],
                           o["centre_id"][acc], o["radius"][acc, 0] / nx,
                           npart_simulation, process_id)
