"""ctypes wrapper of the FP64 oracle (oracle/libpsod_oracle.so).

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP
implementation in ``psod_amd``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import it.  The product path
never does.

Each function names the reference routine it restates; see the C source for
file:line citations (``main file`` = sod/sod_profile_rdfof_parallel_hdf5.f90).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libpsod_oracle.so")
_lib = None

OK, E_ARG, E_CAPACITY, E_UNCLOSED = 0, 1, 3, 4
MAXDELTA = 4


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    cmd = ["make", "-C", _HERE]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class GreedyParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ncut", C.c_int32), ("ncut2", C.c_int32),
        ("ncut3", C.c_int32), ("seuil", C.c_double), ("seuil2", C.c_double),
        ("mmin", C.c_double), ("nsel", C.c_int64), ("xxmin", C.c_double),
        ("local_nx", C.c_int32), ("zmin", C.c_double * 3),
        ("zmax", C.c_double * 3), ("nbins", C.c_int32),
        ("rmin_frac", C.c_double), ("rmax_frac", C.c_double),
        ("ref_order", C.c_int32),
    ]


class GreedyHalo(C.Structure):
    _fields_ = [
        ("ipart", C.c_int64), ("seed", C.c_int64), ("centre", C.c_int64),
        ("iamas", C.c_int64), ("masse", C.c_int64), ("masse2", C.c_int64),
        ("npart_sel", C.c_int64), ("amass", C.c_double), ("xc", C.c_double),
        ("yc", C.c_double), ("zc", C.c_double), ("rayon", C.c_double),
        ("overdens", C.c_double), ("overdens2", C.c_double),
        ("accepted", C.c_int32),
    ]


class FofParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ndelta", C.c_int32),
        ("delta", C.c_double * MAXDELTA), ("delta2", C.c_double),
        ("nbins", C.c_int32), ("rmin_frac", C.c_double),
        ("rmax_frac", C.c_double), ("mmin", C.c_double),
        ("mpart", C.c_double), ("nmin_count", C.c_int32),
        ("rsearch", C.c_double), ("eps", C.c_double), ("tol", C.c_double),
        ("tol_abs", C.c_double),
    ]


class FofHalo(C.Structure):
    _fields_ = [
        ("seed", C.c_int64), ("centre", C.c_int64), ("centre_id", C.c_int64),
        ("status", C.c_int32), ("robust_centre", C.c_int32), ("w", C.c_int32),
        ("accepted", C.c_int32), ("ncand", C.c_int64), ("nneigh", C.c_int64),
        ("ngather", C.c_int64), ("s_best", C.c_double),
        ("s_second", C.c_double), ("count", C.c_int64 * MAXDELTA),
        ("radius", C.c_double * MAXDELTA), ("overdens", C.c_double * MAXDELTA),
        ("rint", C.c_double * MAXDELTA), ("mint", C.c_double * MAXDELTA),
        ("margin", C.c_double * MAXDELTA), ("rgap", C.c_double * MAXDELTA),
        ("count2", C.c_int64),
        ("overdens2", C.c_double), ("margin2", C.c_double),
        ("profile_margin", C.c_double), ("nboundary", C.c_int64),
        ("com_pos", C.c_double * 3), ("com_vel", C.c_double * 3),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("npart", C.c_int64), ("nclump", C.c_int64), ("nlo", C.c_double),
        ("nhi", C.c_double), ("alpha", C.c_double), ("clo", C.c_double),
        ("chi", C.c_double), ("trunc", C.c_double), ("seed", C.c_uint64),
        ("ncentral", C.c_int64), ("ccentral", C.c_double),
        ("sig_clump", C.c_double), ("sig_bulk", C.c_double),
        ("sig_bg", C.c_double),
    ]


def _declare(L):
    v, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.or_cic.argtypes = [i64, v, v, d, i32, v]
    L.or_cicm1.argtypes = [i64, v, d, i32, v, v]
    L.or_quick_sort.argtypes = [i64, v, v]
    L.or_tri_cic.argtypes = [i64, v, v, v, d]
    L.or_tri_cic.restype = i64
    L.or_cell_key.argtypes = [i32, i32, i32, i32, i32]
    L.or_cell_key.restype = i64
    L.or_cell_sort.argtypes = [i64, v, i32, i32, v, v]
    L.or_cell_sort.restype = C.c_int
    L.or_sod_greedy.argtypes = [i64, v, v, v, v, i64, C.POINTER(GreedyParams),
                                i64, C.POINTER(i64), v, v, v, v, v]
    L.or_sod_greedy.restype = C.c_int
    L.or_fof_seeds.argtypes = [i64, v, v, i64, v, v]
    L.or_sod_fof.argtypes = [i64, v, v, v, v, v, i32, i64, v, v,
                             C.POINTER(FofParams), v, v]
    L.or_sod_fof.restype = C.c_int
    L.or_sod_fof_at.argtypes = [i64, v, v, v, v, v, i32, i64, v, v, v,
                                C.POINTER(FofParams), v, v]
    L.or_sod_fof_at.restype = C.c_int
    L.or_computecom.argtypes = [i64, v, v, v, v]
    L.or_delta_crit_to_mean.argtypes = [d, d]
    L.or_delta_crit_to_mean.restype = d
    L.or_bryan_norman_mean.argtypes = [d]
    L.or_bryan_norman_mean.restype = d
    L.or_gen_clumps.argtypes = [C.POINTER(GenParams), v, v, v, v, v, v]
    L.or_gen_clumps.restype = i64
    L.or_gen_particles.argtypes = [C.POINTER(GenParams), v, v, v, v, v, i64,
                                   i64, v, v, v, v]


def _pos(pos):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    assert pos.ndim == 2 and pos.shape[1] == 3
    return pos


# --- a2 / a3 ----------------------------------------------------------------
def cic(pos, ng, mass=None, xxmin=0.0):
    """`cic` (main file:1077-1265): FP64 grid [iz, iy, ix] of rho/rho_bar."""
    pos = _pos(pos)
    if mass is not None:
        mass = np.ascontiguousarray(mass, dtype=np.float32)
    grid = np.zeros((ng, ng, ng), dtype=np.float64)
    lib().or_cic(len(pos), _p(pos), _p(mass), float(xxmin), int(ng), _p(grid))
    return grid


def cicm1(pos, grid, xxmin=0.0):
    """`cicm1` (main file:1268-1390): FP64 density at each particle."""
    pos = _pos(pos)
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    rho = np.zeros(len(pos), dtype=np.float64)
    lib().or_cicm1(len(pos), _p(pos), float(xxmin), int(grid.shape[0]),
                   _p(grid), _p(rho))
    return rho


# --- a11 / a4 / a5 ----------------------------------------------------------
def quick_sort(values):
    """`quick_sort` (main file:2477-2571): (sorted copy, 1-based order)."""
    lst = np.array(values, dtype=np.float64, copy=True)
    order = np.zeros(len(lst), dtype=np.int32)
    lib().or_quick_sort(len(lst), _p(lst), _p(order))
    return lst, order


def tri_cic(rho, threshold=100.0):
    """`tri_cic` (main file:998-1037): (pdens, isort 1-based, npart_new)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    pdens = np.zeros_like(rho)
    isort = np.zeros(len(rho), dtype=np.int32)
    nnew = lib().or_tri_cic(len(rho), _p(rho), _p(pdens), _p(isort),
                            float(threshold))
    return pdens, isort, int(nnew)


def cell_key(nx, bs, ix, iy, iz):
    return int(lib().or_cell_key(nx, bs, ix, iy, iz))


def cell_sort(pos, nx, bs):
    """Stable counting sort by cell (row a5): (cell_start int64, perm int32)."""
    pos = _pos(pos)
    cs = np.zeros(nx ** 3 + 1, dtype=np.int64)
    perm = np.zeros(len(pos), dtype=np.int32)
    rc = lib().or_cell_sort(len(pos), _p(pos), nx, bs, _p(cs), _p(perm))
    if rc != OK:
        raise ValueError(f"or_cell_sort failed: {rc}")
    return cs, perm


# --- a6-a10 greedy ----------------------------------------------------------
def greedy_params(nx, ncut=4, seuil=200.0, seuil2=800.0, mmin=100.0,
                  nsel=4_000_000, xxmin=-0.5, nbins=60, rmin_frac=0.01,
                  rmax_frac=1.0, ref_order=1, local_nx=None):
    p = GreedyParams()
    p.nx, p.ncut, p.ncut2, p.ncut3 = nx, ncut, 2, 1
    p.seuil, p.seuil2, p.mmin, p.nsel = seuil, seuil2, mmin, nsel
    p.xxmin = xxmin
    p.local_nx = 2 * nx if local_nx is None else local_nx
    for d in range(3):
        p.zmin[d], p.zmax[d] = 0.0, float(nx)
    p.nbins, p.rmin_frac, p.rmax_frac = nbins, rmin_frac, rmax_frac
    p.ref_order = ref_order
    return p


def _struct_to_dict(arr, n, cls):
    out = {}
    for name, ctype in cls._fields_:
        vals = [getattr(arr[i], name) for i in range(n)]
        if hasattr(ctype, "_length_"):
            out[name] = np.array([list(v) for v in vals]).reshape(
                n, ctype._length_)
        else:
            out[name] = np.array(vals)
    return out


def sod_greedy(pos, ids, isort, npart_new, params, mass=None, maxrec=None):
    """Greedy `spherover` (main file:1396-2201) on unscaled positions."""
    pos = _pos(pos)
    n = len(pos)
    ids = np.ascontiguousarray(ids, dtype=np.int64)
    isort = np.ascontiguousarray(isort, dtype=np.int32)
    if mass is not None:
        mass = np.ascontiguousarray(mass, dtype=np.float32)
    maxrec = maxrec or max(int(npart_new), 1)
    rec = (GreedyHalo * maxrec)()
    rb = np.zeros((maxrec, params.nbins), dtype=np.float64)
    mb = np.zeros((maxrec, params.nbins), dtype=np.int64)
    alive = np.zeros(n, dtype=np.uint8)
    structure = np.zeros(n, dtype=np.int64)
    nrec = C.c_int64(0)
    rc = lib().or_sod_greedy(n, _p(pos), _p(ids), _p(mass), _p(isort),
                             int(npart_new), C.byref(params), maxrec,
                             C.byref(nrec), rec, _p(rb), _p(mb), _p(alive),
                             _p(structure))
    k = nrec.value
    return dict(status=rc, halos=_struct_to_dict(rec, k, GreedyHalo),
                rad_bins=rb[:k], mass_bins=mb[:k], alive=alive,
                structure=structure)


# --- FoF-seeded independent mode ---------------------------------------------
def fof_params(nx, npart, deltas=(200.0,), delta2=0.0, nbins=32,
               rmin_frac=0.01, rmax_frac=2.0, mmin=0.0, nmin_count=10,
               tol=1e-5, tol_abs=None):
    p = FofParams()
    p.nx, p.ndelta = nx, len(deltas)
    for i, dv in enumerate(deltas):
        p.delta[i] = dv
    p.delta2 = delta2
    p.nbins, p.rmin_frac, p.rmax_frac = nbins, rmin_frac, rmax_frac
    p.mmin = mmin
    p.mpart = float(nx) ** 3 / float(npart)
    p.nmin_count = nmin_count
    p.rsearch, p.eps = 1.0, 1.0 / 64.0
    p.tol = tol
    p.tol_abs = 1e-7 * nx if tol_abs is None else tol_abs
    return p


def fof_seeds(labels, rho, nhalo):
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    seeds = np.zeros(nhalo, dtype=np.int64)
    nmem = np.zeros(nhalo, dtype=np.int64)
    lib().or_fof_seeds(len(labels), _p(labels), _p(rho), nhalo, _p(seeds),
                       _p(nmem))
    return seeds, nmem


def sod_fof(pos, cell_start, perm, bs, seeds, nmem, params, vel=None,
            ids=None, centres=None):
    """FoF-seeded SOD: centre search, growth, profile and centre of mass of
    each halo.  With ``centres`` (one particle per halo) the centre search
    is skipped and the growth runs around the given centres."""
    pos = _pos(pos)
    seeds = np.ascontiguousarray(seeds, dtype=np.int64)
    nmem = np.ascontiguousarray(nmem, dtype=np.int64)
    if vel is not None:
        vel = np.ascontiguousarray(vel, dtype=np.float32)
    if ids is not None:
        ids = np.ascontiguousarray(ids, dtype=np.int64)
    nh = len(seeds)
    out = (FofHalo * max(nh, 1))()
    prof = np.zeros((max(nh, 1), params.nbins), dtype=np.int64)
    cs = np.ascontiguousarray(cell_start, np.int64)
    pm = np.ascontiguousarray(perm, np.int32)
    if centres is None:
        rc = lib().or_sod_fof(len(pos), _p(pos), _p(vel), _p(ids), _p(cs),
                              _p(pm), bs, nh, _p(seeds), _p(nmem),
                              C.byref(params), out, _p(prof))
    else:
        cen = np.ascontiguousarray(centres, dtype=np.int64)
        assert len(cen) == nh
        rc = lib().or_sod_fof_at(len(pos), _p(pos), _p(vel), _p(ids), _p(cs),
                                 _p(pm), bs, nh, _p(seeds), _p(nmem),
                                 _p(cen), C.byref(params), out, _p(prof))
    if rc != OK:
        raise ValueError(f"or_sod_fof failed: {rc}")
    res = _struct_to_dict(out, nh, FofHalo)
    res["profile"] = prof[:nh]
    return res


def computecom(pos, vel=None):
    pos = _pos(pos)
    if vel is not None:
        vel = np.ascontiguousarray(vel, dtype=np.float32)
    cp = np.zeros(3)
    cv = np.zeros(3)
    lib().or_computecom(len(pos), _p(pos), _p(vel), _p(cp), _p(cv))
    return cp, cv


def delta_crit_to_mean(dc, omega_m):
    return lib().or_delta_crit_to_mean(dc, omega_m)


def bryan_norman_mean(omega_m):
    return lib().or_bryan_norman_mean(omega_m)


# --- synthetic catalogue -----------------------------------------------------
def gen_params(npart, nclump, nlo=20, nhi=20000, alpha=1.9, clo=5.0, chi=10.0,
               trunc=2.0, seed=1, ncentral=0, ccentral=6.0, sig_clump=300.0,
               sig_bulk=300.0, sig_bg=50.0):
    g = GenParams()
    g.npart, g.nclump = npart, nclump
    g.nlo, g.nhi, g.alpha = nlo, nhi, alpha
    g.clo, g.chi, g.trunc = clo, chi, trunc
    g.seed, g.ncentral, g.ccentral = seed, ncentral, ccentral
    g.sig_clump, g.sig_bulk, g.sig_bg = sig_clump, sig_bulk, sig_bg
    return g


def gen_clumps(g):
    nc = g.nclump + (1 if g.ncentral > 0 else 0)
    size = np.zeros(nc, np.int64)
    off = np.zeros(nc + 1, np.int64)
    cen = np.zeros((nc, 3))
    conc = np.zeros(nc)
    r200 = np.zeros(nc)
    bulk = np.zeros((nc, 3))
    ncl = lib().or_gen_clumps(C.byref(g), _p(size), _p(off), _p(cen),
                              _p(conc), _p(r200), _p(bulk))
    return dict(size=size, offset=off, centre=cen, conc=conc, r200=r200,
                bulk=bulk, nclumped=int(ncl))


def generate(g, with_vel=True):
    """Whole catalogue on the CPU: (pos, vel, labels, ids, clump table)."""
    t = gen_clumps(g)
    if t["nclumped"] > g.npart:
        raise ValueError("clumps hold more particles than npart")
    n = g.npart
    pos = np.zeros((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32) if with_vel else None
    lab = np.zeros(n, np.int32)
    ids = np.zeros(n, np.int64)
    lib().or_gen_particles(C.byref(g), _p(t["offset"]), _p(t["centre"]),
                           _p(t["conc"]), _p(t["r200"]), _p(t["bulk"]), 0, n,
                           _p(pos), _p(vel), _p(lab), _p(ids))
    return pos, vel, lab, ids, t
