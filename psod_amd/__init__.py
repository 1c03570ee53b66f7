"""psod_amd -- MI355X-native spherical-overdensity (SOD) pass.

Python host side of the C ABI in include/psod.h.  The module-level functions
``cic``, ``cicm1``, ``tri_cic`` and ``spherover`` keep the names and argument
meaning of the reference subroutines (sod/sod_profile_rdfof_parallel_hdf5.f90)
and work on host (numpy) arrays; ``Context`` exposes the device-resident entry
points on torch tensors that already live in HBM.

Every computation runs in the HIP library psod_amd/libpsod.so; importing this
package without it raises ImportError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BLOCK, E_ARG, E_CAPACITY, E_DEVICE, E_NOMEM, E_UNCLOSED,
                   F_CENTRE_ONLY, F_VELOCITY, H_ACCEPTED, H_CLOSED, H_LARGE,
                   H_UNCLOSED, MAXDELTA, OK, PsodError)

_L = _lib.load()

__all__ = ["Context", "PsodError", "cic", "cicm1", "tri_cic", "spherover",
           "sod_params", "gen_params", "default_context"]


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(C.c_void_p)
    return C.c_void_p(t.data_ptr())


def sod_params(nx, npart, deltas=(200.0,), delta2=0.0, nbins=32,
               rmin_frac=0.01, rmax_frac=2.0, mmin=0.0, nmin_count=10,
               tol_abs=None, small_max=0, velocity=False, centre_only=False):
    """psod_sod_params for the FoF-seeded pass (units: SOD cells).
    ``centre_only`` applies to the plain FP64 path: the centre search
    alone."""
    p = _lib.SodParams()
    p.nx, p.ndelta = nx, len(deltas)
    for i, d in enumerate(deltas):
        p.delta[i] = d
    p.delta2 = delta2
    p.nbins, p.rmin_frac, p.rmax_frac = nbins, rmin_frac, rmax_frac
    p.mmin = mmin
    p.mpart = float(nx) ** 3 / float(npart)
    p.nmin_count = nmin_count
    p.rsearch, p.eps = 1.0, 1.0 / 64.0
    p.tol_abs = 1e-7 * nx if tol_abs is None else tol_abs
    p.small_max = small_max
    p.flags = (F_VELOCITY if velocity else 0) | (
        F_CENTRE_ONLY if centre_only else 0)
    return p


def gen_params(npart, nclump, nlo=20, nhi=20000, alpha=1.9, clo=5.0, chi=10.0,
               trunc=2.0, seed=1, ncentral=0, ccentral=6.0, sig_clump=300.0,
               sig_bulk=300.0, sig_bg=50.0):
    g = _lib.GenParams()
    g.npart, g.nclump = npart, nclump
    g.nlo, g.nhi, g.alpha = nlo, nhi, alpha
    g.clo, g.chi, g.trunc = clo, chi, trunc
    g.seed, g.ncentral, g.ccentral = seed, ncentral, ccentral
    g.sig_clump, g.sig_bulk, g.sig_bg = sig_clump, sig_bulk, sig_bg
    return g


def gen_clumps(g):
    """Clump table of the synthetic catalogue (host part of psod_generate)."""
    nc = g.nclump + (1 if g.ncentral > 0 else 0)
    t = dict(size=np.zeros(nc, np.int64), offset=np.zeros(nc + 1, np.int64),
             centre=np.zeros((nc, 3)), conc=np.zeros(nc), r200=np.zeros(nc),
             bulk=np.zeros((nc, 3)))
    ncl = _L.psod_gen_clumps(C.byref(g), _ptr(t["size"]), _ptr(t["offset"]),
                             _ptr(t["centre"]), _ptr(t["conc"]),
                             _ptr(t["r200"]), _ptr(t["bulk"]))
    t["nclumped"] = int(ncl)
    return t


class Context:
    """One device, one stream (psod_create)."""

    def __init__(self, device: int = 0):
        import torch  # device memory for the resident entry points
        self.torch = torch
        self.device = device
        h = C.c_void_p()
        _lib.check(_L.psod_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _L.psod_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return _L.psod_stream(self.h)

    def synchronize(self):
        _lib.check(_L.psod_synchronize(self.h))

    def _empty(self, shape, dtype):
        return self.torch.empty(shape, dtype=dtype,
                                device=f"cuda:{self.device}")

    def _after_torch(self):
        """Orders the context's stream after the work queued so far on
        torch's current stream.  The library's streams are non-blocking, so
        without this a tensor that torch fills or copies just before a call
        (the zeroed seed keys, a caller's upload) could still be written by
        torch while the library's kernels use it.  The entry points
        synchronise before returning, so their outputs need no such step."""
        torch = self.torch
        ext = getattr(self, "_ext", None)
        if ext is None:
            ext = torch.cuda.ExternalStream(
                self.stream, device=torch.device("cuda", self.device))
            self._ext = ext
        ext.wait_stream(torch.cuda.current_stream(self.device))

    # ---- synthetic catalogue -------------------------------------------
    def generate(self, g, with_vel=True):
        torch = self.torch
        t = gen_clumps(g)
        n = g.npart
        pos = self._empty((n, 3), torch.float32)
        vel = self._empty((n, 3), torch.float32) if with_vel else None
        lab = self._empty((n,), torch.int32)
        ids = self._empty((n,), torch.int64)
        self._after_torch()
        _lib.check(_L.psod_generate(
            self.h, C.byref(g), _ptr(t["offset"]), _ptr(t["centre"]),
            _ptr(t["conc"]), _ptr(t["r200"]), _ptr(t["bulk"]), _ptr(pos),
            _ptr(vel), _ptr(lab), _ptr(ids)))
        return pos, vel, lab, ids, t

    # ---- a5 --------------------------------------------------------------
    def cell_sort(self, pos, nx, labels=None, want_perm=True):
        torch = self.torch
        n = pos.shape[0]
        cs = self._empty((nx ** 3 + 1,), torch.int32)
        perm = self._empty((n,), torch.int32) if want_perm else None
        rec = self._empty((n, 4), torch.float32)
        labs = self._empty((n,), torch.int32) if labels is not None else None
        self._after_torch()
        _lib.check(_L.psod_cell_sort(self.h, n, _ptr(pos), _ptr(labels), nx,
                                     _ptr(cs), _ptr(perm), _ptr(rec),
                                     _ptr(labs)))
        return dict(cell_start=cs, perm=perm, rec=rec, lab_sorted=labs, nx=nx,
                    npart=n)

    # ---- a2 / a3 ---------------------------------------------------------
    def cic_deposit(self, cat, ng, shift=0, out=None):
        grid = out if out is not None else self._empty((ng, ng, ng),
                                                       self.torch.float32)
        self._after_torch()
        _lib.check(_L.psod_cic_deposit(self.h, cat["npart"], cat["nx"],
                                       _ptr(cat["cell_start"]),
                                       _ptr(cat["rec"]), ng, shift,
                                       _ptr(grid)))
        return grid

    def cic_interp(self, cat, grid, shift=0, sorted_out=False, orig_out=True,
                   nhalo=0, seed_key=None):
        torch = self.torch
        n = cat["npart"]
        rs = self._empty((n,), torch.float32) if sorted_out else None
        ro = self._empty((n,), torch.float32) if orig_out else None
        self._after_torch()
        _lib.check(_L.psod_cic_interp(
            self.h, n, cat["nx"], _ptr(cat["cell_start"]), _ptr(cat["rec"]),
            grid.shape[0], shift, _ptr(grid), _ptr(rs), _ptr(ro),
            _ptr(cat["lab_sorted"]), nhalo, _ptr(seed_key)))
        return ro, rs

    def seeds(self, cat, grid, nhalo, shift=0):
        """FoF seeds: densest member of each label (ties: smallest index)."""
        key = self._empty((nhalo,), self.torch.int64)  # cleared by the call
        self.cic_interp(cat, grid, shift=shift, orig_out=False, nhalo=nhalo,
                        seed_key=key)
        seed = self._empty((nhalo,), self.torch.int32)
        self._after_torch()
        _lib.check(_L.psod_seed_index(self.h, nhalo, _ptr(key), _ptr(seed)))
        return seed

    def label_counts(self, labels, nhalo):
        out = self._empty((nhalo,), self.torch.int32)
        self._after_torch()
        _lib.check(_L.psod_label_counts(self.h, labels.shape[0], _ptr(labels),
                                        nhalo, _ptr(out)))
        return out

    # ---- a4 --------------------------------------------------------------
    def density_order(self, rho, threshold=100.0):
        n = rho.shape[0]
        isort = self._empty((n,), self.torch.int32)
        nn = C.c_int64(0)
        self._after_torch()
        _lib.check(_L.psod_density_order(self.h, n, _ptr(rho), threshold,
                                         _ptr(isort), C.byref(nn)))
        return isort, nn.value

    # ---- a6-a12 ----------------------------------------------------------
    def alloc_halo_out(self, nhalo, nbins):
        torch = self.torch
        o = dict(
            centre=self._empty((nhalo,), torch.int32),
            centre_id=self._empty((nhalo,), torch.int64),
            status=self._empty((nhalo,), torch.int32),
            w=self._empty((nhalo,), torch.int32),
            count=self._empty((nhalo, MAXDELTA), torch.int32),
            radius=self._empty((nhalo, MAXDELTA), torch.float64),
            overdens=self._empty((nhalo, MAXDELTA), torch.float64),
            rint=self._empty((nhalo, MAXDELTA), torch.float64),
            mint=self._empty((nhalo, MAXDELTA), torch.float64),
            count2=self._empty((nhalo,), torch.int32),
            overdens2=self._empty((nhalo,), torch.float64),
            profile=self._empty((nhalo, nbins), torch.int32),
            com_pos=self._empty((nhalo, 3), torch.float64),
            com_vel=self._empty((nhalo, 3), torch.float64),
            ngather=self._empty((nhalo,), torch.int32),
            nboundary=self._empty((nhalo,), torch.int32),
        )
        return o

    def sod_fof(self, p, cat, pos, seeds, nmem=None, vel=None, ids=None,
                plain=False, out=None):
        nh = seeds.shape[0]
        o = out if out is not None else self.alloc_halo_out(nh, p.nbins)
        ho = _lib.HaloOut(**{k: _ptr(o[k]).value for k in o})
        c = _lib.Catalogue(cat["npart"], cat["nx"], _ptr(pos).value,
                           _ptr(vel).value if vel is not None else None,
                           _ptr(ids).value if ids is not None else None,
                           _ptr(cat["cell_start"]).value,
                           _ptr(cat["rec"]).value)
        f = _L.psod_sod_fof_plain if plain else _L.psod_sod_fof
        self._after_torch()
        _lib.check(f(self.h, C.byref(p), C.byref(c), nh, _ptr(seeds),
                     _ptr(nmem), C.byref(ho)))
        return o

    def run_pass(self, p, pos, labels, nhalo, ng=None, vel=None, ids=None,
                 grid=None, out=None):
        """psod_pass: cell sort, CIC, seeds, member counts and the batched
        SOD of every label in one call (device tensors)."""
        nx = p.nx
        ng = nx if ng is None else ng
        o = out if out is not None else self.alloc_halo_out(nhalo, p.nbins)
        ho = _lib.HaloOut(**{k: _ptr(o[k]).value for k in o})
        self._after_torch()
        _lib.check(_L.psod_pass(self.h, C.byref(p), pos.shape[0], nx, ng,
                                _ptr(pos), _ptr(vel), _ptr(ids), _ptr(labels),
                                nhalo, _ptr(grid), C.byref(ho)))
        return o

    # ---- tunables ----------------------------------------------------------
    def set_tunables(self, small_max=0, bound_min=0, dense_tile=0,
                     small_wg_per_cu=0, giant_min=0):
        """psod_set_tunables: 0 leaves a field to the tuning file or the
        built-in default."""
        t = _lib.Tunables(small_max, bound_min, dense_tile, small_wg_per_cu,
                          giant_min)
        _lib.check(_L.psod_set_tunables(self.h, C.byref(t)))

    def tunables(self, nx):
        """The tunables in force for grids of nx cells per side."""
        t = _lib.Tunables()
        _lib.check(_L.psod_get_tunables(self.h, nx, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.Tunables._fields_}

    def autotune(self, p, pos, labels, nhalo, vel=None, ids=None, reps=2,
                 cache_file=None):
        """psod_autotune on a device catalogue: returns (chosen tunables,
        best deposit ms, best SOD ms)."""
        t = _lib.Tunables()
        ms = (C.c_double * 2)()
        self._after_torch()
        _lib.check(_L.psod_autotune(
            self.h, C.byref(p), pos.shape[0], p.nx, _ptr(pos), _ptr(vel),
            _ptr(ids), _ptr(labels), nhalo, reps,
            cache_file.encode() if cache_file else None, C.byref(t), ms))
        return ({f: getattr(t, f) for f, _ in _lib.Tunables._fields_},
                ms[0], ms[1])

    def load_tunables(self, cache_file):
        n = C.c_int32(0)
        _lib.check(_L.psod_load_tunables(self.h, cache_file.encode(),
                                         C.byref(n)))
        return n.value

    def run_pass_graph(self, p, pos, labels, nhalo, ng=None, vel=None,
                       ids=None, grid=None, out=None):
        """psod_pass_graph: the pass of run_pass, captured as one HIP graph
        on the first call with these tensors and replayed afterwards.
        Returns (outputs, replayed)."""
        nx = p.nx
        ng = nx if ng is None else ng
        o = out if out is not None else self.alloc_halo_out(nhalo, p.nbins)
        ho = _lib.HaloOut(**{k: _ptr(o[k]).value for k in o})
        rp = C.c_int32(0)
        self._after_torch()
        _lib.check(_L.psod_pass_graph(self.h, C.byref(p), pos.shape[0], nx,
                                      ng, _ptr(pos), _ptr(vel), _ptr(ids),
                                      _ptr(labels), nhalo, _ptr(grid),
                                      C.byref(ho), C.byref(rp)))
        return o, bool(rp.value)

    def sod_stats(self):
        """Execution-path counts of the last sod_fof call: halos grown by
        one wave, by one workgroup, re-run with larger collection slices,
        the keys per slice of that re-run, and the centre search of the
        large halos: its tasks (64 candidates of one cell) in all, units
        evaluated exactly (pilot tasks and groups of 8 candidates), units
        skipped by their bound, groups abandoned part-way, pair tests and
        the search's time; giant halos grown by the staged launches and
        their cycles; workgroup halos whose growth cube was sorted whole in
        LDS."""
        a = (C.c_int64 * 13)()
        _lib.check(_L.psod_sod_stats(self.h, a, 13))
        return dict(wave=a[0], workgroup=a[1], retried=a[2], retry_slice=a[3],
                    centre_tasks=a[4], centre_skipped=a[5], pair_tests=a[6],
                    centre_tasks_all=a[7], centre_ms=a[8] * 1e-6,
                    centre_aborted=a[9], giants=a[10], giant_cycles=a[11],
                    light=a[12])

    # ---- host drop-ins ---------------------------------------------------
    def host_cic(self, pos, ng, mass=None, xxmin=0.0):
        """`cic` in FP64: grid [iz, iy, ix] of rho/rho_bar."""
        pos = np.ascontiguousarray(pos, np.float32)
        m = None if mass is None else np.ascontiguousarray(mass, np.float32)
        grid = np.zeros((ng, ng, ng), np.float64)
        _lib.check(_L.psod_host_cic(self.h, len(pos), _ptr(pos), _ptr(m),
                                    float(xxmin), ng, _ptr(grid)))
        return grid

    def host_cicm1(self, pos, grid, xxmin=0.0):
        """`cicm1` in FP64: density at each particle."""
        pos = np.ascontiguousarray(pos, np.float32)
        grid = np.ascontiguousarray(grid, np.float64)
        rho = np.zeros(len(pos), np.float64)
        _lib.check(_L.psod_host_cicm1(self.h, len(pos), _ptr(pos),
                                      float(xxmin), grid.shape[0], _ptr(grid),
                                      _ptr(rho)))
        return rho

    def host_tri_cic(self, rho, threshold=100.0):
        """`tri_cic`: (isort 1-based by ascending 1/rho, npart_new)."""
        rho = np.ascontiguousarray(rho, np.float64)
        isort = np.zeros(len(rho), np.int32)
        nn = C.c_int64(0)
        _lib.check(_L.psod_host_tri_cic(self.h, len(rho), _ptr(rho),
                                        float(threshold), _ptr(isort),
                                        C.byref(nn)))
        return isort, nn.value

    def density_order_f64(self, rho, threshold=100.0):
        """psod_density_order_f64 on a device tensor of FP64 densities."""
        n = rho.shape[0]
        isort = self._empty((n,), self.torch.int32)
        nn = C.c_int64(0)
        self._after_torch()
        _lib.check(_L.psod_density_order_f64(self.h, n, _ptr(rho), threshold,
                                             _ptr(isort), C.byref(nn)))
        return isort, nn.value

    def host_spherover(self, pos, ids, isort, npart_new, nx, ncut=4,
                       seuil=200.0, seuil2=800.0, mmin=100.0, nsel=4_000_000,
                       nbins=60, rmin_frac=0.01, rmax_frac=1.0, maxrec=None):
        """Greedy `spherover` (main file:1396-2201) on the device."""
        pos = np.ascontiguousarray(pos, np.float32)
        ids = np.ascontiguousarray(ids, np.int64)
        isort = np.ascontiguousarray(isort, np.int32)
        n = len(pos)
        maxrec = maxrec or max(int(npart_new), 1)
        gp = _lib.GreedyParams(nx, ncut, seuil, seuil2, mmin, nsel, nbins,
                               rmin_frac, rmax_frac)
        ri = np.zeros((maxrec, 6), np.int64)
        rf = np.zeros((maxrec, 8), np.float64)
        rb = np.zeros((maxrec, nbins), np.float64)
        mb = np.zeros((maxrec, nbins), np.int32)
        alive = np.zeros(n, np.uint8)
        structure = np.zeros(n, np.int64)
        nrec = C.c_int64(0)
        rc = _L.psod_host_spherover(self.h, n, _ptr(pos), _ptr(ids),
                                    _ptr(isort), int(npart_new),
                                    C.byref(gp), maxrec, C.byref(nrec),
                                    _ptr(ri), _ptr(rf), _ptr(rb), _ptr(mb),
                                    _ptr(alive), _ptr(structure))
        if rc not in (OK, E_CAPACITY, E_UNCLOSED):
            _lib.check(rc)
        k = nrec.value
        names_i = ["ipart", "seed", "centre", "iamas", "masse", "masse2"]
        names_f = ["amass", "xc", "yc", "zc", "rayon", "overdens",
                   "overdens2", "accepted"]
        halos = {nm: ri[:k, j] for j, nm in enumerate(names_i)}
        halos.update({nm: rf[:k, j] for j, nm in enumerate(names_f)})
        return dict(status=rc, halos=halos, rad_bins=rb[:k],
                    mass_bins=mb[:k], alive=alive, structure=structure)


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


# ---- reference-named drop-ins (host arrays) --------------------------------
def cic(pos, ng, mass=None, xxmin=0.0):
    """`cic` (main file:1077): FP64 grid [iz, iy, ix] of rho/rho_bar."""
    return default_context().host_cic(pos, ng, mass, xxmin)


def cicm1(pos, grid, xxmin=0.0):
    """`cicm1` (main file:1268): FP64 density at each particle."""
    return default_context().host_cicm1(pos, grid, xxmin)


def tri_cic(rho, threshold=100.0):
    """`tri_cic` (main file:998): (isort 1-based by ascending 1/rho, ties
    by index, npart_new)."""
    return default_context().host_tri_cic(rho, threshold)


def spherover(pos, ids, isort, npart_new, nx, **kw):
    """`spherover` (main file:1396): greedy SOD with exclusion."""
    return default_context().host_spherover(pos, ids, isort, npart_new, nx,
                                            **kw)
