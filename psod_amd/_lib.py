"""ctypes declarations of include/psod.h (psod_amd/libpsod.so)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PSOD_LIBRARY selects another build of the same library (A/B timing of
# builds); by default the in-tree build is loaded
LIBPATH = os.environ.get("PSOD_LIBRARY") or os.path.join(HERE, "libpsod.so")

OK, E_ARG, E_DEVICE, E_CAPACITY, E_UNCLOSED, E_NOMEM = 0, 1, 2, 3, 4, 5
MAXDELTA = 4
BLOCK = 8
H_CLOSED, H_ACCEPTED, H_LARGE, H_UNCLOSED = 1, 2, 4, 8
F_VELOCITY = 1
F_CENTRE_ONLY = 2

EXPORTS = [
    "psod_create", "psod_destroy", "psod_last_error", "psod_synchronize",
    "psod_stream", "psod_scratch_bytes", "psod_sod_stats", "psod_pass", "psod_cell_sort", "psod_cic_deposit",
    "psod_cic_interp", "psod_seed_index", "psod_density_order",
    "psod_sod_fof", "psod_sod_fof_plain", "psod_label_counts",
    "psod_gen_clumps", "psod_generate", "psod_host_cic", "psod_host_cicm1",
    "psod_host_tri_cic", "psod_host_spherover", "psod_density_order_f64",
    "psod_cic_deposit_noalloc", "psod_set_tunables", "psod_get_tunables",
    "psod_autotune", "psod_load_tunables", "psod_pass_graph",
]


class PsodError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"psod error {code}: {msg}")
        self.code = code


class SodParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ndelta", C.c_int32),
        ("delta", C.c_double * MAXDELTA), ("delta2", C.c_double),
        ("nbins", C.c_int32), ("rmin_frac", C.c_double),
        ("rmax_frac", C.c_double), ("mmin", C.c_double),
        ("mpart", C.c_double), ("nmin_count", C.c_int32),
        ("rsearch", C.c_double), ("eps", C.c_double),
        ("tol_abs", C.c_double), ("small_max", C.c_int32),
        ("flags", C.c_int32),
    ]


class HaloOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "centre", "centre_id", "status", "w", "count", "radius", "overdens",
        "rint", "mint", "count2", "overdens2", "profile", "com_pos",
        "com_vel", "ngather", "nboundary")]


class Catalogue(C.Structure):
    _fields_ = [
        ("npart", C.c_int64), ("nx", C.c_int32), ("pos", C.c_void_p),
        ("vel", C.c_void_p), ("ids", C.c_void_p), ("cell_start", C.c_void_p),
        ("rec", C.c_void_p),
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


class Tunables(C.Structure):
    _fields_ = [("small_max", C.c_int32), ("bound_min", C.c_int32),
                ("dense_tile", C.c_int32), ("small_wg_per_cu", C.c_int32),
                ("giant_min", C.c_int32)]


class GreedyParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ncut", C.c_int32), ("seuil", C.c_double),
        ("seuil2", C.c_double), ("mmin", C.c_double), ("nsel", C.c_int64),
        ("nbins", C.c_int32), ("rmin_frac", C.c_double),
        ("rmax_frac", C.c_double),
    ]


_lib = None


def load():
    """Load libpsod.so.  There is no fallback: without the HIP library the
    package cannot compute anything."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(
            f"{LIBPATH} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIBPATH)
    v, i32, i64, d, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint64
    P = C.POINTER
    sig = {
        "psod_create": (C.c_int, [C.c_int, P(v)]),
        "psod_destroy": (C.c_int, [v]),
        "psod_last_error": (C.c_char_p, []),
        "psod_synchronize": (C.c_int, [v]),
        "psod_stream": (v, [v]),
        "psod_scratch_bytes": (i64, [v]),
        "psod_sod_stats": (C.c_int, [v, P(i64), i32]),
        "psod_cell_sort": (C.c_int, [v, i64, v, v, i32, v, v, v, v]),
        "psod_cic_deposit": (C.c_int, [v, i64, i32, v, v, i32, i32, v]),
        "psod_cic_deposit_noalloc": (C.c_int, [v, i64, i32, v, v, i32, i32,
                                               v]),
        "psod_cic_interp": (C.c_int, [v, i64, i32, v, v, i32, i32, v, v, v,
                                      v, i64, v]),
        "psod_seed_index": (C.c_int, [v, i64, v, v]),
        "psod_density_order": (C.c_int, [v, i64, v, d, v, P(i64)]),
        "psod_sod_fof": (C.c_int, [v, P(SodParams), P(Catalogue), i64, v, v,
                                   P(HaloOut)]),
        "psod_sod_fof_plain": (C.c_int, [v, P(SodParams), P(Catalogue), i64,
                                         v, v, P(HaloOut)]),
        "psod_label_counts": (C.c_int, [v, i64, v, i64, v]),
        "psod_pass": (C.c_int, [v, P(SodParams), i64, i32, i32, v, v, v, v,
                                i64, v, P(HaloOut)]),
        "psod_pass_graph": (C.c_int, [v, P(SodParams), i64, i32, i32, v, v,
                                      v, v, i64, v, P(HaloOut), P(i32)]),
        "psod_gen_clumps": (i64, [P(GenParams), v, v, v, v, v, v]),
        "psod_generate": (C.c_int, [v, P(GenParams), v, v, v, v, v, v, v, v,
                                    v]),
        "psod_host_cic": (C.c_int, [v, i64, v, v, d, i32, v]),
        "psod_host_cicm1": (C.c_int, [v, i64, v, d, i32, v, v]),
        "psod_host_tri_cic": (C.c_int, [v, i64, v, d, v, P(i64)]),
        "psod_density_order_f64": (C.c_int, [v, i64, v, d, v, P(i64)]),
        "psod_set_tunables": (C.c_int, [v, P(Tunables)]),
        "psod_get_tunables": (C.c_int, [v, i32, P(Tunables)]),
        "psod_autotune": (C.c_int, [v, P(SodParams), i64, i32, v, v, v, v,
                                    i64, i32, C.c_char_p, P(Tunables),
                                    P(d)]),
        "psod_load_tunables": (C.c_int, [v, C.c_char_p, P(i32)]),
        "psod_host_spherover": (C.c_int, [v, i64, v, v, v, i64,
                                          P(GreedyParams), i64, P(i64), v, v,
                                          v, v, v, v]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        msg = load().psod_last_error().decode(errors="replace")
        raise PsodError(rc, msg)
    return rc
