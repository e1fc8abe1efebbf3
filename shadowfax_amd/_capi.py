"""ctypes binding of libshadowfax_amd.so (the C ABI in include/shadowfax_amd.h).

The library is loaded from this package directory.  There is no fallback: if
the HIP library is missing, importing the entry points raises.
"""
import ctypes as ct
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libshadowfax_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "shadowfax_amd.h")

SFX_EXACT, SFX_HLLC = 0, 1
SFX_PREC_FP64, SFX_PREC_MIXED = 0, 1
SFX_OK = 0
SFX_ERR_ARG = -1
SFX_STATIC_SCRATCH_BYTES = 64 << 20
ST_SOLVER, ST_NANFLUX, ST_ITERCAP, ST_UPDATEQ = 1, 2, 4, 8

vp = ct.c_void_p
i64 = ct.c_int64
dbl = ct.c_double
_dp = ct.POINTER(ct.c_double)
_u8p = ct.POINTER(ct.c_uint8)


class Stats(ct.Structure):
    _fields_ = [("neval", ct.c_uint64), ("nbrent", ct.c_uint64),
                ("nvacuum", ct.c_uint64), ("niter", ct.c_uint64),
                ("nfail", ct.c_uint64)]


class MeshDesc(ct.Structure):
    """sfx_mesh (include/shadowfax_amd.h); device pointers."""
    _fields_ = [("ndim", ct.c_int), ("C", i64), ("F", i64),
                ("W", vp), ("vel", vp), ("mass", vp), ("total_area", vp), ("dt", vp),
                ("volume", vp), ("active", vp),
                ("left", vp), ("right", vp), ("ghost", vp), ("twin", vp),
                ("area", vp), ("dL", vp), ("dR", vp), ("rRL", vp), ("mc", vp), ("rcent", vp),
                ("cf_off", vp), ("cf_entry", vp), ("egeo", vp), ("gather_off", vp),
                ("gather_entry", vp), ("egeo_group", vp), ("egeo_rows", i64),
                ("egeo_slot", vp), ("closure", vp), ("frec", vp)]


_lib = None


class SfxError(RuntimeError):
    pass


def declared_symbols():
    """Function names declared in include/shadowfax_amd.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sfx_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise SfxError("libshadowfax_amd.so not built (run __graft_entry__.build() "
                       "or python -m shadowfax_amd.build); there is no CPU fallback")
    L = ct.CDLL(LIBPATH)
    L.sfx_last_error.restype = ct.c_char_p
    L.sfx_version.restype = ct.c_char_p
    L.sfx_solver_create.argtypes = [ct.c_int, ct.c_int, dbl, dbl, dbl, ct.c_int,
                                    ct.POINTER(vp)]
    L.sfx_solver_create_by_name.argtypes = [ct.c_char_p, ct.c_int, dbl, dbl, dbl,
                                            ct.c_int, ct.POINTER(vp)]
    L.sfx_solver_destroy.argtypes = [vp]
    L.sfx_solver_tag.argtypes = [vp]
    L.sfx_solver_tag.restype = ct.c_char_p
    L.sfx_solver_gamma.argtypes = [vp]
    L.sfx_solver_gamma.restype = dbl
    L.sfx_solver_ndim.argtypes = [vp]
    L.sfx_solver_precision.argtypes = [vp]
    L.sfx_solver_get_stats.argtypes = [vp, ct.POINTER(Stats)]
    L.sfx_solver_reset_stats.argtypes = [vp]
    L.sfx_solver_enable_stats.argtypes = [vp, ct.c_int]
    L.sfx_solve.argtypes = [vp, _dp, _dp, _dp, _dp, ct.c_int, _dp, _u8p]
    L.sfx_solve_for_flux.argtypes = [vp, _dp, _dp, _dp, _dp, ct.c_int, _dp, _u8p]
    L.sfx_solve_for_flux_batch.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.sfx_solve_for_flux_batch_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.sfx_exact_solve_batch.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.sfx_time_derivative_batch.argtypes = [vp, i64, vp, vp, vp, vp]
    L.sfx_solver_set_split.argtypes = [vp, ct.c_int]
    L.sfx_solver_get_split.argtypes = [vp]
    L.sfx_get_W_batch.argtypes = [vp, i64, vp, vp, vp, vp]
    L.sfx_get_Q_batch.argtypes = [vp, i64, vp, vp, vp, vp]
    L.sfx_hydro_timestep.argtypes = [vp, i64, vp, vp, vp, ct.POINTER(Timeline), vp, vp, vp, vp,
                                     vp, vp, vp]
    for name, args in _mesh_sigs().items():
        if hasattr(L, name):
            getattr(L, name).argtypes = args
    L.sfx_mesh_entry_groups.restype = ct.c_int64
    L.sfx_mesh_entry_groups_sorted.restype = ct.c_int64
    _lib = L
    return L


class Timeline(ct.Structure):
    """sfx_timeline: the integer timeline of TimeLine (src/TimeLine.hpp:52-113)."""
    _fields_ = [("integer_maxtime", ct.c_uint64), ("current_time", ct.c_uint64),
                ("timestep", ct.c_uint64), ("max_timestep", ct.c_uint64),
                ("min_timestep", ct.c_uint64), ("tottime", ct.c_double),
                ("mintime", ct.c_double), ("cfl", ct.c_double),
                ("global_timestep", ct.c_int32), ("pad", ct.c_int32)]


def _mesh_sigs():
    md = ct.POINTER(MeshDesc)
    return {
        "sfx_mesh_entry_geometry": [vp, md, i64, vp, vp],
        "sfx_mesh_entry_record": [ct.c_int, ct.c_int],
        "sfx_mesh_face_records": [vp, md, vp, vp],
        "sfx_mesh_face_record_planes": [ct.c_int, ct.c_int],
        "sfx_mesh_entry_groups": [i64, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)],
        "sfx_mesh_entry_groups_sorted": [i64, ct.POINTER(ct.c_int64), i64,
                                         ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)],
        "sfx_gradients": [vp, md, vp, vp],
        "sfx_face_fluxes": [vp, md, vp, vp, vp, vp],
        "sfx_gather_dQ": [vp, md, vp, vp, vp],
        "sfx_face_flux_pass": [vp, md, vp, vp, vp, vp, vp],
        "sfx_face_flux_pass_static": [vp, md, vp, vp, vp, vp],
        "sfx_update_step": [vp, i64, vp, vp, vp, vp, vp, vp, vp],
        "sfx_update_step_timed": [vp, i64, vp, vp, vp, vp, vp, vp, ct.c_uint64, vp, vp],
    }


def check(rc, what="sfx call"):
    if rc != SFX_OK:
        raise SfxError("%s failed (%d): %s" % (what, rc, lib().sfx_last_error().decode()))
