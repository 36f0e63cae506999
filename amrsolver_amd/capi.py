"""ctypes binding of the C ABI declared in include/amrsolver_hip.h.

The shared library is built in-tree (``python -m amrsolver_amd.build``) into
``amrsolver_amd/lib/libamrsolver_hip.so``.  There is no fallback: if the
library is missing every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libamrsolver_hip.so")

AMR_OK = 0
AMR_ERR_ARG = -1
AMR_ERR_HIP = -2
AMR_ERR_STATE = -3
NCONS = 5
NGHOST = 3


class Closures(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double), ("mw", ctypes.c_double), ("Ru", ctypes.c_double),
        ("cv", ctypes.c_double), ("cp", ctypes.c_double), ("Rspec", ctypes.c_double),
        ("Cshock", ctypes.c_double), ("Cdamp", ctypes.c_double),
        ("visc_law", ctypes.c_int32), ("visc_ref", ctypes.c_double), ("Tvisc_ref", ctypes.c_double),
        ("cond_law", ctypes.c_int32), ("cond_ref", ctypes.c_double), ("Tcond_ref", ctypes.c_double),
        ("Pr", ctypes.c_double),
    ]


class Scheme(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "rhs_euler", "rhs_visc", "rhs_source", "flux_euler", "order_keep", "art_diss",
        "order_rk", "stages_rk")]


class Geom(ctypes.Structure):
    _fields_ = [
        ("dom_lo", ctypes.c_int32 * 3), ("dom_hi", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3), ("lo_bc", ctypes.c_int32 * 3), ("hi_bc", ctypes.c_int32 * 3),
        ("dx", ctypes.c_double * 3), ("prob_lo", ctypes.c_double * 3), ("prob_hi", ctypes.c_double * 3),
    ]


class Problem(ctypes.Structure):
    _fields_ = [("wall_T", ctypes.c_double), ("body_force_x", ctypes.c_double)]


AMR_TAG_DENGRAD = 1
AMR_TAG_WALLBAND = 2
AMR_TAG_VORT = 3


class Tagger(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad_", ctypes.c_int32), ("threshold", ctypes.c_double),
                ("h", ctypes.c_double), ("factor", ctypes.c_double)]


class Fab(ctypes.Structure):
    _fields_ = [
        ("lo", ctypes.c_int32 * 3), ("hi", ctypes.c_int32 * 3), ("ng", ctypes.c_int32),
        ("pad_", ctypes.c_int32), ("data", ctypes.c_void_p),
        ("jstride", ctypes.c_int64), ("kstride", ctypes.c_int64), ("nstride", ctypes.c_int64),
    ]


# entry point -> argtypes (the exact list of functions include/amrsolver_hip.h declares)
_P = ctypes.c_void_p
_PF = ctypes.POINTER(Fab)
SIGNATURES = {
    "amr_last_error": ([], ctypes.c_char_p),
    "amr_ctx_create": ([ctypes.c_int, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "amr_ctx_destroy": ([_P], ctypes.c_int),
    "amr_ctx_synchronize": ([_P], ctypes.c_int),
    "amr_boxtab_create": ([_P, ctypes.POINTER(Geom), ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                           ctypes.POINTER(_P)], ctypes.c_int),
    "amr_boxtab_destroy": ([_P], ctypes.c_int),
    "amr_fill_boundary": ([_P, _PF], ctypes.c_int),
    "amr_fill_two_levels": ([_P, _P, ctypes.c_int, ctypes.c_double, _PF, ctypes.c_double, _PF, ctypes.c_int,
                             _PF], ctypes.c_int),
    "amr_fill_two_levels_cf": ([_P, _P, ctypes.c_int, ctypes.c_double, _PF, ctypes.c_double, _PF, ctypes.c_int,
                                _PF], ctypes.c_int),
    "amr_average_down": ([_P, _P, ctypes.c_int, _PF, _PF], ctypes.c_int),
    "amr_set_user_source": ([_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int], ctypes.c_int),
    "amr_face_fluxes": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, _PF, _PF, _PF], ctypes.c_int),
    "amr_tag_cells": ([_P, ctypes.POINTER(Tagger), _PF, _P], ctypes.c_int),
    "amr_tag_blocks": ([_P, _P, ctypes.c_int, _P], ctypes.c_int),
    "amr_regrid_fill": ([_P, _P, _PF, _P, ctypes.c_int, ctypes.c_double, _PF, ctypes.c_double, _PF, _PF],
                        ctypes.c_int),
    "amr_plot_vars": ([_P, ctypes.POINTER(Closures), _PF, ctypes.c_int, ctypes.c_int, _PF], ctypes.c_int),
    "amr_write_plotfile": ([ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(_PF), ctypes.c_int,
                            ctypes.POINTER(ctypes.c_char_p), ctypes.c_double, ctypes.POINTER(ctypes.c_int32),
                            ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "amr_rhs": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, _PF], ctypes.c_int),
    "amr_rk_stage": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), ctypes.c_int,
                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                      _PF, _PF, _PF], ctypes.c_int),
    "amr_stage_coeffs": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                         ctypes.c_int),
    "amr_advance": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, ctypes.c_double,
                     ctypes.c_double], ctypes.c_int),
    "amr_advance_graph": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF,
                           ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "amr_clock_set": ([_P, ctypes.c_double], ctypes.c_int),
    "amr_clock_get": ([_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "amr_advance_cfl": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, ctypes.c_double,
                         ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "amr_advance_cfl_graph": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "amr_new_dt_levels": ([ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(Closures), ctypes.POINTER(_PF),
                           ctypes.POINTER(ctypes.c_int32), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                           ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "amr_est_dt": ([_P, ctypes.POINTER(Closures), _PF, ctypes.c_double,
                    ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "amr_est_dt_noalloc": ([_P, ctypes.POINTER(Closures), _PF, ctypes.c_double,
                            ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "amr_diagnostics": ([_P, ctypes.POINTER(Closures), _PF, ctypes.c_double] + [ctypes.POINTER(ctypes.c_double)] * 5,
                        ctypes.c_int),
    "amr_check_state": ([_P, ctypes.POINTER(Closures), _PF, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "amr_set_state_check": ([_P, ctypes.c_int], ctypes.c_int),
    "amr_set_tiling_form": ([_P, ctypes.POINTER(Scheme), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
                            ctypes.c_int),
    "amr_totals": ([_P, _PF, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "amr_batch_create": ([_P, ctypes.c_int, ctypes.POINTER(Geom), ctypes.POINTER(_P)], ctypes.c_int),
    "amr_batch_advance": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF,
                           ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "amr_batch_est_dt": ([_P, ctypes.POINTER(Closures), _PF, ctypes.c_double,
                          ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "amr_set_tiling": ([_P, ctypes.POINTER(Scheme), ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "amr_set_box_tiling": ([_P, ctypes.POINTER(Scheme), ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "amr_step_plan": ([_P, ctypes.POINTER(Scheme), _PF, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "amr_autotune": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF,
                      ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "amr_set_source": ([_P, _PF], ctypes.c_int),
    "amr_set_problem": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Problem)], ctypes.c_int),
    "amr_vjp_rhs": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, _PF, _PF], ctypes.c_int),
    "amr_vjp_stage": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), ctypes.c_int,
                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                       _PF, _PF, _PF, _PF, ctypes.c_int], ctypes.c_int),
    "amr_adjoint_run": ([_P, ctypes.POINTER(Closures), ctypes.POINTER(Scheme), _PF, ctypes.c_double,
                         ctypes.c_int, ctypes.c_int, ctypes.c_int, _PF, _PF,
                         ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
}


class AmrError(RuntimeError):
    pass


_lib = None


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise AmrError(f"HIP library not built: {LIBPATH} missing "
                       "(run `python -m amrsolver_amd.build`)")
    lib = ctypes.CDLL(LIBPATH)
    for name, (args, res) in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != AMR_OK:
        msg = lib.amr_last_error().decode()
        raise AmrError(f"{name} failed ({rc}): {msg}")
    return rc


def stage_coeffs(order_rk, stages_rk, dt, stage):
    n = ctypes.c_int()
    mode = ctypes.c_int()
    a, b, c, d = (ctypes.c_double() for _ in range(4))
    call("amr_stage_coeffs", order_rk, stages_rk, dt, stage, ctypes.byref(n), ctypes.byref(mode),
         ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
    return n.value, mode.value, a.value, b.value, c.value, d.value
