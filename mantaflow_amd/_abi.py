"""ctypes binding of libmfx.so (declared in include/mfx.h).

The library is loaded from this package directory; there is no fallback: if
it is missing or fails to load, importing raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# MFX_LIB names another build of the library (the sanitizer test's)
LIB_PATH = os.environ.get("MFX_LIB") or os.path.join(_HERE, "libmfx.so")

GRID_REAL, GRID_INT, GRID_VEC3, GRID_MAC, GRID_LEVELSET, GRID_FLAGS = 1, 2, 4, 12, 17, 34
RED_MAX, RED_MIN, RED_MAXABS, RED_L1, RED_L2, RED_SUM = 0, 1, 2, 3, 4, 5
SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYLINDER = 1, 2, 3


class PressureOpts(C.Structure):
    _fields_ = [("cgAccuracy", C.c_float), ("cgMaxIterFac", C.c_float), ("preconditioner", C.c_int),
                ("enforceCompatibility", C.c_int), ("useL2Norm", C.c_int),
                ("zeroPressureFixing", C.c_int), ("relTol", C.c_double),
                ("maxIterOverride", C.c_int), ("precision", C.c_int),
                ("trueResidualEvery", C.c_int), ("deferCorrection", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("resNorm", C.c_float), ("relResidual", C.c_double),
                ("fixPidx", C.c_int), ("maxIter", C.c_int), ("preconditionerUsed", C.c_int),
                ("converged", C.c_int), ("seconds", C.c_double)]


class PressureBnd(C.Structure):
    _fields_ = [("phi", C.c_void_p), ("perCellCorr", C.c_void_p), ("fractions", C.c_void_p),
                ("obvel", C.c_void_p), ("curv", C.c_void_p), ("gfClamp", C.c_float), ("surfTens", C.c_float)]


class ServeOpts(C.Structure):
    _fields_ = [("maxBatch", C.c_int), ("workers", C.c_int), ("hbmReserve", C.c_size_t),
                ("memoryLimit", C.c_size_t), ("solve", PressureOpts)]


class ServeSession(C.Structure):
    _fields_ = [("flags", C.c_void_p), ("vel", C.c_void_p), ("pressure", C.c_void_p), ("density", C.c_void_p),
                ("keep", C.c_void_p), ("srcval", C.c_void_p), ("gravity", C.c_float * 3)]


class ServeResult(C.Structure):
    _fields_ = [("tag", C.c_longlong), ("session", C.c_int), ("kind", C.c_int), ("status", C.c_int),
                ("batch", C.c_int), ("tick", C.c_longlong), ("submitted", C.c_double), ("done", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("relResidual", C.c_double), ("error", C.c_char * 200)]


REQ_STEP, REQ_SOLVE = 0, 1

# every exported symbol of include/mfx.h with its ctypes signature
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "mfx_last_error": (C.c_char_p, []),
    "mfx_device_count": (_I, [C.POINTER(_I)]),
    "mfx_device_memory": (_I, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mfx_scene_create": (_I, [_I, _I, _I, _I, C.POINTER(_P)]),
    "mfx_scene_destroy": (_I, [_P]),
    "mfx_scene_set_dt": (_I, [_P, _F]),
    "mfx_scene_get_dt": (_I, [_P, _FP]),
    "mfx_scene_sync": (_I, [_P]),
    "mfx_scene_memory": (_I, [_P, C.POINTER(C.c_size_t)]),
    "mfx_grid_create": (_I, [_P, _I, C.POINTER(_P)]),
    "mfx_grid_destroy": (_I, [_P]),
    "mfx_grid_type": (_I, [_P, C.POINTER(_I)]),
    "mfx_grid_upload": (_I, [_P, _P, C.c_size_t]),
    "mfx_grid_download": (_I, [_P, _P, C.c_size_t]),
    "mfx_grid_clear": (_I, [_P]),
    "mfx_grid_copy": (_I, [_P, _P]),
    "mfx_grid_swap": (_I, [_P, _P]),
    "mfx_grid_set_const": (_I, [_P, _FP]),
    "mfx_grid_add_const": (_I, [_P, _FP]),
    "mfx_grid_mult_const": (_I, [_P, _FP]),
    "mfx_grid_add": (_I, [_P, _P]),
    "mfx_grid_sub": (_I, [_P, _P]),
    "mfx_grid_mult": (_I, [_P, _P]),
    "mfx_grid_mult_add": (_I, [_P, _P, _P]),
    "mfx_grid_add_scaled": (_I, [_P, _P, _FP]),
    "mfx_grid_clamp": (_I, [_P, _F, _F]),
    "mfx_grid_reduce": (_I, [_P, _I, _I, C.POINTER(C.c_double)]),
    "mfx_grid_set_bound": (_I, [_P, _FP, _I]),
    "mfx_grid_set_bound_neumann": (_I, [_P, _I]),
    "mfx_grid_set_bound_mac": (_I, [_P, _FP, _I, _I]),
    "mfx_grid_max_diff": (_I, [_P, _P, C.POINTER(C.c_double)]),
    "mfx_get_component": (_I, [_P, _P, _I]),
    "mfx_set_component": (_I, [_P, _P, _I]),
    "mfx_copy_mac_to_vec3": (_I, [_P, _P]),
    "mfx_flags_init_domain": (_I, [_P, _I, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
    "mfx_flags_fill_grid": (_I, [_P, _I]),
    "mfx_flags_count_cells": (_I, [_P, _I, _I, _P, C.POINTER(C.c_longlong)]),
    "mfx_set_open_bound": (_I, [_P, _I, C.c_char_p, _I]),
    "mfx_reset_outflow": (_I, [_P, _P, _P]),
    "mfx_apply_shape": (_I, [_P, _I, _FP, _FP, _F, _FP, _P]),
    "mfx_advect_semi_lagrange": (_I, [_P, _P, _P, _I, _F, _I, _I, _I]),
    "mfx_set_wall_bcs": (_I, [_P, _P, _P]),
    "mfx_set_wall_bcs_frac": (_I, [_P, _P, _P, _P, _P, _I]),
    "mfx_pressure_bnd_default": (None, [C.POINTER(PressureBnd)]),
    "mfx_solve_pressure_bnd": (_I, [_P, _P, _P, C.POINTER(PressureOpts), C.POINTER(PressureBnd),
                                    C.POINTER(SolveStats), _P]),
    "mfx_compute_pressure_rhs_bnd": (_I, [_P, _P, _P, C.POINTER(PressureOpts), C.POINTER(PressureBnd)]),
    "mfx_solve_pressure_system_bnd": (_I, [_P, _P, _P, C.POINTER(PressureOpts), C.POINTER(PressureBnd),
                                           C.POINTER(SolveStats)]),
    "mfx_correct_velocity_bnd": (_I, [_P, _P, _P, C.POINTER(PressureBnd)]),
    "mfx_add_buoyancy": (_I, [_P, _P, _P, _FP, _F, _I]),
    "mfx_set_wall_bcs_buoyancy": (_I, [_P, _P, _P, _FP, _F, _I]),
    "mfx_pressure_opts_default": (None, [C.POINTER(PressureOpts)]),
    "mfx_solve_pressure": (_I, [_P, _P, _P, C.POINTER(PressureOpts), C.POINTER(SolveStats), _P]),
    "mfx_solve_pressure_noalloc": (_I, [_P, _P, _P, C.POINTER(PressureOpts), C.POINTER(SolveStats)]),
    "mfx_cg_solve_diffusion": (_I, [_P, _P, _F, _F, _F, C.POINTER(SolveStats)]),
    "mfx_cg_solve_we": (_I, [_P, _P, _P, _P, _I, _F, _F, _F, C.POINTER(SolveStats)]),
    "mfx_compute_pressure_rhs": (_I, [_P, _P, _P, C.POINTER(PressureOpts)]),
    "mfx_solve_pressure_system": (_I, [_P, _P, _P, C.POINTER(PressureOpts), C.POINTER(SolveStats)]),
    "mfx_solve_pressure_batch": (_I, [C.c_int, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                      C.POINTER(PressureOpts), C.POINTER(SolveStats)]),
    "mfx_correct_velocity": (_I, [_P, _P, _P]),
    "mfx_correct_velocity_wall_bcs": (_I, [_P, _P, _P]),
    "mfx_release_mg": (_I, [_P]),
    "mfx_serve_opts_default": (None, [C.POINTER(ServeOpts)]),
    "mfx_serve_create": (_I, [C.POINTER(ServeOpts), C.POINTER(_P)]),
    "mfx_serve_add_session": (_I, [_P, C.POINTER(ServeSession), C.POINTER(_I)]),
    "mfx_serve_submit": (_I, [_P, _I, _I, C.c_longlong]),
    "mfx_serve_submit_many": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_longlong)]),
    "mfx_serve_poll": (_I, [_P, C.POINTER(ServeResult), _I, C.POINTER(_I), C.c_double]),
    "mfx_serve_drain": (_I, [_P, C.c_double]),
    "mfx_serve_time": (_I, [_P, C.POINTER(C.c_double)]),
    "mfx_serve_destroy": (_I, [_P]),
    "mfx_kernel_timing": (_I, [_I]),
    "mfx_kernel_timing_reset": (_I, []),
    "mfx_kernel_timing_get": (_I, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
}

_lib = None


def lib():
    """Load libmfx.so (once).  Raises if it is not built or cannot be loaded."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmfx.so is not built (%s); run python -m mantaflow_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MfxError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise MfxError(lib().mfx_last_error().decode(errors="replace"))
    return rc


def floats(*v):
    arr = (C.c_float * max(3, len(v)))(*v)
    return arr
