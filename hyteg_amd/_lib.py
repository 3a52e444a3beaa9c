"""ctypes binding of libhymi.so (the C-ABI declared in include/hymi.h).

The library is loaded from the package directory; a missing library raises instead of falling back to
anything else.
"""
import atexit
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhymi.so")

OK, ERR_INVALID, ERR_LEVEL, ERR_OOM, ERR_HIP, ERR_UNSUPPORTED, ERR_INTERNAL = range(7)
INNER, DIRICHLET, NEUMANN, FREESLIP, ALL = 1, 2, 4, 8, 15
REPLACE, ADD = 0, 1
F64, F32 = 0, 1
VERTEX, EDGE, FACE, CELL = 0, 1, 2, 3


# Set when the interpreter starts shutting down.  From then on the Python wrappers release nothing: module
# teardown destroys objects in no particular order (a context could go before the functions on it), and the
# process exit returns every device allocation anyway.
closing = False


def _mark_closing():
    global closing
    closing = True


atexit.register(_mark_closing)


def releasable():
    """True while handles may still be destroyed (library loaded, interpreter not shutting down)."""
    return _lib is not None and not closing


class HymiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hymi error {code}: {msg}")
        self.code = code


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_P = ctypes.POINTER

# name -> argument types (all return int status except where noted)
SIGNATURES = {
    "hymi_version": [],
    "hymi_ctx_create": [_i, _P(_vp)],
    "hymi_ctx_destroy": [_vp],
    "hymi_ctx_synchronize": [_vp],
    "hymi_ctx_bytes_in_use": [_vp, _P(_i64)],
    "hymi_mesh_create_cuboid": [_vp, _P(_d), _P(_d), _i, _i, _i, _P(_vp)],
    "hymi_mesh_create": [_vp, _i, _P(_d), _i, _P(_i), _P(_vp)],
    "hymi_mesh_from_gmsh": [_vp, ctypes.c_char_p, _P(_vp)],
    "hymi_mesh_destroy": [_vp],
    "hymi_mesh_counts": [_vp, _P(_i)],
    "hymi_mesh_primitive_type": [_vp, _i, _i, _P(_i)],
    "hymi_mesh_primitive_coordinates": [_vp, _i, _i, _P(_d)],
    "hymi_mesh_owned_count": [_vp, _i, _P(_i64)],
    "hymi_mesh_owned_coords": [_vp, _i, _P(_d)],
    "hymi_mesh_owned_types": [_vp, _i, _P(_i)],
    "hymi_function_create": [_vp, _i, _i, _i, _i, _P(_vp)],
    "hymi_function_destroy": [_vp],
    "hymi_function_upload_owned": [_vp, _i, _P(_d)],
    "hymi_function_download_owned": [_vp, _i, _P(_d)],
    "hymi_function_primitive_size": [_vp, _i, _i, _i, _P(_i64)],
    "hymi_function_download_primitive": [_vp, _i, _i, _i, _P(_d)],
    "hymi_function_download_micro_vertices": [_vp, _i, _i, _P(_d)],
    "hymi_function_upload_primitive": [_vp, _i, _i, _i, _P(_d)],
    "hymi_interpolate_const": [_vp, _i, _i, _d],
    "hymi_assign": [_vp, _i, _i, _i, _P(_d), _P(_vp)],
    "hymi_add": [_vp, _i, _i, _i, _P(_d), _P(_vp)],
    "hymi_dot": [_vp, _vp, _i, _i, _P(_d)],
    "hymi_dot_noalloc": [_vp, _vp, _i, _i, _vp],
    "hymi_convert": [_vp, _vp, _i, _i],
    "hymi_laplace_create": [_vp, _i, _i, _P(_vp)],
    "hymi_mass_create": [_vp, _i, _i, _P(_vp)],
    "hymi_operator_destroy": [_vp],
    "hymi_operator_stencil": [_vp, _i, _i, _i, _P(_d), _P(_i)],
    "hymi_apply": [_vp, _vp, _vp, _i, _i, _i],
    "hymi_residual": [_vp, _vp, _vp, _vp, _i, _i],
    "hymi_smooth_sor": [_vp, _vp, _vp, _d, _i, _i],
    "hymi_smooth_sor_backwards": [_vp, _vp, _vp, _d, _i, _i],
    "hymi_restrict": [_vp, _i, _i],
    "hymi_prolongate_add": [_vp, _i, _i],
    "hymi_prolongate": [_vp, _i, _i],
    "hymi_smooth_gs": [_vp, _vp, _vp, _i, _i],
    "hymi_smooth_gs_backwards": [_vp, _vp, _vp, _i, _i],
    "hymi_gmg_set_smoothing": [_vp, _i, _i, _i],
    "hymi_cg_solve": [_vp, _vp, _vp, _i, _i, _d],
    "hymi_gmg_create": [_vp, _vp, _i, _i, _i, _i, _d, _i, _d, _P(_vp)],
    "hymi_gmg_destroy": [_vp],
    "hymi_gmg_vcycle": [_vp, _vp, _vp, _i],
    "hymi_gmg_vcycle_graph": [_vp, _vp, _vp, _i, _i],
    "hymi_gmg_set_variant": [_vp, _i, _i],
    "hymi_fmg_solve": [_vp, _vp, _vp, _i, _i],
    "hymi_time_kernel": [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _P(_d)],
    "hymi_mesh_set_schedule": [_vp, _i, _i, _i, _i, _i],
    "hymi_interpolate_modes": [_vp, _i, _i, _i, _P(_d)],
    "hymi_solve_refined": [_vp, _vp, _vp, _i, _d, _i, _i, _P(_i), _P(_d), _P(_d)],
    "hymi_refine_step": [_vp, _vp, _vp, _i, _i, ctypes.c_uint, _P(_d)],
    "hymi_function_select_columns": [_vp, _i, ctypes.c_uint],
    "hymi_function_download_owned_column": [_vp, _i, _i, _P(_d)],
    "hymi_mesh_get_schedule": [_vp, _i, _P(_i)],
    "hymi_mesh_set_cell_sor": [_vp, _i, _i, _i],
    "hymi_mesh_get_cell_sor": [_vp, _i, _P(_i)],
    "hymi_mesh_set_residual_restrict": [_vp, _i, _i],
    "hymi_mesh_get_residual_restrict": [_vp, _i, _P(_i)],
    "hymi_mesh_autotune": [_vp, _i, _i, _i, _P(_d), _i, _P(_i)],
}

_lib = None


def load():
    """Load libhymi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m hyteg_amd.build` first")
        lib = ctypes.CDLL(LIB_PATH)
        lib.hymi_last_error_string.restype = ctypes.c_char_p
        lib.hymi_last_error_string.argtypes = []
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != OK:
        raise HymiError(rc, load().hymi_last_error_string().decode())
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
