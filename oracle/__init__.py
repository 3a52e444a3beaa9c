"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle restates the reference's CPU algorithm for the P1 multigrid path
(see hyteg_oracle.c for the reference files it follows).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
as the checker or as the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

INNER, DIRICHLET, NEUMANN, FREESLIP, ALL = 1, 2, 4, 8, 15


def build(force=False):
    """Compile the oracle with gcc (OpenMP over primitives, as the reference's optional OpenMP)."""
    src = os.path.join(_HERE, "hyteg_oracle.c")
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared",
                               "-o", _SO, src, "-lm"])
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        vp, i, d, l = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        P = ctypes.POINTER
        sig = {
            "om_mesh_cuboid": (vp, [d, d, d, d, d, d, i, i, i]),
            "om_mesh_create": (vp, [i, P(d), i, P(i)]),
            "om_mesh_free": (None, [vp]),
            "om_mesh_counts": (None, [vp, P(i)]),
            "om_prim_type": (i, [vp, i, i]),
            "om_edge_vertices": (None, [vp, i, P(i)]),
            "om_face_vertices": (None, [vp, i, P(i)]),
            "om_cell_vertices": (None, [vp, i, P(i)]),
            "om_face_num_cells": (i, [vp, i]),
            "om_face_cells": (None, [vp, i, P(i)]),
            "om_cell_faces": (None, [vp, i, P(i)]),
            "om_cell_face_map": (None, [vp, i, P(i)]),
            "om_owned_count": (l, [vp, i]),
            "om_owned_coords": (None, [vp, i, P(d)]),
            "om_owned_types": (None, [vp, i, P(i)]),
            "om_num_dofs": (l, [vp, i, i]),
            "om_func_create": (vp, [vp, i, i]),
            "om_func_free": (None, [vp]),
            "om_func_get_owned": (None, [vp, i, P(d)]),
            "om_func_set_owned": (None, [vp, i, P(d)]),
            "om_func_prim_size": (l, [vp, i, i, i]),
            "om_func_get_prim": (None, [vp, i, i, i, P(d)]),
            "om_func_set_prim": (None, [vp, i, i, i, P(d)]),
            "om_communicate": (None, [vp, i, i]),
            "om_interpolate_const": (None, [vp, i, i, d]),
            "om_assign": (None, [vp, i, i, i, P(d), P(vp)]),
            "om_add": (None, [vp, i, i, i, P(d), P(vp)]),
            "om_dot": (d, [vp, vp, i, i]),
            "om_op_create": (vp, [vp, i, i, i]),
            "om_op_free": (None, [vp]),
            "om_cell_stencil": (None, [vp, i, i, P(d)]),
            "om_vertex_stencil": (i, [vp, i, i, P(d)]),
            "om_edge_stencil": (i, [vp, i, i, P(d)]),
            "om_face_stencil": (i, [vp, i, i, i, P(d)]),
            "om_apply": (None, [vp, vp, vp, i, i, i]),
            "om_smooth_sor": (None, [vp, vp, vp, d, i, i, i]),
            "om_smooth_sor_backward": (None, [vp, vp, vp, d, i, i, i]),
            "om_gmg_set_variant": (None, [vp, i, i]),
            "om_gmg_set_smoothing": (None, [vp, i, i, i]),
            "om_fmg_solve": (None, [vp, vp, vp, vp, i, i]),
            "om_restrict": (None, [vp, i, i]),
            "om_prolongate": (None, [vp, i, i, i]),
            "om_cg_create": (vp, [vp, i, i, i, d]),
            "om_cg_free": (None, [vp]),
            "om_cg_solve": (i, [vp, vp, vp, vp, i]),
            "om_gmg_create": (vp, [vp, i, i, i, i, d, i, i, d]),
            "om_gmg_free": (None, [vp]),
            "om_gmg_solve": (None, [vp, vp, vp, vp, i]),
            "om_cell_index": (l, [l, l, l, l]),
            "om_face_index": (l, [l, l, l]),
            "om_wtime": (d, []),
            "om_num_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Mesh:
    def __init__(self, handle):
        self.h = handle

    @staticmethod
    def cuboid(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        return Mesh(lib().om_mesh_cuboid(*lo, *hi, nx, ny, nz))

    @staticmethod
    def from_arrays(coords, cells):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        return Mesh(lib().om_mesh_create(len(coords), _dp(coords), len(cells), _ip(cells)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_mesh_free(self.h)
            self.h = None

    def counts(self):
        out = np.zeros(4, dtype=np.int32)
        lib().om_mesh_counts(self.h, _ip(out))
        return tuple(int(x) for x in out)

    def prim_type(self, cls, idx):
        return lib().om_prim_type(self.h, cls, idx)

    def owned_count(self, level):
        return lib().om_owned_count(self.h, level)

    def owned_coords(self, level):
        n = self.owned_count(level)
        out = np.zeros((n, 3))
        lib().om_owned_coords(self.h, level, _dp(out))
        return out

    def owned_types(self, level):
        n = self.owned_count(level)
        out = np.zeros(n, dtype=np.int32)
        lib().om_owned_types(self.h, level, _ip(out))
        return out

    def num_dofs(self, level, flag=ALL):
        return lib().om_num_dofs(self.h, level, flag)

    def cell_vertices(self, c):
        out = np.zeros(4, dtype=np.int32)
        lib().om_cell_vertices(self.h, c, _ip(out))
        return out

    def face_vertices(self, f):
        out = np.zeros(3, dtype=np.int32)
        lib().om_face_vertices(self.h, f, _ip(out))
        return out

    def edge_vertices(self, e):
        out = np.zeros(2, dtype=np.int32)
        lib().om_edge_vertices(self.h, e, _ip(out))
        return out

    def face_cells(self, f):
        out = np.zeros(2, dtype=np.int32)
        lib().om_face_cells(self.h, f, _ip(out))
        return out

    def cell_face_map(self, c):
        out = np.zeros(12, dtype=np.int32)
        lib().om_cell_face_map(self.h, c, _ip(out))
        return out.reshape(4, 3)


class Function:
    def __init__(self, mesh, minL, maxL):
        self.mesh = mesh
        self.minL, self.maxL = minL, maxL
        self.h = lib().om_func_create(mesh.h, minL, maxL)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_func_free(self.h)
            self.h = None

    def get(self, level):
        out = np.zeros(self.mesh.owned_count(level))
        lib().om_func_get_owned(self.h, level, _dp(out))
        return out

    def set(self, level, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.mesh.owned_count(level)
        lib().om_func_set_owned(self.h, level, _dp(v))

    def prim(self, level, cls, idx):
        n = lib().om_func_prim_size(self.h, level, cls, idx)
        out = np.zeros(n)
        lib().om_func_get_prim(self.h, level, cls, idx, _dp(out))
        return out

    def set_prim(self, level, cls, idx, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        lib().om_func_set_prim(self.h, level, cls, idx, _dp(v))

    def communicate(self, level, direction):
        lib().om_communicate(self.h, level, direction)

    def interpolate(self, level, value, flag=ALL):
        lib().om_interpolate_const(self.h, level, flag, float(value))

    def assign(self, level, coeffs, funcs, flag=ALL):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        arr = (ctypes.c_void_p * len(funcs))(*[f.h for f in funcs])
        lib().om_assign(self.h, level, flag, len(funcs), _dp(c), arr)

    def add(self, level, coeffs, funcs, flag=ALL):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        arr = (ctypes.c_void_p * len(funcs))(*[f.h for f in funcs])
        lib().om_add(self.h, level, flag, len(funcs), _dp(c), arr)

    def dot(self, other, level, flag=ALL):
        return lib().om_dot(self.h, other.h, level, flag)


class Operator:
    """form 0: P1 diffusion (P1ConstantLaplaceOperator); form 1: P1 mass."""

    def __init__(self, mesh, minL, maxL, form=0):
        self.mesh = mesh
        self.minL, self.maxL = minL, maxL
        self.h = lib().om_op_create(mesh.h, minL, maxL, form)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_op_free(self.h)
            self.h = None

    def cell_stencil(self, level, c):
        out = np.zeros(15)
        lib().om_cell_stencil(self.h, level, c, _dp(out))
        return out

    def vertex_stencil(self, level, v):
        n = lib().om_vertex_stencil(self.h, level, v, None)
        out = np.zeros(n)
        lib().om_vertex_stencil(self.h, level, v, _dp(out))
        return out

    def edge_stencil(self, level, e):
        n = lib().om_edge_stencil(self.h, level, e, None)
        out = np.zeros(n)
        lib().om_edge_stencil(self.h, level, e, _dp(out))
        return out

    def face_stencil(self, level, f, n):
        out = np.zeros(15)
        lib().om_face_stencil(self.h, level, f, n, _dp(out))
        return out

    def apply(self, src, dst, level, flag=INNER | NEUMANN, update=0):
        lib().om_apply(self.h, src.h, dst.h, level, flag, update)

    def smooth_sor(self, x, b, level, omega=1.0, flag=INNER | NEUMANN, order=0):
        lib().om_smooth_sor(self.h, x.h, b.h, float(omega), level, flag, order)

    def smooth_sor_backward(self, x, b, level, omega=1.0, flag=INNER | NEUMANN, order=0):
        lib().om_smooth_sor_backward(self.h, x.h, b.h, float(omega), level, flag, order)


def restrict(f, src_level, flag=INNER | NEUMANN):
    lib().om_restrict(f.h, src_level, flag)


def prolongate(f, src_level, flag=INNER | NEUMANN, add=True):
    lib().om_prolongate(f.h, src_level, flag, 1 if add else 0)


class CG:
    def __init__(self, mesh, minL, maxL, max_iter=10000, tol=1e-16):
        self.h = lib().om_cg_create(mesh.h, minL, maxL, max_iter, tol)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_cg_free(self.h)
            self.h = None

    def solve(self, A, x, b, level):
        return lib().om_cg_solve(self.h, A.h, x.h, b.h, level)


class GMG:
    def __init__(self, mesh, minL, maxL, nu1=3, nu2=3, omega=1.0, order=0, cg_max_iter=10000, cg_tol=1e-16,
                 wcycle=False, symmetric=False):
        self.h = lib().om_gmg_create(mesh.h, minL, maxL, nu1, nu2, float(omega), order, cg_max_iter, cg_tol)
        lib().om_gmg_set_variant(self.h, int(wcycle), int(symmetric))

    def set_smoothing(self, pre, post, increment=0):
        """setSmoothingSteps(pre, post, smoothIncrement)."""
        lib().om_gmg_set_smoothing(self.h, pre, post, increment)

    def fmg(self, A, x, b, level, cycles_per_level=1):
        """FullMultigridSolver: cycles on every level from the coarsest up, prolongation (Replace) between."""
        lib().om_fmg_solve(self.h, A.h, x.h, b.h, level, cycles_per_level)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_gmg_free(self.h)
            self.h = None

    def solve(self, A, x, b, level):
        lib().om_gmg_solve(self.h, A.h, x.h, b.h, level)


def cell_index(w, x, y, z):
    return lib().om_cell_index(w, x, y, z)


def face_index(w, x, y):
    return lib().om_face_index(w, x, y)


def wtime():
    return lib().om_wtime()


def num_threads():
    return lib().om_num_threads()
