"""hyteg_amd -- MI355X-native matrix-free P1 geometric multigrid with HyTeG's operator/solver API.

Python mirror of the reference's C++ interface for this path (classes and methods carry the reference's
names and argument meaning); every call goes through the C-ABI of libhymi.so (include/hymi.h), whose
kernels run on the GPU.  Nothing here computes on the CPU.

    storage = PrimitiveStorage(MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 2, 2, 2))
    u = P1Function("u", storage, 0, 7); f = P1Function("f", storage, 0, 7)
    A = P1ConstantLaplaceOperator(storage, 0, 7)
    gmg = GeometricMultigridSolver(storage, SORSmoother(1.0), CGSolver(storage, 0, 0),
                                   P1toP1LinearRestriction(), P1toP1LinearProlongation(), 0, 7, 3, 3)
    gmg.solve(A, u, f, 7)
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ADD, ALL, CELL, DIRICHLET, EDGE, F32, F64, FACE, FREESLIP, INNER, NEUMANN, REPLACE,
                   VERTEX, HymiError)
from .dataexport import DataFormat, FixedSizeSQLDB, VTKOutput

__all__ = ["Context", "context", "MeshInfo", "PrimitiveStorage", "P1Function", "P1ConstantLaplaceOperator",
           "P1ConstantMassOperator", "P1toP1LinearRestriction", "P1toP1LinearProlongation", "SORSmoother",
           "CGSolver", "GeometricMultigridSolver", "GaussSeidelSmoother", "SymmetricSORSmoother", "CycleType",
           "FullMultigridSolver", "MixedPrecisionRefinementSolver", "HymiError", "INNER", "DIRICHLET", "NEUMANN", "FREESLIP", "ALL",
           "REPLACE", "ADD", "F64", "F32", "VERTEX", "EDGE", "FACE", "CELL", "DataFormat", "VTKOutput",
           "FixedSizeSQLDB"]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _releasable():
    """True while native handles may be destroyed; False during interpreter shutdown."""
    try:
        return _lib is not None and _lib.releasable()
    except Exception:
        return False


class Context:
    """A device + HIP stream (device=-1: host-only context for topology and stencil queries)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _lib.call("hymi_ctx_create", device, ctypes.byref(h))
        self.h = h
        self.device = device

    def synchronize(self):
        _lib.call("hymi_ctx_synchronize", self.h)

    def bytes_in_use(self):
        v = ctypes.c_int64()
        _lib.call("hymi_ctx_bytes_in_use", self.h, ctypes.byref(v))
        return v.value

    def __del__(self, _releasable=_releasable):
        if getattr(self, "h", None) and _releasable():
            _lib._lib.hymi_ctx_destroy(self.h)
            self.h = None


_contexts = {}


def context(device=0):
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


class MeshInfo:
    """Macro mesh description (MeshInfo::meshCuboid, MeshInfo::fromGmshFile)."""

    def __init__(self, kind, args):
        self.kind, self.args = kind, args

    @staticmethod
    def meshCuboid(lowerLeftFront, upperRightBack, nx, ny, nz):
        return MeshInfo("cuboid", (tuple(lowerLeftFront), tuple(upperRightBack), nx, ny, nz))

    @staticmethod
    def fromGmshFile(path):
        return MeshInfo("gmsh", (str(path),))

    @staticmethod
    def fromArrays(coords, cells):
        return MeshInfo("arrays", (np.ascontiguousarray(coords, np.float64), np.ascontiguousarray(cells, np.int32)))


class PrimitiveStorage:
    """Macro primitives + boundary flags (SetupPrimitiveStorage + setMeshBoundaryFlagsOnBoundary(1, 0, true))."""

    def __init__(self, meshInfo, ctx=None):
        self.ctx = ctx if ctx is not None else context(0)
        h = ctypes.c_void_p()
        if meshInfo.kind == "cuboid":
            lo, hi, nx, ny, nz = meshInfo.args
            lo = np.array(lo, np.float64)
            hi = np.array(hi, np.float64)
            _lib.call("hymi_mesh_create_cuboid", self.ctx.h, _dp(lo), _dp(hi), nx, ny, nz, ctypes.byref(h))
        elif meshInfo.kind == "gmsh":
            _lib.call("hymi_mesh_from_gmsh", self.ctx.h, meshInfo.args[0].encode(), ctypes.byref(h))
        else:
            X, C = meshInfo.args
            _lib.call("hymi_mesh_create", self.ctx.h, len(X), _dp(X), len(C), _ip(C), ctypes.byref(h))
        self.h = h

    def __del__(self, _releasable=_releasable):
        if getattr(self, "h", None) and _releasable():
            _lib._lib.hymi_mesh_destroy(self.h)
            self.h = None

    def counts(self):
        out = np.zeros(4, np.int32)
        _lib.call("hymi_mesh_counts", self.h, _ip(out))
        return tuple(int(v) for v in out)

    def primitive_type(self, cls, idx):
        v = ctypes.c_int()
        _lib.call("hymi_mesh_primitive_type", self.h, cls, idx, ctypes.byref(v))
        return v.value

    def primitive_coordinates(self, cls, idx):
        """Corner coordinates of one macro primitive in its local order (Primitive::getCoordinates)."""
        out = np.zeros((cls + 1, 3))
        _lib.call("hymi_mesh_primitive_coordinates", self.h, cls, idx, _dp(out))
        return out

    def owned_count(self, level):
        v = ctypes.c_int64()
        _lib.call("hymi_mesh_owned_count", self.h, level, ctypes.byref(v))
        return v.value

    def owned_coords(self, level):
        out = np.zeros((self.owned_count(level), 3))
        _lib.call("hymi_mesh_owned_coords", self.h, level, _dp(out))
        return out

    def owned_types(self, level):
        out = np.zeros(self.owned_count(level), np.int32)
        _lib.call("hymi_mesh_owned_types", self.h, level, _ip(out))
        return out

    def set_schedule(self, level, rows=0, waves=0, block=0, kind=-1):
        """Launch schedule of the macro-cell kernels on `level`: kind 0 row windows, 1 flat slab windows,
        2 LDS slab tiles, 3 columns marched along z (rows = slabs per unit, block = prefetch distance);
        -1 / 0 = level default."""
        _lib.call("hymi_mesh_set_schedule", self.h, level, kind, rows, waves, block)

    def use_column_schedules(self, maxLevel, minLevel=5, fusedSorFrom=7, fusedSorFp32From=7):
        """Select the column march (kind 3) with its measured best parameters on levels minLevel..maxLevel:
        16 slabs per unit and 4 waves per workgroup from level 7 up, 8 and 4 on level 6, 4 and 2 below;
        loads one slab ahead.  The SOR cell phase runs as one fused launch (set_cell_sor(level, 24)) for FP64
        from level fusedSorFrom up and for FP32 from level fusedSorFp32From up: where it beats four colour
        passes (the library defaults).  On other levels of the range it is switched off."""
        for L in range(minLevel, maxLevel + 1):
            rows, waves = (16, 4) if L >= 7 else (8, 4) if L == 6 else (4, 2)
            self.set_schedule(L, rows, waves, 1, 3)
            fp64, fp32 = L >= max(fusedSorFrom, 3), L >= max(fusedSorFp32From, 3)
            self.set_cell_sor(L, 24 if fp64 or fp32 else 0, fp64=fp64, fp32=fp32)

    def autotune(self, level, fp64=True, fp32=True, reps=5):
        """Measure the launch schedules of the macro-cell kernels and the fused SOR cell phase on `level` and keep
        the fastest (hymi_mesh_autotune; results are unchanged, bit for bit).  Returns the measured ms per
        candidate: schedules first, then (colour passes, fused) per selected precision."""
        report = np.zeros(64)
        count = ctypes.c_int()
        _lib.call("hymi_mesh_autotune", self.h, level, (1 if fp64 else 0) | (2 if fp32 else 0), reps, _dp(report),
                  len(report), ctypes.byref(count))
        return report[: count.value].tolist()

    def set_cell_sor(self, level, tile, fp64=True, fp32=False):
        """SOR cell phase on `level`: tile 0 = four colour passes, 24 = all four colours in one launch over
        24 x 24 tiles marched along z, for functions of the selected precisions with one right-hand side (same
        values bit for bit)."""
        _lib.call("hymi_mesh_set_cell_sor", self.h, level, int(tile), (1 if fp64 else 0) | (2 if fp32 else 0))

    def set_residual_restrict(self, level, mode):
        """Residual and restriction of a V-cycle from fine `level`: 1 = the macro-cell interiors in one pass, 0 = two
        passes, -1 = default (one pass from level 8 up).  Results are the same bit for bit."""
        _lib.call("hymi_mesh_set_residual_restrict", self.h, level, int(mode))

    def get_residual_restrict(self, level):
        """1 where `level` runs the fused residual and restriction."""
        out = ctypes.c_int()
        _lib.call("hymi_mesh_get_residual_restrict", self.h, level, ctypes.byref(out))
        return out.value

    def get_cell_sor(self, level):
        """Tile of the fused cell phase on `level` (0: colour passes)."""
        return self.get_cell_sor_setting(level)[0]

    def get_cell_sor_setting(self, level):
        """(tile, precisions): precisions bit 0 FP64, bit 1 FP32."""
        out = np.zeros(2, np.int32)
        _lib.call("hymi_mesh_get_cell_sor", self.h, level, _ip(out))
        return int(out[0]), int(out[1])

    def get_schedule(self, level):
        out = np.zeros(4, np.int32)
        _lib.call("hymi_mesh_get_schedule", self.h, level, _ip(out))
        return {"rows": int(out[0]), "waves": int(out[1]), "block": int(out[2]), "kind": int(out[3])}


class P1Function:
    """P1Function<real_t>(name, storage, minLevel, maxLevel); precision F64/F32, nrhs interleaved RHS."""

    def __init__(self, name, storage, minLevel, maxLevel, precision=F64, nrhs=1):
        self.name, self.storage = name, storage
        self.minLevel, self.maxLevel = minLevel, maxLevel
        self.precision, self.nrhs = precision, nrhs
        h = ctypes.c_void_p()
        _lib.call("hymi_function_create", storage.h, minLevel, maxLevel, precision, nrhs, ctypes.byref(h))
        self.h = h

    def __del__(self, _releasable=_releasable):
        if getattr(self, "h", None) and _releasable():
            _lib._lib.hymi_function_destroy(self.h)
            self.h = None

    # host transfers (owned order, see PrimitiveStorage.owned_coords)
    def download(self, level):
        out = np.zeros(self.storage.owned_count(level) * self.nrhs)
        _lib.call("hymi_function_download_owned", self.h, level, _dp(out))
        return out if self.nrhs == 1 else out.reshape(-1, self.nrhs)

    def upload(self, level, values):
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        if v.size != self.storage.owned_count(level) * self.nrhs:
            raise ValueError("wrong number of values")
        _lib.call("hymi_function_upload_owned", self.h, level, _dp(v))

    def download_column(self, level, column):
        """Column `column` of an interleaved function (owned order)."""
        out = np.zeros(self.storage.owned_count(level))
        _lib.call("hymi_function_download_owned_column", self.h, level, column, _dp(out))
        return out

    def select_columns(self, level, keep):
        """Zero every column not in `keep` (an iterable of column indices) on the whole level."""
        mask = 0
        for r in keep:
            mask |= 1 << int(r)
        _lib.call("hymi_function_select_columns", self.h, level, mask)

    def download_primitive(self, level, cls, idx):
        n = ctypes.c_int64()
        _lib.call("hymi_function_primitive_size", self.h, level, cls, idx, ctypes.byref(n))
        out = np.zeros(n.value)
        _lib.call("hymi_function_download_primitive", self.h, level, cls, idx, _dp(out))
        return out

    def download_micro_vertices(self, level, cls=CELL):
        """Values of every macro-cell (or macro-face) at all of its micro-vertices, closure copies synced:
        shape (primitives, points per primitive[, nrhs]), points in the primitive's index order."""
        w = (1 << level) + 1
        per = w * (w + 1) * (w + 2) // 6 if cls == CELL else w * (w + 1) // 2
        n = self.storage.counts()[cls]
        out = np.zeros(n * per * self.nrhs)
        _lib.call("hymi_function_download_micro_vertices", self.h, level, cls, _dp(out))
        return out.reshape((n, per) if self.nrhs == 1 else (n, per, self.nrhs))

    def upload_primitive(self, level, cls, idx, values):
        v = np.ascontiguousarray(values, np.float64)
        _lib.call("hymi_function_upload_primitive", self.h, level, cls, idx, _dp(v))

    def interpolate_modes(self, modes, level, flag=ALL):
        """interpolate(sum of Fourier modes) on the device; modes: (nModes, 4) or (nrhs, nModes, 4) rows of
        (amplitude, k, l, m) for a * sin(k pi x) sin(l pi y) sin(m pi z)."""
        m = np.ascontiguousarray(modes, dtype=np.float64)
        if m.ndim == 2:
            m = np.broadcast_to(m, (self.nrhs,) + m.shape).copy()
        if m.shape[0] != self.nrhs or m.shape[2] != 4:
            raise ValueError("modes must have shape (nModes, 4) or (nrhs, nModes, 4)")
        _lib.call("hymi_interpolate_modes", self.h, level, flag, m.shape[1], _dp(m))

    def interpolate(self, value, level, flag=ALL):
        """interpolate(constant | callable(xyz)->values, level, flag)."""
        if callable(value):
            X = self.storage.owned_coords(level)
            vals = np.asarray(value(X), np.float64)
            types = self.storage.owned_types(level)
            cur = self.download(level)
            sel = (types & flag) != 0
            if self.nrhs == 1:
                cur[sel] = vals[sel]
            else:
                cur[sel] = vals.reshape(len(X), -1)[sel]
            self.upload(level, cur)
        else:
            _lib.call("hymi_interpolate_const", self.h, level, flag, float(value))

    def _lincomb(self, name, scalars, functions, level, flag):
        c = np.ascontiguousarray(scalars, np.float64)
        arr = (ctypes.c_void_p * max(1, len(functions)))(*[f.h for f in functions])
        _lib.call(name, self.h, level, flag, len(functions), _dp(c), arr)

    def assign(self, scalars, functions, level, flag=ALL):
        self._lincomb("hymi_assign", scalars, functions, level, flag)

    def add(self, scalars, functions, level, flag=ALL):
        self._lincomb("hymi_add", scalars, functions, level, flag)

    def dotGlobal(self, other, level, flag=ALL):
        out = np.zeros(self.nrhs)
        _lib.call("hymi_dot", self.h, other.h, level, flag, _dp(out))
        return float(out[0]) if self.nrhs == 1 else out

    def copyFrom(self, other, level, flag=ALL):
        """Copy another function of any precision on one level (mixed-precision refinement)."""
        _lib.call("hymi_convert", self.h, other.h, level, flag)


class _P1ConstantOperator:
    _create = None

    def __init__(self, storage, minLevel, maxLevel):
        self.storage = storage
        self.minLevel, self.maxLevel = minLevel, maxLevel
        h = ctypes.c_void_p()
        _lib.call(self._create, storage.h, minLevel, maxLevel, ctypes.byref(h))
        self.h = h

    def __del__(self, _releasable=_releasable):
        if getattr(self, "h", None) and _releasable():
            _lib._lib.hymi_operator_destroy(self.h)
            self.h = None

    def stencil(self, level, cls, idx):
        n = ctypes.c_int()
        _lib.call("hymi_operator_stencil", self.h, level, cls, idx, None, ctypes.byref(n))
        out = np.zeros(n.value)
        _lib.call("hymi_operator_stencil", self.h, level, cls, idx, _dp(out), ctypes.byref(n))
        return out

    def apply(self, src, dst, level, flag, updateType=REPLACE):
        _lib.call("hymi_apply", self.h, src.h, dst.h, level, flag, updateType)

    def residual(self, x, b, r, level, flag):
        """r = b - A x on flag DoFs."""
        _lib.call("hymi_residual", self.h, x.h, b.h, r.h, level, flag)

    def smooth_sor(self, dst, rhs, relax, level, flag):
        _lib.call("hymi_smooth_sor", self.h, dst.h, rhs.h, float(relax), level, flag)

    def smooth_sor_backwards(self, dst, rhs, relax, level, flag):
        _lib.call("hymi_smooth_sor_backwards", self.h, dst.h, rhs.h, float(relax), level, flag)

    def smooth_gs(self, dst, rhs, level, flag):
        _lib.call("hymi_smooth_gs", self.h, dst.h, rhs.h, level, flag)

    def smooth_gs_backwards(self, dst, rhs, level, flag):
        _lib.call("hymi_smooth_gs_backwards", self.h, dst.h, rhs.h, level, flag)


class P1ConstantLaplaceOperator(_P1ConstantOperator):
    _create = "hymi_laplace_create"


class P1ConstantMassOperator(_P1ConstantOperator):
    _create = "hymi_mass_create"


class P1toP1LinearRestriction:
    def restrict(self, function, sourceLevel, flag):
        _lib.call("hymi_restrict", function.h, sourceLevel, flag)


class P1toP1LinearProlongation:
    def prolongate(self, function, sourceLevel, flag):
        """UpdateType Replace: the flag DoFs of sourceLevel + 1 become the prolongation of sourceLevel."""
        _lib.call("hymi_prolongate", function.h, sourceLevel, flag)

    def prolongateAndAdd(self, function, sourceLevel, flag):
        _lib.call("hymi_prolongate_add", function.h, sourceLevel, flag)


class SORSmoother:
    symmetric = False

    def __init__(self, relax=1.0):
        self.relax = relax

    def solve(self, A, x, b, level):
        A.smooth_sor(x, b, self.relax, level, INNER | NEUMANN)


class GaussSeidelSmoother(SORSmoother):
    """GaussSeidelSmoother (src/hyteg/solvers/GaussSeidelSmoother.hpp:38-50): smooth_gs, i.e. SOR with relax 1."""

    def __init__(self):
        super().__init__(1.0)

    def solve(self, A, x, b, level):
        A.smooth_gs(x, b, level, INNER | NEUMANN)


class SymmetricSORSmoother(SORSmoother):
    """SymmetricSORSmoother (src/hyteg/solvers/SymmetricSORSmoother.hpp:38-60): a forward SOR sweep, then a
    backward one."""
    symmetric = True

    def solve(self, A, x, b, level):
        A.smooth_sor(x, b, self.relax, level, INNER | NEUMANN)
        A.smooth_sor_backwards(x, b, self.relax, level, INNER | NEUMANN)


class CycleType:
    """GeometricMultigridSolver's CycleType (V: one coarse correction per level, W: two)."""
    VCYCLE = 0
    WCYCLE = 1


class CGSolver:
    """CGSolver(storage, minLevel, maxLevel, maxIter, tolerance) run on the device."""

    def __init__(self, storage, minLevel, maxLevel, maxIter=0, tolerance=1e-16):
        self.maxIter, self.tolerance = maxIter, tolerance
        self.minLevel = minLevel

    def solve(self, A, x, b, level):
        it = self.maxIter if self.maxIter > 0 else max(1, x.storage.owned_count(level))
        _lib.call("hymi_cg_solve", A.h, x.h, b.h, level, it, float(self.tolerance))


class GeometricMultigridSolver:
    """GeometricMultigridSolver(storage, smoother, coarseSolver, restriction, prolongation, minLevel, maxLevel,
    preSmoothSteps, postSmoothSteps, smoothIncrementOnCoarserGrids, cycleType) -- V- or W-cycles with SOR,
    Gauss-Seidel or symmetric SOR smoothing and a device CG coarse solver (GeometricMultigridSolver.hpp:42-235)."""

    def __init__(self, storage, smoother, coarseSolver, restrictionOperator, prolongationOperator, minLevel,
                 maxLevel, preSmoothSteps=3, postSmoothSteps=3, smoothIncrement=0, cycleType=CycleType.VCYCLE):
        if smoothIncrement < 0:
            raise ValueError("smoothIncrement must be >= 0")
        if cycleType not in (CycleType.VCYCLE, CycleType.WCYCLE):
            raise ValueError("cycleType must be CycleType.VCYCLE or CycleType.WCYCLE")
        self.storage = storage
        self.smoother, self.coarse = smoother, coarseSolver
        self.minLevel, self.maxLevel = minLevel, maxLevel
        self.nu1, self.nu2 = preSmoothSteps, postSmoothSteps
        self.smoothIncrement = smoothIncrement
        self.cycleType = cycleType
        self._h = {}

    def setSmoothingSteps(self, preSmoothingSteps, postSmoothingSteps, smoothIncrement=0):
        """GeometricMultigridSolver::setSmoothingSteps (GeometricMultigridSolver.hpp:120-125)."""
        self.nu1, self.nu2, self.smoothIncrement = preSmoothingSteps, postSmoothingSteps, smoothIncrement
        for h, _ in self._h.values():
            _lib.call("hymi_gmg_set_smoothing", h, self.nu1, self.nu2, self.smoothIncrement)

    def _handle(self, A, x):
        key = (id(A), x.precision, x.nrhs)
        if key not in self._h:
            h = ctypes.c_void_p()
            cg_it = getattr(self.coarse, "maxIter", 0)
            cg_tol = getattr(self.coarse, "tolerance", 1e-16)
            _lib.call("hymi_gmg_create", A.h, x.h, self.minLevel, self.maxLevel, self.nu1, self.nu2,
                      float(self.smoother.relax), int(cg_it), float(cg_tol), ctypes.byref(h))
            self._h[key] = (h, A)
            if self.smoothIncrement:
                _lib.call("hymi_gmg_set_smoothing", h, self.nu1, self.nu2, self.smoothIncrement)
            if self.cycleType != CycleType.VCYCLE or getattr(self.smoother, "symmetric", False):
                _lib.call("hymi_gmg_set_variant", h, int(self.cycleType),
                          int(getattr(self.smoother, "symmetric", False)))
        return self._h[key][0]

    def solve(self, A, x, b, level):
        _lib.call("hymi_gmg_vcycle", self._handle(A, x), x.h, b.h, level)

    def solve_graph(self, A, x, b, level, count=1):
        """Replay a hipGraph-captured V-cycle `count` times."""
        _lib.call("hymi_gmg_vcycle_graph", self._handle(A, x), x.h, b.h, level, count)

    def __del__(self, _releasable=_releasable):
        if not _releasable():
            return
        for h, _ in getattr(self, "_h", {}).values():
            _lib._lib.hymi_gmg_destroy(h)
        self._h = {}


class FullMultigridSolver:
    """FullMultigridSolver(storage, gmgSolver, fmgProlongation, minLevel, maxLevel, cyclesPerLevel)
    (src/hyteg/solvers/FullMultigridSolver.hpp:31-68): cycles of gmgSolver on every level from minLevel up,
    the iterate prolongated (Replace) to the next level in between.  b must hold the right-hand side on
    every level; the cycles run from captured graphs."""

    def __init__(self, storage, gmgSolver, fmgProlongation, minLevel, maxLevel, cyclesPerLevel=1):
        if not isinstance(fmgProlongation, P1toP1LinearProlongation):
            raise TypeError("only the linear prolongation is available")
        if minLevel != gmgSolver.minLevel or maxLevel != gmgSolver.maxLevel:
            raise ValueError("the FMG levels must be the multigrid solver's levels")
        self.gmg, self.minLevel, self.maxLevel = gmgSolver, minLevel, maxLevel
        self.cyclesPerLevel = cyclesPerLevel

    def solve(self, A, x, b, level):
        _lib.call("hymi_fmg_solve", self.gmg._handle(A, x), x.h, b.h, level, self.cyclesPerLevel)


class MixedPrecisionRefinementSolver:
    """Iterative refinement: residual, convergence check and update in the precision of x (FP64), the
    correction from `cyclesPerOuter` V-cycles of `innerSolver` run in `innerPrecision` (FP32).

    solve(A, x, b, level) iterates until max_rhs ||b - A x|| / ||b - A x_0|| <= tolerance or maxOuter
    correction steps; it returns the final relative residual per right-hand side and records
    `outerSteps` and `history` (relative residual after 0, 1, ... steps)."""

    def __init__(self, storage, innerSolver, tolerance=1e-10, maxOuter=50, cyclesPerOuter=2, innerPrecision=F32):
        self.storage, self.inner = storage, innerSolver
        self.tolerance, self.maxOuter, self.cyclesPerOuter = tolerance, maxOuter, cyclesPerOuter
        self.innerPrecision = innerPrecision
        self._templates = {}
        self.outerSteps, self.history = 0, []

    def _innerHandle(self, A, nrhs):
        if nrhs not in self._templates:
            # the solver takes only precision and nrhs from its template; one coarse level suffices
            self._templates[nrhs] = P1Function("refinement-template", self.storage, self.inner.minLevel,
                                               self.inner.minLevel, precision=self.innerPrecision, nrhs=nrhs)
        return self.inner._handle(A, self._templates[nrhs])

    def step(self, A, x, b, level, cycles=None, columns=None):
        """One refinement step (hymi_refine_step): with cycles = 0 only ||b - A x|| per column; else the
        correction from the residual the previous step left, added to the given columns (default all), and
        the new residual norms.  Returns the norms (one per column)."""
        h = self._innerHandle(A, x.nrhs)
        cycles = self.cyclesPerOuter if cycles is None else cycles
        mask = 0
        for r in (range(x.nrhs) if columns is None else columns):
            mask |= 1 << int(r)
        out = np.zeros(x.nrhs)
        _lib.call("hymi_refine_step", h, x.h, b.h, level, int(cycles), mask, _dp(out))
        return out

    def solve(self, A, x, b, level):
        h = self._innerHandle(A, x.nrhs)
        outer = ctypes.c_int()
        rel = np.zeros(x.nrhs)
        hist = np.zeros(self.maxOuter + 1)
        _lib.call("hymi_solve_refined", h, x.h, b.h, level, float(self.tolerance), int(self.maxOuter),
                  int(self.cyclesPerOuter), ctypes.byref(outer), _dp(rel), _dp(hist))
        self.outerSteps = outer.value
        self.history = hist[: outer.value + 1].tolist()
        return rel
