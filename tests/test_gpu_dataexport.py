"""GPU tests of the VTK output of device-resident P1 functions (VTKOutput / VTKP1Writer, src/hyteg/dataexport/).

The written point data must be the oracle's primitive memory for the same function after it has communicated
every direction (the reference writes after VTKOutput::syncAllFunctions): bitwise for uploaded values, to
1e-12 relative after SOR sweeps (FP64, different summation orders only).  The reference's own VTKOutputTest3D
case (cube_6el.msh, level 2, coordinate functions) is checked through data = coordinate at every point."""
import base64
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say

pytestmark = pytest.mark.gpu


def _arrays(path):
    piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")

    def get(el):
        dt = {"Float64": "<f8", "Int32": "<i4", "UInt32": "<u4", "UInt16": "<u2"}[el.get("type")]
        if el.get("format") == "ascii":
            return np.array((el.text or "").split(), dtype=np.float64).astype(dt)
        raw = base64.b64decode(el.text.strip())
        return np.frombuffer(raw[4:], dt)
    out = {"points": get(piece.find("Points/DataArray")).reshape(-1, 3)}
    for el in piece.find("PointData"):
        out[el.get("Name")] = get(el)
    return out


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_coordinate_functions_cube_6el(tmp_path, gpu_ctx, oracle_lib, fmt):
    import hyteg_amd as H
    from hyteg_amd import dataexport as D
    L = 2
    say(f"[vtk] cube_6el level={L} coordinate functions format={fmt}")
    om, S = mesh_pair(("gmsh", "cube_6el.msh"), gpu_ctx)
    funcs = []
    for d, name in enumerate(("P1 scalar function 1", "P1 scalar function 2", "P1 scalar function 3")):
        f = H.P1Function(name, S, L, L)
        f.interpolate(lambda X, d=d: X[:, d], L, H.ALL)
        funcs.append(f)
    out = D.VTKOutput(str(tmp_path), "VTKOutputTest3D", S)
    out.setVTKDataFormat(fmt)
    for f in funcs:
        out.add(f)
    A = _arrays(out.write(L))
    tol = 5e-7 if fmt == "ascii" else 1e-15
    for d, f in enumerate(funcs):
        assert np.allclose(A[f.name], A["points"][:, d], rtol=tol, atol=tol), f.name


@pytest.mark.parametrize("kind", [("gmsh", "cube_6el.msh"), ("cube", 2), ("gmsh", "regular_octahedron_8el.msh")])
@pytest.mark.parametrize("level", [2, 4])
def test_point_data_is_the_oracle_primitive_memory(tmp_path, gpu_ctx, oracle_lib, kind, level):
    import hyteg_amd as H
    from hyteg_amd import dataexport as D
    O = oracle_lib
    say(f"[vtk] mesh={kind} level={level} uploaded and smoothed functions, 3D and 2D")
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(5 + level)
    x, b = rng.standard_normal((2, om.owned_count(level)))
    u, s, f = (H.P1Function(n, S, 0, level) for n in ("u", "s", "f"))
    u.upload(level, x)
    s.upload(level, x)
    f.upload(level, b)
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    ou, os_, of = (O.Function(om, 0, level) for _ in range(3))
    ou.set(level, x)
    os_.set(level, x)
    of.set(level, b)
    OA = O.Operator(om, 0, level)
    for _ in range(2):
        A.smooth_sor(s, f, 0.9, level, H.INNER)
        OA.smooth_sor(os_, of, level, omega=0.9, flag=O.INNER, order=1)
    for g in (ou, os_):
        for d in range(6):
            g.communicate(level, d)
    nV, nE, nF, nC = om.counts()
    w = (1 << level) + 1
    for two_d in (False, True):
        out = D.VTKOutput(str(tmp_path), "parity", S)
        out.setVTKDataFormat(D.DataFormat.BINARY)
        if two_d:
            out.set2D()
        out.add(u)
        out.add(s)
        V = _arrays(out.write(level, timestep=int(two_d)))
        cls, n, per = (H.FACE, nF, w * (w + 1) // 2) if two_d else (H.CELL, nC, w * (w + 1) * (w + 2) // 6)
        ref_u = np.concatenate([ou.prim(level, cls, p)[:per] for p in range(n)])
        ref_s = np.concatenate([os_.prim(level, cls, p)[:per] for p in range(n)])
        assert np.array_equal(V["u"], ref_u), ("uploaded", two_d)
        assert rel_err(V["s"], ref_s) <= 1e-12, ("smoothed", two_d)


def test_batched_and_fp32_functions(tmp_path, gpu_ctx, oracle_lib):
    """A function with nrhs = 32 writes 32 arrays, each equal to the single-RHS export of its column; an FP32
    function writes its float values widened to Float64."""
    import hyteg_amd as H
    from hyteg_amd import dataexport as D
    L = 3
    say(f"[vtk] cube n=1 level={L} nrhs=32 and FP32")
    om, S = mesh_pair(("cube", 1), gpu_ctx)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((om.owned_count(L), 32))
    ub = H.P1Function("batch", S, L, L, nrhs=32)
    ub.upload(L, X)
    u32 = H.P1Function("single", S, L, L, precision=H.F32)
    u32.upload(L, X[:, 0])
    out = D.VTKOutput(str(tmp_path), "batched", S)
    out.setVTKDataFormat(D.DataFormat.BINARY)
    out.add(ub)
    out.add(u32)
    V = _arrays(out.write(L))
    one = H.P1Function("one", S, L, L)
    assert sorted(k for k in V if k.startswith("batch_")) == sorted("batch_%d" % k for k in range(32))
    for k in (0, 7, 31):
        one.upload(L, X[:, k])
        assert np.array_equal(V["batch_%d" % k], one.download_micro_vertices(L, H.CELL).reshape(-1)), k
    one.upload(L, X[:, 0].astype(np.float32).astype(np.float64))
    assert np.array_equal(V["single"], one.download_micro_vertices(L, H.CELL).reshape(-1))
