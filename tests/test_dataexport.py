"""CPU tests of the VTK and SQLite output (hyteg_amd.dataexport; src/hyteg/dataexport/ in the reference).

The micro-mesh is pinned by what the geometry fixes rather than by a restated table: the micro-tetrahedra
(micro-triangles) tile their macro-cell (macro-face) conformingly -- equal volumes (areas) summing to the macro
one, every interior micro-face (micro-edge) shared by exactly two elements, every boundary one lying on the
macro boundary -- and the point order is recovered from the coordinates and checked against the reference's
index formula.  A wrong vertex offset in one of the six micro-cell types breaks the volume or face-sharing
checks.  The files are parsed back with xml.etree and checked against the same arrays; the function values
come from a stand-in object here (the GPU tests write real functions)."""
import base64
import os
import sqlite3
import xml.etree.ElementTree as ET
from collections import Counter

import numpy as np
import pytest

from helpers import mesh_pair

from hyteg_amd import dataexport as D

SKEWED = np.array([[0.3, -0.2, 0.1], [1.7, 0.1, -0.3], [0.2, 1.1, 0.4], [0.5, 0.6, 2.0]])
UNIT = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def _vol(p):
    return np.linalg.det(np.stack([p[..., 1, :] - p[..., 0, :], p[..., 2, :] - p[..., 0, :],
                                   p[..., 3, :] - p[..., 0, :]], axis=-2)) / 6.0


@pytest.mark.parametrize("level", [1, 2, 3, 4])
@pytest.mark.parametrize("corners", [UNIT, SKEWED], ids=["unit", "skewed"])
def test_micro_cells_tile_the_macro_cell(level, corners):
    w = (1 << level) + 1
    P = D.micro_vertex_points(corners, level)
    C = D.micro_cell_connectivity(level)
    assert P.shape == (w * (w + 1) * (w + 2) // 6, 3)
    assert C.shape == (8 ** level, 4)
    assert C.min() >= 0 and C.max() < len(P)
    assert all(len(set(row)) == 4 for row in C.tolist())
    V = np.abs(_vol(P[C]))
    macro = abs(_vol(corners))
    assert np.allclose(V, macro / 8 ** level, rtol=1e-10, atol=0)
    assert np.isclose(V.sum(), macro, rtol=1e-12)
    faces = Counter()
    for a, b, c, d in C.tolist():
        for t in ((a, b, c), (a, b, d), (a, c, d), (b, c, d)):
            faces[tuple(sorted(t))] += 1
    assert set(faces.values()) <= {1, 2}
    boundary = [f for f, k in faces.items() if k == 1]
    assert len(boundary) == 4 * 4 ** level
    # boundary micro-faces lie on a macro face: one barycentric coordinate is zero at all three points
    bary = np.linalg.solve(np.stack([corners[1] - corners[0], corners[2] - corners[0], corners[3] - corners[0]]).T,
                           (P - corners[0]).T).T
    lam = np.column_stack([1 - bary.sum(1), bary])
    for f in boundary:
        assert np.any(np.all(np.abs(lam[list(f)]) < 1e-12, axis=0)), f


@pytest.mark.parametrize("level", [0, 1, 3, 5])
def test_micro_vertex_order_is_the_reference_index(level):
    """Point k sits at micro-vertex (x, y, z) with index(x, y, z) = k, x running fastest; point 0 is corner 0
    exactly, the other corners up to the rounding of c0 + (c_k - c0)."""
    w = (1 << level) + 1
    P = D.micro_vertex_points(SKEWED, level)
    M = np.stack([SKEWED[1] - SKEWED[0], SKEWED[2] - SKEWED[0], SKEWED[3] - SKEWED[0]]).T
    xyz = np.rint(np.linalg.solve(M, (P - SKEWED[0]).T).T * (w - 1)).astype(np.int64)
    assert np.all(xyz >= 0) and np.all(xyz.sum(1) <= w - 1)
    k = D._tet_index(w, xyz[:, 0], xyz[:, 1], xyz[:, 2])
    assert np.array_equal(k, np.arange(len(P)))
    # brute-force order of the reference's iterator: z outer, y, x inner
    order = [(x, y, z) for z in range(w) for y in range(w - z) for x in range(w - z - y)]
    assert [tuple(t) for t in xyz.tolist()] == order
    for corner, (x, y, z) in zip(SKEWED, [(0, 0, 0), (w - 1, 0, 0), (0, w - 1, 0), (0, 0, w - 1)]):
        assert np.allclose(P[D._tet_index(w, x, y, z)], corner, rtol=0, atol=4e-16)
    assert np.array_equal(P[0], SKEWED[0])


@pytest.mark.parametrize("level", [0, 1, 2, 4])
def test_micro_faces_tile_the_macro_face(level):
    c = SKEWED[:3]
    w = (1 << level) + 1
    P = D.micro_face_points(c, level)
    T = D.micro_face_connectivity(level)
    assert P.shape == (w * (w + 1) // 2, 3)
    assert T.shape == (4 ** level, 3)
    assert T.min() >= 0 and T.max() < len(P)
    # point (x, y) of row y: c0 + x d0 + y d2, rows along c2
    pts = [(x, y) for y in range(w) for x in range(w - y)]
    d0, d2 = (c[1] - c[0]) / (w - 1), (c[2] - c[0]) / (w - 1)
    want = np.array([c[0] + x * d0 + y * d2 for x, y in pts])
    assert np.allclose(P, want, rtol=0, atol=1e-14)
    area = 0.5 * np.linalg.norm(np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]]), axis=1)
    macro = 0.5 * np.linalg.norm(np.cross(c[1] - c[0], c[2] - c[0]))
    assert np.allclose(area, macro / 4 ** level, rtol=1e-10)
    edges = Counter()
    for a, b, d in T.tolist():
        for e in ((a, b), (a, d), (b, d)):
            edges[tuple(sorted(e))] += 1
    assert set(edges.values()) <= {1, 2}
    assert sum(1 for k in edges.values() if k == 1) == 3 * 2 ** level


def test_stream_encodings():
    a = np.array([0.1, -2.5e-7, 3.0, 1e300])
    assert D._Stream.text(a, np.float64, D.DataFormat.ASCII) == \
        "1.000000e-01\n-2.500000e-07\n3.000000e+00\n1.000000e+300\n"
    assert D._Stream.text(np.array([4, 8, 12]), np.uint32, D.DataFormat.ASCII) == "4\n8\n12\n"
    assert D._Stream.text(np.zeros(0), np.float64, D.DataFormat.ASCII) == ""
    raw = base64.b64decode(D._Stream.text(a, np.float64, D.DataFormat.BINARY))
    assert np.frombuffer(raw[:4], "<u4")[0] == 32
    assert np.array_equal(np.frombuffer(raw[4:], "<f8"), a)
    raw = base64.b64decode(D._Stream.text(np.array([10, 10]), np.uint16, D.DataFormat.BINARY))
    assert raw == np.uint32(4).tobytes() + np.array([10, 10], "<u2").tobytes()
    with pytest.raises(ValueError):
        D._Stream.text(a, np.float64, "hex")


class _StandIn:
    """A function-shaped object whose micro-vertex values are the x coordinate of each point (host test)."""

    def __init__(self, name, storage, level, nrhs=1):
        self.name, self.storage, self.minLevel, self.maxLevel, self.nrhs = name, storage, 0, level, nrhs

    def download_micro_vertices(self, level, cls):
        n = self.storage.counts()[cls]
        pts = D.micro_vertex_points if cls == 3 else D.micro_face_points
        x = np.stack([pts(self.storage.primitive_coordinates(cls, p), level)[:, 0] for p in range(n)])
        return x if self.nrhs == 1 else np.stack([x * (k + 1) for k in range(self.nrhs)], axis=-1)


def _arrays(piece):
    def get(el):
        fmt, typ = el.get("format"), el.get("type")
        dt = {"Float64": "<f8", "Int32": "<i4", "UInt32": "<u4", "UInt16": "<u2"}[typ]
        if fmt == "ascii":
            return np.array((el.text or "").split(), dtype=np.float64).astype(dt)
        raw = base64.b64decode(el.text.strip())
        n = np.frombuffer(raw[:4], "<u4")[0]
        assert n == len(raw) - 4
        return np.frombuffer(raw[4:], dt)
    out = {"points": get(piece.find("Points/DataArray")).reshape(-1, 3)}
    for el in piece.find("Cells"):
        out[el.get("Name")] = get(el)
    for el in piece.find("PointData"):
        out["data:" + el.get("Name")] = get(el)
    return out


@pytest.mark.parametrize("fmt", [D.DataFormat.ASCII, D.DataFormat.BINARY])
@pytest.mark.parametrize("two_d", [False, True])
def test_vtu_document(tmp_path, host_ctx, oracle_lib, fmt, two_d):
    """cube_6el.msh at level 2, the reference's VTKOutputTest3D case, with a stand-in function."""
    om, S = mesh_pair(("gmsh", "cube_6el.msh"), host_ctx)
    L = 2
    out = D.VTKOutput(str(tmp_path), "VTKOutputTest3D", S)
    assert out.write(L) is None                      # nothing registered: no file
    out.setVTKDataFormat(fmt)
    if two_d:
        out.set2D()
    out.add(_StandIn("P1 scalar function 1", S, L))
    out.add(_StandIn("batched", S, L, nrhs=2))
    path = out.write(L, timestep=0)
    assert path == os.path.join(str(tmp_path), "VTKOutputTest3D_VertexDoF_level2_ts0.vtu")
    root = ET.parse(path).getroot()
    assert root.tag == "VTKFile" and root.get("type") == "UnstructuredGrid"
    piece = root.find("UnstructuredGrid/Piece")
    nV, nE, nF, nC = S.counts()
    w = (1 << L) + 1
    n, per, ncell = (nF, w * (w + 1) // 2, 4 ** L) if two_d else (nC, w * (w + 1) * (w + 2) // 6, 8 ** L)
    assert int(piece.get("NumberOfPoints")) == n * per
    assert int(piece.get("NumberOfCells")) == n * ncell
    A = _arrays(piece)
    nv = 3 if two_d else 4
    assert np.array_equal(A["offsets"], nv * np.arange(1, n * ncell + 1))
    assert np.all(A["types"] == (5 if two_d else 10))
    conn = A["connectivity"].reshape(-1, nv)
    local = D.micro_face_connectivity(L) if two_d else D.micro_cell_connectivity(L)
    assert np.array_equal(conn, (np.arange(n)[:, None, None] * per + local[None]).reshape(-1, nv))
    # corners agree with the oracle's macro mesh (vertex coordinates = level-0 owned coordinates)
    X0 = om.owned_coords(0)
    for p in range(n):
        ids = om.face_vertices(p) if two_d else om.cell_vertices(p)
        assert np.array_equal(S.primitive_coordinates(2 if two_d else 3, p), X0[ids])
    tol = 5e-7 if fmt == D.DataFormat.ASCII else 0.0
    pts = A["points"]
    assert np.allclose(A["data:P1 scalar function 1"], pts[:, 0], rtol=tol, atol=tol)
    assert np.allclose(A["data:batched_1"], 2 * pts[:, 0], rtol=tol, atol=tol)
    # every macro-cell's micro-tetrahedra are positively or negatively oriented as a whole and fill the cube
    if not two_d:
        vol = np.abs(_vol(pts[conn])).sum()
        assert np.isclose(vol, 1.0, rtol=1e-5 if tol else 1e-12)


def test_write_frequency_and_argument_errors(tmp_path, host_ctx, oracle_lib):
    om, S = mesh_pair(("gmsh", "tet_1el.msh"), host_ctx)
    out = D.VTKOutput(str(tmp_path), "f", S, writeFrequency=3)
    out.add(_StandIn("u", S, 2))
    assert out.write(1, timestep=1) is None
    assert out.write(1, timestep=3).endswith("f_VertexDoF_level1_ts3.vtu")
    with pytest.raises(ValueError):
        out.write(3, timestep=0)                     # level above the function's maxLevel
    assert D.VTKOutput(str(tmp_path), "g", S, writeFrequency=0).write(1) is None
    with pytest.raises(TypeError):
        out.add(object())
    with pytest.raises(ValueError):
        out.setVTKDataFormat("hex")


def test_fixed_size_sql_db(tmp_path):
    db = str(tmp_path / "runs.db")
    q = D.FixedSizeSQLDB(db)
    q.setConstantEntry("level", 7)
    q.setConstantEntry("mesh", "cube_48el")
    q.setConstantEntry("fp64", True)
    with pytest.raises(RuntimeError):
        q.setConstantEntry("level", 8)               # overwriting a constant is refused
    q.setVariableEntry("cycle", 0)
    q.setVariableEntry("residual", 1.5)
    with pytest.raises(RuntimeError):
        q.setVariableEntry("cycle", 1)
    assert q.writeRowOnRoot() == 1
    with pytest.raises(RuntimeError):
        q.writeRowOnRoot()                           # variables must be set again for every row
    q.setVariableEntry("cycle", 1)
    with pytest.raises(RuntimeError):
        q.setVariableEntry("new", 2.0)               # columns are fixed after the first row
    with pytest.raises(RuntimeError):
        q.setConstantEntry("other", 1)
    q.setVariableEntry("residual", 0.25)
    assert q.writeRowOnRoot() == 2
    con = sqlite3.connect(db)
    rows = con.execute('SELECT runId, level, mesh, fp64, cycle, residual, timestamp FROM runs ORDER BY runId').fetchall()
    types = {r[1]: r[2] for r in con.execute("PRAGMA table_info(runs)")}
    con.close()
    assert [r[:6] for r in rows] == [(1, 7, "cube_48el", 1, 0, 1.5), (2, 7, "cube_48el", 1, 1, 0.25)]
    assert all(r[6] for r in rows)
    assert types["level"] == "INTEGER" and types["residual"] == "DOUBLE" and types["mesh"] == "TEXT"
    # a second database object on the same file appends rows and adds its new columns
    q2 = D.FixedSizeSQLDB(db, allowOverwritingValues=True)
    q2.setConstantEntry("level", 6)
    q2.setConstantEntry("level", 5)
    q2.setVariableEntry("gdofs", 3.09)
    assert q2.writeRowOnRoot() == 3
    con = sqlite3.connect(db)
    assert con.execute("SELECT level, gdofs, cycle FROM runs WHERE runId = 3").fetchone() == (5, 3.09, None)
    con.close()
    with pytest.raises(TypeError):
        D.FixedSizeSQLDB(db).setVariableEntry("bad", [1, 2])
