"""VTK (.vtu) and SQLite output of P1 functions.

Host-side mirror of the reference's src/hyteg/dataexport/: VTKOutput (VTKOutput.hpp:72-147,
VTKOutput.cpp:129-250), the P1 writer (VTKP1Writer.cpp:34-150), the micro-mesh writer (VTKMeshWriter.cpp:35-92,
230-320, 322-459) and FixedSizeSQLDB (SQL.hpp:37-134, SQL.cpp:24-59).

Function values leave the device in one transfer per function and level (hymi_function_download_micro_vertices,
which first refreshes the closure copies as VTKOutput::syncAllFunctions does); points and connectivity are
generated with numpy from the macro-primitive corners (hymi_mesh_primitive_coordinates).  The files hold what the
reference writes for the same storage: per macro-cell (3D) or macro-face (2D) every micro-vertex in that
primitive's index order, the micro-tetrahedra (triangles) in the reference's six-type (two-type) order, and one
DataArray per registered function.
"""
import base64
import os
import sqlite3

import numpy as np

from ._lib import CELL, FACE

__all__ = ["DataFormat", "VTKOutput", "FixedSizeSQLDB", "micro_vertex_points", "micro_cell_connectivity",
           "micro_face_points", "micro_face_connectivity"]


class DataFormat:
    """vtk::DataFormat (src/hyteg/dataexport/VTKHelpers.hpp:26-30)."""
    ASCII = "ascii"
    BINARY = "binary"


_VTK_TYPE = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int32): "Int32",
             np.dtype(np.int64): "Int64", np.dtype(np.uint32): "UInt32", np.dtype(np.uint16): "UInt16"}

# VTK cell type ids: VTK_TRIANGLE, VTK_TETRA
_VTK_TRIANGLE, _VTK_TETRA = 5, 10

# micro-tetrahedra of a macro-cell: (iterator width below the vertex width w, 4 vertex offsets from the base
# index), in the order VTKMeshWriter::writeCells3D emits them (white up, blue up, green up, white down, blue down,
# green down; vertex offsets of celldof::macrocell::getMicroVerticesFromMicroCell, CellDoFIndexing.hpp:129-166)
_MICRO_CELL_TYPES = (
    (1, ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))),
    (2, ((1, 0, 0), (1, 1, 0), (0, 1, 0), (1, 0, 1))),
    (2, ((1, 0, 0), (0, 1, 0), (1, 0, 1), (0, 0, 1))),
    (3, ((1, 1, 0), (1, 1, 1), (0, 1, 1), (1, 0, 1))),
    (2, ((1, 0, 1), (0, 1, 1), (0, 0, 1), (0, 1, 0))),
    (2, ((0, 1, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1))),
)


def _tet_points(n):
    """All (x, y, z) with x + y + z <= n - 1 in the order indexing::CellIterator(n) visits them (x fastest, then
    y, then z; MacroCellIndexing.cpp:135-162), as three int64 arrays."""
    if n <= 0:
        e = np.zeros(0, np.int64)
        return e, e, e
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    keep = (x + y + z) <= n - 1
    return x[keep].astype(np.int64), y[keep].astype(np.int64), z[keep].astype(np.int64)


def _tet_size(w):
    return w * (w + 1) * (w + 2) // 6


def _tet_index(w, x, y, z):
    """Position of micro-vertex (x, y, z) in a macro-cell of width w (vertexdof::macrocell::index; the
    calcVTKPointArrayPosition of VTKMeshWriter.cpp:335-342)."""
    r = w - z
    return (_tet_size(w) - r * (r + 1) * (r + 2) // 6) + (y * (r + 1) - y * (y + 1) // 2) + x


def micro_vertex_points(corners, level):
    """Micro-vertex coordinates of one macro-cell, (tetSize(w), 3), in index order:
    c0 + (c1 - c0) h x + (c2 - c0) h y + (c3 - c0) h z with h = 1 / 2^level, evaluated in the reference's
    operation order (vertexdof::macrocell::coordinateFromIndex, VertexDoFMacroCell.hpp:69-76)."""
    c = np.asarray(corners, np.float64)
    w = (1 << level) + 1
    x, y, z = _tet_points(w)
    h = 1.0 / (w - 1)
    xs, ys, zs = (c[1] - c[0]) * h, (c[2] - c[0]) * h, (c[3] - c[0]) * h
    return c[0] + xs * x[:, None].astype(np.float64) + ys * y[:, None].astype(np.float64) + \
        zs * z[:, None].astype(np.float64)


def micro_cell_connectivity(level):
    """Point indices (within one macro-cell) of its 8^level micro-tetrahedra, (8^level, 4) int64, in the
    order of VTKMeshWriter::writeCells3D (VTKMeshWriter.cpp:347-409)."""
    w = (1 << level) + 1
    parts = []
    for shrink, offs in _MICRO_CELL_TYPES:
        x, y, z = _tet_points(w - shrink)
        parts.append(np.stack([_tet_index(w, x + dx, y + dy, z + dz) for dx, dy, dz in offs], axis=1))
    return np.concatenate(parts, axis=0)


def micro_face_points(corners, level):
    """Micro-vertex coordinates of one macro-face, (w(w+1)/2, 3), rows along c2 and points along c1 within a
    row, accumulated step by step as VTKMeshWriter::writePointsForMicroVertices does (VTKMeshWriter.cpp:43-74):
    row start c0 + i d2, then repeated additions of d0, with d = (c_k - c0) / (w - 1)."""
    c = np.asarray(corners, np.float64)
    w = (1 << level) + 1
    d0 = (c[1] - c[0]) / (float(w) - 1.0)
    d2 = (c[2] - c[0]) / (float(w) - 1.0)
    rows = []
    for i in range(w):
        steps = np.empty((w - i, 3))
        steps[0] = c[0] + float(i) * d2
        steps[1:] = d0
        rows.append(np.cumsum(steps, axis=0))
    return np.concatenate(rows, axis=0)


def micro_face_connectivity(level):
    """Point indices (within one macro-face) of its 4^level micro-triangles, (4^level, 3) int64, in the order
    of VTKMeshWriter::writeCells2D (VTKMeshWriter.cpp:245-271)."""
    w = (1 << level) + 1
    tris = []
    start = 0
    for i in range(w - 1):
        inner = w - 1 - i
        j = np.arange(inner - 1, dtype=np.int64)
        s = start + j
        pair = np.stack([np.stack([s, s + 1, s + inner + 1], axis=1),
                         np.stack([s + 1, s + inner + 2, s + inner + 1], axis=1)], axis=1).reshape(-1, 3)
        last = start + inner - 1
        tris.append(pair)
        tris.append(np.array([[last, last + 1, last + inner + 1]], np.int64))
        start += inner + 1
    return np.concatenate(tris, axis=0)


class _Stream:
    """VTKOutput::VTKStreamWriter (VTKOutput.hpp:141-178): ASCII writes one value per line (floating point in
    std::scientific's default form, %.6e); BINARY writes the base64 encoding of a UInt32 byte count followed by
    the raw little-endian values (the VTK XML format's inline binary encoding)."""

    @staticmethod
    def text(values, dtype, fmt):
        a = np.ascontiguousarray(values, dtype=dtype).reshape(-1)
        if fmt == DataFormat.ASCII:
            if a.size == 0:
                return ""
            form = "%.6e\n" if a.dtype.kind == "f" else "%d\n"
            return (form * a.size) % tuple(a.tolist())
        if fmt == DataFormat.BINARY:
            raw = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
            if len(raw) >= 1 << 32:
                raise ValueError("a binary VTK data array is limited to 4 GiB by its UInt32 header")
            return base64.b64encode(np.uint32(len(raw)).tobytes() + raw).decode("ascii")
        raise ValueError("VTK format not supported.")


def _open_data_element(dtype, name, components, fmt):
    """vtk::openDataElement (VTKHelpers.hpp:78-109)."""
    s = '<DataArray type="%s"' % _VTK_TYPE[np.dtype(dtype)]
    if name:
        s += ' Name="%s"' % name
    if components > 0:
        s += ' NumberOfComponents="%d"' % components
    if fmt not in (DataFormat.ASCII, DataFormat.BINARY):
        raise ValueError("VTK format not supported.")
    return s + ' format="%s">\n' % fmt


def _data_array(parts, values, dtype, name, components, fmt):
    parts.append(_open_data_element(dtype, name, components, fmt))
    parts.append(_Stream.text(values, dtype, fmt))
    parts.append("\n</DataArray>\n")


class VTKOutput:
    """VTKOutput(dir, filename, storage, writeFrequency=1) for P1 functions
    (src/hyteg/dataexport/VTKOutput.hpp:72-147, VTKOutput.cpp:26-41).

    write(level, timestep) writes "<dir>/<filename>_VertexDoF_level<L>_ts<timestep>.vtu" when writeFrequency > 0
    and timestep is a multiple of it (VTKOutput.cpp:129-132, 188-250); nothing is written when no function is
    registered.  The storage is three-dimensional, so output starts in 3D (macro-cells); set2D() writes the
    macro-faces instead.  A function with nrhs > 1 is written as nrhs arrays named "<name>_<k>".
    """

    def __init__(self, dir, filename, storage, writeFrequency=1):
        self.dir, self.filename, self.storage = str(dir), str(filename), storage
        self.writeFrequency = int(writeFrequency)
        self.write2D = False
        self.vtkDataFormat = DataFormat.ASCII
        self.p1Functions = []

    def setVTKDataFormat(self, fmt):
        if fmt not in (DataFormat.ASCII, DataFormat.BINARY):
            raise ValueError("VTK format not supported.")
        self.vtkDataFormat = fmt

    def set2D(self):
        self.write2D = True

    def set3D(self):
        self.write2D = False

    def add(self, function):
        if not hasattr(function, "download_micro_vertices"):
            raise TypeError("VTKOutput: Failed to add function object!")
        if function.storage is not self.storage:
            raise ValueError("VTKOutput: function lives on another storage")
        self.p1Functions.append(function)

    def fileNameExtension(self, level, timestep):
        return "_VertexDoF_level%d_ts%d" % (level, timestep)

    def path(self, level, timestep=0):
        return os.path.join(self.dir, self.filename + self.fileNameExtension(level, timestep) + ".vtu")

    def write(self, level, timestep=0):
        """Write the registered functions on `level`; returns the file path, or None when nothing was due."""
        if self.writeFrequency <= 0 or timestep % self.writeFrequency != 0 or not self.p1Functions:
            return None
        for f in self.p1Functions:
            if not f.minLevel <= level <= f.maxLevel:
                raise ValueError("function %r has no level %d" % (f.name, level))
        text = self.to_string(level)
        path = self.path(level, timestep)
        with open(path, "w") as fh:
            fh.write(text)
        return path

    def to_string(self, level):
        """The complete .vtu document for `level` (header, one piece, footer)."""
        parts = ['<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1">\n<UnstructuredGrid>\n']
        self._piece(parts, level)
        parts.append("</UnstructuredGrid>\n</VTKFile>\n")
        return "".join(parts)

    def _piece(self, parts, level):
        """VTKP1Writer::write (VTKP1Writer.cpp:34-100)."""
        cls = FACE if self.write2D else CELL
        n = self.storage.counts()[cls]
        w = (1 << level) + 1
        per = w * (w + 1) // 2 if self.write2D else _tet_size(w)
        if self.write2D:
            local, pts = micro_face_connectivity(level), micro_face_points
            ctype, cdtype = _VTK_TRIANGLE, np.uint32
        else:
            local, pts = micro_cell_connectivity(level), micro_vertex_points
            ctype, cdtype = _VTK_TETRA, np.int32
        parts.append('<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (n * per, n * len(local)))

        points = np.concatenate([pts(self.storage.primitive_coordinates(cls, p), level) for p in range(n)]) \
            if n else np.zeros((0, 3))
        parts.append("<Points>\n")
        _data_array(parts, points, np.float64, "", 3, self.vtkDataFormat)
        parts.append("</Points>\n")

        nv = local.shape[1]
        conn = (np.arange(n, dtype=np.int64)[:, None, None] * per + local[None]).reshape(-1)
        if conn.size and conn.max() > np.iinfo(cdtype).max:
            raise ValueError("too many points for the VTK connectivity type")
        parts.append("<Cells>\n")
        _data_array(parts, conn, cdtype, "connectivity", 0, self.vtkDataFormat)
        _data_array(parts, nv * np.arange(1, n * len(local) + 1, dtype=np.int64), np.uint32, "offsets", 0,
                    self.vtkDataFormat)
        _data_array(parts, np.full(n * len(local), ctype), np.uint16, "types", 0, self.vtkDataFormat)
        parts.append("</Cells>\n")

        parts.append("<PointData>\n")
        for f in self.p1Functions:
            vals = f.download_micro_vertices(level, cls)
            if f.nrhs == 1:
                _data_array(parts, vals, np.float64, f.name, 1, self.vtkDataFormat)
            else:
                for k in range(f.nrhs):
                    _data_array(parts, vals[..., k], np.float64, "%s_%d" % (f.name, k), 1, self.vtkDataFormat)
        parts.append("</PointData>\n")
        parts.append("</Piece>\n")


class FixedSizeSQLDB:
    """FixedSizeSQLDB(dbFile, allowOverwritingValues=False) (src/hyteg/dataexport/SQL.hpp:37-134, SQL.cpp:24-59).

    Constant entries go into every row; variable entries must be set before every writeRowOnRoot(), which
    refuses a row with a variable key missing, so no cell stays empty.  Columns are fixed after the first row.
    Rows go to table "runs" (runId INTEGER PRIMARY KEY, timestamp, one column per key: INTEGER for integers and
    booleans, DOUBLE for reals, TEXT for strings), the layout of walberla's SQLiteDB::storeRun, which the
    reference calls; a key that is new to an existing file adds a column.
    """

    def __init__(self, dbFile, allowOverwritingValues=False):
        self.dbFile = str(dbFile)
        self.allowOverwritingValues = bool(allowOverwritingValues)
        self.columnsFixed = False
        self.constant = {}
        self.variable = {}
        self.variableKeys = []

    @staticmethod
    def _check_value(value):
        if isinstance(value, (bool, np.bool_, int, np.integer, float, np.floating, str)):
            return value.item() if isinstance(value, np.generic) else value
        raise TypeError("SQL entries are strings, reals, integers or booleans")

    def setConstantEntry(self, key, value):
        if self.columnsFixed:
            raise RuntimeError('Cannot insert "%s" since columns are fixed after first write.' % key)
        if not self.allowOverwritingValues and key in self.constant:
            raise RuntimeError('Attempt to overwrite "%s" dismissed.' % key)
        self.constant[key] = self._check_value(value)

    def setVariableEntry(self, key, value):
        if self.columnsFixed:
            if key not in self.variableKeys:
                raise RuntimeError('New variable key "%s" cannot be inserted since columns are fixed.' % key)
        elif key not in self.variableKeys:
            self.variableKeys.append(key)
        if not self.allowOverwritingValues and key in self.variable:
            raise RuntimeError('Attempt to overwrite "%s" dismissed.' % key)
        self.variable[key] = self._check_value(value)

    @staticmethod
    def _sql_type(value):
        if isinstance(value, (bool, int)):
            return "INTEGER"
        if isinstance(value, float):
            return "DOUBLE"
        return "TEXT"

    def writeRowOnRoot(self):
        """Store the current row; returns its runId."""
        for key in self.variableKeys:
            if key not in self.variable:
                raise RuntimeError('Variable key "%s" not set for this row.' % key)
        row = dict(self.constant)
        row.update(self.variable)
        con = sqlite3.connect(self.dbFile)
        try:
            with con:
                con.execute("CREATE TABLE IF NOT EXISTS runs (runId INTEGER PRIMARY KEY, timestamp)")
                have = {r[1] for r in con.execute("PRAGMA table_info(runs)")}
                for key, value in row.items():
                    if key not in have:
                        con.execute('ALTER TABLE runs ADD COLUMN "%s" %s' % (key.replace('"', '""'),
                                                                             self._sql_type(value)))
                keys = list(row)
                cols = ", ".join(['timestamp'] + ['"%s"' % k.replace('"', '""') for k in keys])
                marks = ", ".join(["CURRENT_TIMESTAMP"] + ["?"] * len(keys))
                cur = con.execute("INSERT INTO runs (%s) VALUES (%s)" % (cols, marks),
                                  [int(v) if isinstance(v, bool) else v for v in row.values()])
                run = cur.lastrowid
        finally:
            con.close()
        self.columnsFixed = True
        self.variable = {}
        return run
