"""CPU tests of the product's host side: the C-ABI library loads and exports every declared entry point,
topology / owned numbering / stencil weights agree with the oracle (host-only context, no GPU needed),
and errors come back as status codes."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from helpers import mesh_pair

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "hymi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hymi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import hyteg_amd._lib as L
    lib = L.load()
    names = _declared_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    assert lib.hymi_version() == 1


def test_every_declared_symbol_is_bound_in_python():
    import hyteg_amd._lib as L
    bound = set(L.SIGNATURES) | {"hymi_last_error_string"}
    assert set(_declared_symbols()) <= bound


@pytest.fixture(scope="module")
def H():
    import hyteg_amd
    return hyteg_amd


@pytest.fixture(scope="module")
def hctx(H):
    return H.Context(-1)


MESHES = [("cube", 1), ("cube", 2), ("cube", 3), ("gmsh", "regular_octahedron_8el.msh"), ("gmsh", "cube_48el.msh"),
          ("gmsh", "pyramid_2el.msh"), ("gmsh", "tet_1el.msh")]


@pytest.mark.parametrize("kind", MESHES)
def test_topology_and_owned_numbering_match_oracle(H, hctx, oracle_lib, kind):
    om, S = mesh_pair(kind, hctx)
    assert S.counts() == om.counts()
    nV, nE, nF, nC = om.counts()
    for cls, n in ((0, nV), (1, nE), (2, nF), (3, nC)):
        for i in range(n):
            assert S.primitive_type(cls, i) == om.prim_type(cls, i)
    for L in range(0, 5):
        assert S.owned_count(L) == om.owned_count(L)
        assert np.array_equal(S.owned_types(L), om.owned_types(L))
        assert np.abs(S.owned_coords(L) - om.owned_coords(L)).max() <= 1e-15


@pytest.mark.parametrize("kind", MESHES)
@pytest.mark.parametrize("form", [0, 1])
def test_stencil_weights_match_oracle(H, hctx, oracle_lib, kind, form):
    O = oracle_lib
    om, S = mesh_pair(kind, hctx)
    A = (H.P1ConstantLaplaceOperator if form == 0 else H.P1ConstantMassOperator)(S, 0, 4)
    B = O.Operator(om, 0, 4, form)
    nV, nE, nF, nC = om.counts()
    for L in range(1, 5):
        def close(a, b):
            return a.shape == b.shape and np.abs(a - b).max() <= 1e-14 * np.abs(b).max()

        for v in range(nV):
            assert close(A.stencil(L, H.VERTEX, v), B.vertex_stencil(L, v))
        for e in range(nE):
            b = B.edge_stencil(L, e)  # reference layout [f0 W, f0 E, W, C, E, f1 W, f1 E, ..., cells]
            assert close(A.stencil(L, H.EDGE, e), np.concatenate([b[2:5], b[0:2], b[5:]]))
        if L < 2:
            continue
        for c in range(nC):
            assert close(A.stencil(L, H.CELL, c), B.cell_stencil(L, c))
        for f in range(nF):
            a = A.stencil(L, H.FACE, f)
            sides = [B.face_stencil(L, f, n) for n in range(om.face_cells(f).tolist().count(-1) == 0 and 2 or 1)]
            ref = sum(sides)
            scale = abs(ref[0])
            assert abs(a[0] - ref[0]) <= 1e-14 * scale
            assert abs(a.sum() - ref.sum()) <= 1e-14 * scale


def test_errors_are_status_codes(H, hctx):
    import hyteg_amd._lib as L
    lib = L.load()
    out = ctypes.c_void_p()
    lo = (ctypes.c_double * 3)(0, 0, 0)
    hi = (ctypes.c_double * 3)(1, 1, 1)
    assert lib.hymi_mesh_create_cuboid(hctx.h, lo, hi, 0, 1, 1, ctypes.byref(out)) == L.ERR_INTERNAL
    assert b"sub-cube" in lib.hymi_last_error_string()
    assert lib.hymi_mesh_counts(None, None) == L.ERR_INVALID
    assert lib.hymi_mesh_from_gmsh(hctx.h, b"/nonexistent.msh", ctypes.byref(out)) != L.OK
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 1, 1, 1), hctx)
    f = ctypes.c_void_p()
    # functions need a device: a host-only context refuses with a status code
    assert lib.hymi_function_create(S.h, 0, 2, 0, 1, ctypes.byref(f)) == L.ERR_INVALID
    with pytest.raises(H.HymiError):
        H.P1ConstantLaplaceOperator(S, 3, 1)


def test_schedule_arguments(H, hctx):
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 1, 1, 1), hctx)
    for L in (2, 5, 7, 8):
        d = S.get_schedule(L)
        assert 1 <= d["rows"] <= 4096 and d["waves"] in (1, 2, 4) and d["block"] in (1, 2, 4, 8)
        assert d["kind"] in (0, 1, 2, 3)
    S.set_schedule(5, rows=16, waves=4, block=8, kind=1)
    assert S.get_schedule(5) == {"rows": 16, "waves": 4, "block": 8, "kind": 1}
    for bad in (dict(waves=3), dict(block=3), dict(rows=5000), dict(rows=-1), dict(kind=4),
                dict(kind=2, rows=200, waves=4, block=8), dict(kind=3, block=8)):
        with pytest.raises(H.HymiError):
            S.set_schedule(5, **bad)
    assert S.get_schedule(5) == {"rows": 16, "waves": 4, "block": 8, "kind": 1}   # a refused call changes nothing
    S.set_schedule(5)
    assert S.get_schedule(5) == S.get_schedule(5) and S.get_schedule(5)["rows"] >= 1


def test_bench_refuses_multi_gpu():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert res.returncode != 0
    assert "one GPU" in res.stderr


def test_serving_counts_and_stream(H, hctx):
    from hyteg_amd.serving import cuboid_counts, hierarchy_bytes, poisson_stream
    for n in (1, 2, 3):
        S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n), hctx)
        assert cuboid_counts(n) == S.counts()
    # budget estimate grows ~8x per level and covers the finest level's three FP64 and three FP32 copies
    for n in (1, 2, 3):
        fixed = 8 << 20                                            # tables and scalars
        assert 6.5 < (hierarchy_bytes(n, 7) - fixed) / (hierarchy_bytes(n, 6) - fixed) < 8.5
        assert hierarchy_bytes(n, 7) > 36 * ((n * 2 ** 7 + 1) ** 3)
    a, b = poisson_stream(2000, rate=200.0, seed=3), poisson_stream(2000, rate=200.0, seed=3)
    assert [(r.arrival, r.n, r.level) for r in a] == [(r.arrival, r.n, r.level) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0          # exponential inter-arrival, mean 1/rate
    assert {r.n for r in a} == {1, 2, 3} and {r.level for r in a} == {4, 5, 6, 7}
    assert all(r.modes.shape == (4, 4) and 1 <= r.modes[:, 1:].min() and r.modes[:, 1:].max() <= 8 for r in a)


def test_cell_sor_setting(H, hctx):
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 1, 1, 1), hctx)
    assert S.get_cell_sor(6) == 0
    S.set_cell_sor(6, 24)
    S.set_cell_sor(3, 24, fp32=True)
    S.set_cell_sor(8, 24, fp64=False, fp32=True)
    # level 7 keeps the library default: the fused phase for both precisions
    assert [S.get_cell_sor_setting(L) for L in (6, 3, 8, 7)] == [(24, 1), (24, 3), (24, 2), (24, 3)]
    S.set_cell_sor(6, 0)
    assert S.get_cell_sor(6) == 0
    for level, tile, kw in ((6, 8, {}), (6, 16, {}), (6, 32, {}), (2, 24, {}), (6, 24, {"fp64": False})):
        with pytest.raises(H.HymiError):
            S.set_cell_sor(level, tile, **kw)
    S2 = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 1, 1, 1), hctx)
    assert [S2.get_cell_sor_setting(L) for L in (5, 6, 7, 8)] == [(0, 0), (0, 0), (24, 3), (24, 3)]
    assert [S2.get_schedule(L)["kind"] for L in (3, 4, 5, 6, 7, 8)] == [1, 1, 3, 3, 3, 3]
    assert [(S2.get_schedule(L)["rows"], S2.get_schedule(L)["waves"]) for L in (5, 6, 7)] == [(4, 2), (8, 4), (16, 4)]
    S.use_column_schedules(8)
    assert [S.get_cell_sor_setting(L) for L in (5, 6, 7, 8)] == [(0, 0), (0, 0), (24, 3), (24, 3)]
    S.use_column_schedules(8, fusedSorFp32From=8)
    assert [S.get_cell_sor_setting(L) for L in (6, 7, 8)] == [(0, 0), (24, 1), (24, 3)]


def _tri_row_f32(W, f):
    """The row lookup the device kernels use (kernels.hip triRow32): a float32 estimate from the smaller root of
    the row-start quadratic R(y) = y (2W + 1 - y) / 2, then one correction step either way."""
    f = np.asarray(f, dtype=np.int64)
    q = np.float32(2 * W + 1)
    y = (np.float32(0.5) * (q - np.sqrt(q * q - np.float32(8.0) * f.astype(np.float32)))).astype(np.int64)
    y += ((y + 1) * (2 * W - y) // 2 <= f).astype(np.int64)
    y -= (y * (2 * W + 1 - y) // 2 > f).astype(np.int64)
    return y


def test_row_lookup_closed_form_at_every_row_boundary():
    """The flat transfer, face-ghost and band-copy kernels find the row of a flat triangle index in closed form.
    Against a search over the row starts it must hold at both sides of every row boundary, for triangle widths
    up to those of level 11 (a perturbed estimate is caught too: the check is not vacuous)."""
    widths = list(range(2, 300)) + list(range(300, 2060, 7)) + [1025, 2049, 2050]
    for W in widths:
        n = W * (W + 1) // 2
        ys = np.arange(W)
        starts = ys * (2 * W + 1 - ys) // 2
        f = np.concatenate([starts, starts + 1, starts - 1, starts + (W - ys) - 1])
        f = f[(f >= 0) & (f < n)]
        exact = np.searchsorted(starts, f, side="right") - 1
        assert np.array_equal(_tri_row_f32(W, f), exact), W
    # the check can fail: with the row length off by one (rows of W + 1 - y points) the lookup misses boundaries
    W = 257
    ys = np.arange(W)
    starts = ys * (2 * W + 1 - ys) // 2
    f = np.concatenate([starts, starts - 1])[1:]
    assert not np.array_equal(_tri_row_f32(W + 1, f), np.searchsorted(starts, f, side="right") - 1)
