"""GPU parity of the drop-in entry points that mirror the reference's interface one to one: prolongate with
Replace (P1toP1LinearProlongation::prolongate), smooth_gs (GSSmoothable::smooth_gs), the multigrid solver's
smoothIncrementOnCoarserGrids, and the per-primitive upload in the reference's FunctionMemory layout.

Tolerances as in test_gpu_parity.py: transfers <= 1e-13, SOR sweeps <= 1e-12, V-cycles <= 1e-11 relative; the
per-primitive transfers and the Gauss-Seidel alias are bit for bit.
"""
import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say, sine_rhs

pytestmark = pytest.mark.gpu

MESHES = [("cube", 2), ("gmsh", "regular_octahedron_8el.msh")]


@pytest.mark.parametrize("kind", MESHES)
@pytest.mark.parametrize("level", [1, 2, 3])
def test_prolongate_replace(kind, level, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(40 + level)
    for flag, oflag in ((H.INNER, O.INNER), (H.ALL, O.ALL), (H.INNER | H.NEUMANN, O.INNER | O.NEUMANN)):
        say(f"[prolongate replace] mesh={kind} level={level} flag={flag}")
        f = H.P1Function("f", S, 0, level + 1)
        of = O.Function(om, 0, level + 1)
        for L in (level, level + 1):
            v = rng.standard_normal(om.owned_count(L))
            f.upload(L, v)
            of.set(L, v)
        H.P1toP1LinearProlongation().prolongate(f, level, flag)
        O.prolongate(of, level, flag=oflag, add=False)
        assert rel_err(f.download(level + 1), of.get(level + 1)) <= 1e-13, flag
        # the source level is untouched
        assert rel_err(f.download(level), of.get(level)) == 0.0


@pytest.mark.parametrize("kind", MESHES)
@pytest.mark.parametrize("level", [2, 3, 5])
def test_smooth_gs_is_sor_with_relax_one(kind, level, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[smooth_gs] mesh={kind} level={level}")
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(70 + level)
    x0, bb = rng.standard_normal((2, om.owned_count(level)))
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    OA = O.Operator(om, 0, level)
    u, v, f = (H.P1Function(n, S, 0, level) for n in "uvf")
    f.upload(level, bb)
    ou, of = O.Function(om, 0, level), O.Function(om, 0, level)
    ou.set(level, x0)
    of.set(level, bb)
    u.upload(level, x0)
    v.upload(level, x0)
    for sweep in range(2):
        A.smooth_gs(u, f, level, H.INNER)
        A.smooth_sor(v, f, 1.0, level, H.INNER)
        OA.smooth_sor(ou, of, level, omega=1.0, flag=O.INNER, order=1)
        assert np.array_equal(u.download(level), v.download(level)), sweep
        assert rel_err(u.download(level), ou.get(level)) <= 1e-12, sweep
    A.smooth_gs_backwards(u, f, level, H.INNER)
    A.smooth_sor_backwards(v, f, 1.0, level, H.INNER)
    OA.smooth_sor_backward(ou, of, level, omega=1.0, flag=O.INNER, order=1)
    assert np.array_equal(u.download(level), v.download(level))
    assert rel_err(u.download(level), ou.get(level)) <= 1e-12


@pytest.mark.parametrize("kind,L", [(("cube", 2), 4), (("gmsh", "regular_octahedron_8el.msh"), 3)])
def test_smooth_increment_on_coarser_grids(kind, L, gpu_ctx, oracle_lib):
    """GeometricMultigridSolver(..., pre, post, smoothIncrementOnCoarserGrids): V(1,1) + 1 per coarser level,
    eager and from a captured graph, against the oracle; then setSmoothingSteps back to V(2,2) + 0."""
    import hyteg_amd as H
    O = oracle_lib
    say(f"[smooth increment] mesh={kind} level={L}")
    om, S = mesh_pair(kind, gpu_ctx)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    OA = O.Operator(om, 0, L)
    b = H.P1Function("b", S, 0, L)
    ob = O.Function(om, 0, L)
    rhs = sine_rhs(om.owned_coords(L))
    b.upload(L, rhs)
    ob.set(L, rhs)
    x, x2 = H.P1Function("x", S, 0, L), H.P1Function("x2", S, 0, L)
    ox = O.Function(om, 0, L)
    gmg = H.GeometricMultigridSolver(S, H.GaussSeidelSmoother(), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, 1, 1, smoothIncrement=1)
    og = O.GMG(om, 0, L, 1, 1, 1.0, 1)
    og.set_smoothing(1, 1, 1)
    for cyc in range(3):
        gmg.solve(A, x, b, L)
        gmg.solve_graph(A, x2, b, L, 1)
        og.solve(OA, ox, ob, L)
        assert rel_err(x.download(L), ox.get(L)) <= 1e-11, cyc
        assert np.array_equal(x.download(L), x2.download(L)), cyc
    gmg.setSmoothingSteps(2, 2, 0)
    og.set_smoothing(2, 2, 0)
    gmg.solve_graph(A, x2, b, L, 1)
    og.solve(OA, ox, ob, L)
    assert rel_err(x2.download(L), ox.get(L)) <= 1e-11


@pytest.mark.parametrize("kind", MESHES + [("gmsh", "cube_48el.msh")])
def test_upload_primitive_round_trip(kind, gpu_ctx, oracle_lib):
    """hymi_function_upload_primitive takes each primitive's owned points from an array in the reference's
    FunctionMemory layout (ghost slots and closure copies ignored): uploading every primitive of the oracle's
    communicated memory rebuilds the function, and downloading gives the oracle's memory back bit for bit."""
    import hyteg_amd as H
    O = oracle_lib
    L = 3
    say(f"[upload primitive] mesh={kind} level={L}")
    om, S = mesh_pair(kind, gpu_ctx)
    vals = np.random.default_rng(9).standard_normal(om.owned_count(L))
    of = O.Function(om, 0, L)
    of.set(L, vals)
    for d in range(6):
        of.communicate(L, d)
    f = H.P1Function("f", S, 0, L)
    f.interpolate(7.0, L, H.ALL)          # every value is overwritten by the uploads
    counts = om.counts()
    for cls in (H.CELL, H.FACE, H.EDGE, H.VERTEX):
        for idx in range(counts[cls]):
            mem = of.prim(L, cls, idx)
            garbled = mem.copy()
            n_owned = {H.VERTEX: 1, H.EDGE: 2 ** L + 1, H.FACE: (2 ** L + 1) * (2 ** L + 2) // 2,
                       H.CELL: len(mem)}[cls]
            garbled[n_owned:] = np.nan     # ghost layers: ignored on upload
            f.upload_primitive(L, cls, idx, garbled)
    assert np.array_equal(f.download(L), vals)
    for cls in (H.VERTEX, H.EDGE, H.FACE, H.CELL):
        for idx in range(counts[cls]):
            assert np.array_equal(f.download_primitive(L, cls, idx), of.prim(L, cls, idx)), (cls, idx)


@pytest.mark.parametrize("kind", MESHES)
@pytest.mark.parametrize("level", [2, 3, 5])
def test_sor_other_flags(kind, level, gpu_ctx, oracle_lib):
    """SOR sweeps with flags that update the boundary (All, Dirichlet), two in a row, against the oracle.  Every
    face copies its edges' new values into its closure, also a face whose own points the flag leaves alone."""
    import hyteg_amd as H
    O = oracle_lib
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(90 + level)
    x0, bb = rng.standard_normal((2, om.owned_count(level)))
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    OA = O.Operator(om, 0, level)
    for flag, oflag in ((H.ALL, O.ALL), (H.DIRICHLET, O.DIRICHLET)):
        say(f"[sor flags] mesh={kind} level={level} flag={flag}")
        u, f = H.P1Function("u", S, 0, level), H.P1Function("f", S, 0, level)
        u.upload(level, x0)
        f.upload(level, bb)
        ou, of = O.Function(om, 0, level), O.Function(om, 0, level)
        ou.set(level, x0)
        of.set(level, bb)
        for sweep in range(2):
            A.smooth_sor(u, f, 0.8, level, flag)
            OA.smooth_sor(ou, of, level, omega=0.8, flag=oflag, order=1)
            assert rel_err(u.download(level), ou.get(level)) <= 1e-12, (flag, sweep)
