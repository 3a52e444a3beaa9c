"""GPU tests of the fused SOR cell phase (all four colours of the macro-cell interiors in one launch, tiles
marched along z): it must give the four colour passes' values bit for bit -- forward and backward sweeps,
levels 3 to 7, FP64 and (where selected for it) FP32 -- and the oracle's to 1e-12 relative; V-cycles with it
must equal V-cycles without it, also through captured graphs that were recorded before the switch, and the
ping-pong pairs of their smoothing sequences must leave every macro-cell closure as the colour passes do."""
import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say, sine_rhs

pytestmark = pytest.mark.gpu


def _sweeps(H, S, L, x, b, tile, backward, prec, sweeps=2):
    S.set_cell_sor(L, tile)
    assert S.get_cell_sor(L) == tile
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    u, f = H.P1Function("u", S, 0, L, precision=prec), H.P1Function("f", S, 0, L, precision=prec)
    u.upload(L, x)
    f.upload(L, b)
    for _ in range(sweeps):
        if backward:
            A.smooth_sor_backwards(u, f, 0.9, L, H.INNER)
        else:
            A.smooth_sor(u, f, 0.9, L, H.INNER)
    return u.download(L)


@pytest.mark.parametrize("kind,level", [(("cube", 1), 3), (("cube", 1), 6), (("gmsh", "regular_octahedron_8el.msh"), 6),
                                        (("cube", 1), 7)])
@pytest.mark.parametrize("tile", [24])
@pytest.mark.parametrize("backward", [False, True])
def test_fused_equals_colour_passes(kind, level, tile, backward, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    say(f"[sor fused] mesh={kind} level={level} tile={tile} backward={backward} FP64")
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(100 + level)
    x, b = rng.standard_normal((2, om.owned_count(level)))
    ref = _sweeps(H, S, level, x, b, 0, backward, H.F64)
    got = _sweeps(H, S, level, x, b, tile, backward, H.F64)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("level", [6, 7])
def test_fused_fp32(level, gpu_ctx, oracle_lib):
    """FP32 functions run the fused phase where it is selected for FP32 (same values as the colour passes, bit
    for bit), and keep the passes where it is selected for FP64 only (timing the tile kernel is then refused)."""
    import ctypes
    import hyteg_amd as H
    L = level
    say(f"[sor fused] cube n=1 level={L} FP32")
    om, S = mesh_pair(("cube", 1), gpu_ctx)
    rng = np.random.default_rng(8)
    x, b = rng.standard_normal((2, om.owned_count(L)))
    ref = _sweeps(H, S, L, x, b, 0, False, H.F32)
    S.set_cell_sor(L, 24, fp32=True)
    A = H.P1ConstantLaplaceOperator(S, L, L)
    u, f = H.P1Function("u", S, L, L, precision=H.F32), H.P1Function("f", S, L, L, precision=H.F32)
    u.upload(L, x)
    f.upload(L, b)
    for _ in range(2):
        A.smooth_sor(u, f, 0.9, L, H.INNER)
    assert np.array_equal(u.download(L), ref)
    got = _sweeps(H, S, L, x, b, 24, False, H.F32)          # FP64 only: FP32 keeps the passes
    assert np.array_equal(got, ref)
    ms = ctypes.c_double()
    with pytest.raises(H.HymiError):
        H._lib.call("hymi_time_kernel", A.h, None, u.h, f.h, None, L, 7, 1, ctypes.byref(ms))


def test_fused_matches_oracle(gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    L = 5
    say(f"[sor fused] octahedron level={L} tile=24 vs oracle")
    om, S = mesh_pair(("gmsh", "regular_octahedron_8el.msh"), gpu_ctx)
    rng = np.random.default_rng(12)
    x, b = rng.standard_normal((2, om.owned_count(L)))
    got = _sweeps(H, S, L, x, b, 24, False, H.F64, sweeps=3)
    OA = O.Operator(om, 0, L)
    ou, of = O.Function(om, 0, L), O.Function(om, 0, L)
    ou.set(L, x)
    of.set(L, b)
    for _ in range(3):
        OA.smooth_sor(ou, of, L, omega=0.9, flag=O.INNER, order=1)
    assert rel_err(got, ou.get(L)) <= 1e-12


def test_vcycle_graphs_with_fused_cell_phase(gpu_ctx, oracle_lib):
    """V(3,3) cycles replayed from captured graphs: the fused cell phase on levels 6 and 7 gives the colour
    passes' iterates bit for bit; switching the phase after a capture re-captures the graph."""
    import hyteg_amd as H
    L = 7
    say(f"[sor fused] cube n=1 V(3,3) graphs level={L}")
    om, S = mesh_pair(("cube", 1), gpu_ctx)
    X = om.owned_coords(L)
    rhs = sine_rhs(X)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    fi, f = H.P1Function("fi", S, 0, L), H.P1Function("f", S, 0, L)
    fi.upload(L, rhs)
    M.apply(fi, f, L, H.ALL)
    gmg = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, 3, 3)
    runs = {}
    u = H.P1Function("u", S, 0, L)
    for tiles in ((0, 0), (24, 24), (0, 0)):
        S.set_cell_sor(6, tiles[0])
        S.set_cell_sor(7, tiles[1])
        u.interpolate(0.0, L, H.ALL)
        gmg.solve_graph(A, u, f, L, count=3)       # same functions: the graph is captured again after a switch
        runs.setdefault(tiles, []).append(u.download(L))
    a, b2 = runs[(0, 0)]
    assert np.array_equal(a, b2)
    assert np.array_equal(runs[(24, 24)][0], a)


def test_chained_sweeps_symmetric_and_w_cycles(gpu_ctx, oracle_lib):
    """Smoothing sequences chain their fused sweeps (the halo of sweep k comes from sweep k-1's band array, no
    copy): symmetric SOR smoothing (forward and backward sweeps alternating) in W-cycles gives the colour
    passes' iterates bit for bit, and so do 5 pre- and post-sweeps of a V-cycle."""
    import hyteg_amd as H
    L = 7
    say(f"[sor fused] cube n=1 level={L} chained sweeps: symmetric W(2,2) and V(5,5)")
    om, S = mesh_pair(("cube", 1), gpu_ctx)
    rhs = sine_rhs(om.owned_coords(L))
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    f = H.P1Function("f", S, 0, L)
    f.upload(L, rhs)
    for smoother, cyc, nu in ((H.SymmetricSORSmoother(1.0), H.CycleType.WCYCLE, 2), (H.SORSmoother(1.0), H.CycleType.VCYCLE, 5)):
        gmg = H.GeometricMultigridSolver(S, smoother, H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                         H.P1toP1LinearProlongation(), 0, L, nu, nu, cycleType=cyc)
        out = []
        for tile in (0, 24):
            S.set_cell_sor(L, tile)
            S.set_cell_sor(6, tile)
            u = H.P1Function("u", S, 0, L)
            gmg.solve_graph(A, u, f, L, count=2)
            out.append(u.download(L))
        assert np.array_equal(out[0], out[1]), (cyc, nu)


@pytest.mark.parametrize("prec_name", ["F64", "F32"])
def test_pingpong_pairs_keep_closures(prec_name, gpu_ctx, oracle_lib):
    """V(2,2) and V(3,3) cycles on the 48-tet cube at level 6 run their smoothing sequences as ping-pong pairs when
    the fused cell phase is selected (closure points read from the faces, one face-to-cell copy per pair).  The
    owned values and every macro-cell array with its closure must equal those of the colour passes bit for bit."""
    import hyteg_amd as H
    prec = getattr(H, prec_name)
    L = 6
    om, S = mesh_pair(("cube", 2), gpu_ctx)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    for nu in (2, 3):
        say(f"[sor fused] ping-pong pairs cube n=2 level={L} {prec_name} V({nu},{nu})")
        rng = np.random.default_rng(60 + nu)
        x0, bv = rng.standard_normal((2, om.owned_count(L)))
        out = []
        for tile in (0, 24):
            S.set_cell_sor(L, tile, fp64=True, fp32=True)
            x, b = H.P1Function("x", S, 0, L, precision=prec), H.P1Function("b", S, 0, L, precision=prec)
            x.upload(L, x0)
            b.upload(L, bv)
            g = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                           H.P1toP1LinearProlongation(), 0, L, nu, nu)
            g.solve(A, x, b, L)
            out.append((x.download(L), [x.download_primitive(L, H.CELL, c) for c in range(48)]))
        assert np.array_equal(out[0][0], out[1][0])
        for c in range(48):
            assert np.array_equal(out[0][1][c], out[1][1][c]), c
